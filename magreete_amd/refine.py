"""Mixed-precision solve: complex64 LU (FP32 or BF16x3 MFMA products) plus
FP64 GMRES iterative refinement, with fallbacks.

GMRES is right-preconditioned by the low-precision factors P = L U and
flexible (the preconditioned directions are kept): x = Z y with
z_j = P^{-1} v_j, the Krylov basis built from A z_j in FP64 (A = M,
complex128, kept resident), starting from x = 0 so that the first direction
is P^{-1} b at no FP64 cost.  All T right-hand sides run together: the
matrix products are one (n x n) @ (n x T) FP64 GEMM and one complex64
multi-RHS solve (GEMM-only, through stored diagonal-block inverses) per
iteration; the per-column Gram-Schmidt runs as HIP multi-vector dot/axpy
passes and the Givens updates are small (T x j) torch operations.
Convergence is declared on the FP64 residual ||b - A x|| / ||b|| <= tol per
column, formed at the end of every cycle from the Arnoldi relation (equal to
the product b - A x up to the rounding of the FP64 products already taken).

The first factorisation is unpivoted (no pivot search, no grid barrier, no
row swap); the refinement restores the accuracy.  If it breaks down (an exact
zero pivot) or the refinement stalls, the complex64 LU is redone with partial
pivoting, and if that stalls too the solve falls back to the complex128 LU.
The returned stats say which tier answered.
"""
import torch

from . import _lib as L
from . import engine as E


def _gmres_ir(M, F, B, X, bnorm, tol, max_iters, restart, history=None):
    """Solve M X = B by flexible GMRES from X = 0, right-preconditioned by F.
    The first Krylov direction is P^{-1} b, so the first cycle starts with no
    FP64 product.  Every cycle ends with the residual b - M x formed from the
    Arnoldi relation M Z_k = V_{k+1} H_k, i.e. r = V_{k+1} (beta e1 - H_k y)
    with the unrotated Hessenberg matrix: it equals the FP64 product b - M x
    up to the rounding of the FP64 products M z_j already taken, at the cost
    of one pass over the basis instead of another product.
    history: optional list that receives the largest per-column relative
    residual estimate |g_{j+1}| / ||b|| after every FP64 product.
    Returns (converged, FP64 products, cycles, residual)."""
    dev = M.device
    n, T = B.shape
    lowp = torch.empty((n, T), dtype=torch.complex64, device=dev)
    fwork = torch.empty(F.work_elems(T), dtype=F.A.dtype, device=dev)
    m_max = max(1, min(restart, max_iters))
    work = torch.empty(L.load().mg_mdot_work_elems(m_max + 1, T), dtype=torch.complex128, device=dev)

    def precond(v, out):
        E.convert(v, lowp)
        F.solve(lowp, work=fwork)
        E.convert(lowp, out)
        return out

    bnorm64 = bnorm.to(torch.float64).contiguous()
    est_buf = torch.empty(1, dtype=torch.float64, device=dev)
    X.zero_()
    R = B.clone()
    beta = torch.linalg.norm(R, dim=0)
    rel = (beta / bnorm).max().item()
    iters = cycles = 0
    prev = None
    while True:
        if rel <= tol:
            return True, iters, cycles, rel
        if not (rel == rel) or iters >= max_iters or (prev is not None and rel > 0.5 * prev):
            return False, iters, cycles, rel
        prev = rel
        cycles += 1
        m = max(1, min(restart, max_iters - iters))
        V = torch.empty((m + 1, n, T), dtype=torch.complex128, device=dev)
        Z = torch.empty((m, n, T), dtype=torch.complex128, device=dev)
        H = torch.zeros((m + 1, m, T), dtype=torch.complex128, device=dev)
        H0 = torch.zeros((m + 1, m, T), dtype=torch.complex128, device=dev)   # unrotated, for the residual
        cs = torch.zeros((m, T), dtype=torch.complex128, device=dev)
        sn = torch.zeros((m, T), dtype=torch.complex128, device=dev)
        g = torch.zeros((m + 1, T), dtype=torch.complex128, device=dev)
        safe = torch.where(beta == 0, torch.ones_like(beta), beta)
        torch.div(R, safe, out=V[0])
        g[0] = beta.to(torch.complex128)
        k = 0
        for j in range(m):
            precond(V[j], Z[j])
            w = V[j + 1]
            E.cgemm(M, Z[j], w, alpha=1.0, beta=0.0, split=True)    # w = A z_j (FP64)
            iters += 1
            for _ in range(2):                                # classical Gram-Schmidt, twice
                h = E.mdot(V[: j + 1], w, work=work)
                E.maxpy(V[: j + 1], h, w, alpha=-1.0)
                H[: j + 1, j] += h
            hn = torch.linalg.norm(w, dim=0)
            w /= torch.where(hn == 0, torch.ones_like(hn), hn)
            # h_{j+1,j} = hn, the unrotated column into H0, the earlier Givens
            # rotations and a new one, g rotated (one HIP kernel)
            E.givens(H, H0, cs, sn, g, hn, bnorm64, j, est_buf)
            k = j + 1
            # |g_{j+1}| is the norm of the Arnoldi residual formed below
            est = est_buf.item()
            if history is not None:
                history.append(est)
            if est <= tol:
                break
        Hk = H[:k, :k].permute(2, 0, 1).contiguous()
        gk = g[:k].T.contiguous().unsqueeze(-1)
        y = torch.linalg.solve_triangular(Hk, gk, upper=True).squeeze(-1).T.contiguous()   # (k, T)
        E.maxpy(Z[:k], y, X, alpha=1.0)                        # x += Z y
        c = -(H0[: k + 1, :k] * y.unsqueeze(0)).sum(1)           # beta e1 - H_k y
        c[0] += beta.to(torch.complex128)
        R.zero_()
        E.maxpy(V[: k + 1], c.contiguous(), R, alpha=1.0)       # r = V_{k+1} c
        del V, Z, H, H0
        beta = torch.linalg.norm(R, dim=0)
        rel = (beta / bnorm).max().item()


class MixedSolver:
    """A resident mixed-precision factorisation of M with FP64 GMRES-IR solves.

    M (complex128) stays intact and must outlive the solver; M32 holds M in
    complex64 on entry and is overwritten by the factors.  The tier starts
    at the unpivoted block LDU ('nopiv', or 'pivoted' with pivot=True); a
    solve whose refinement breaks down or stalls moves the solver to the
    next tier (complex64 LU with partial pivoting, then complex128 LU) for
    this and every later solve."""

    TIERS = ("nopiv", "pivoted", "fp64")

    def __init__(self, M, M32, prec="c64", tol=1e-12, max_iters=30, restart=12, pivot=False):
        self.M, self.M32, self.prec = M, M32, prec
        self.tol, self.max_iters, self.restart = tol, max_iters, restart
        self.tier = "pivoted" if pivot else "nopiv"
        self.fell = False
        self.F = None
        self._factor(fresh=True)

    def _factor(self, fresh=False):
        if self.tier == "fp64":
            self.F = E.Factor(self.M.clone(), "c128")
            return
        if not fresh:
            E.convert(self.M, self.M32)                    # the previous tier overwrote M32
        self.F = E.Factor(self.M32, self.prec, pivot=(self.tier == "pivoted"))
        if self.tier == "nopiv" and int(self.F.info.item()) != 0:
            self._escalate()                               # exact zero pivot: pivot this time

    def _escalate(self):
        self.tier = self.TIERS[self.TIERS.index(self.tier) + 1]
        self.fell = True
        self._factor()

    @property
    def info(self):
        return self.F.info

    def solve(self, B, X):
        """X <- M^{-1} B to a relative FP64 residual <= tol per column.
        Returns stats: tier, fallback, iters (FP64 products), cycles,
        residual."""
        bnorm = torch.linalg.norm(B, dim=0)
        bnorm = torch.where(bnorm == 0, torch.ones_like(bnorm), bnorm)
        total = 0
        hist = []
        while self.tier != "fp64":
            ok, iters, cycles, rel = _gmres_ir(self.M, self.F, B, X, bnorm, self.tol, self.max_iters, self.restart,
                                               hist)
            total += iters
            if ok:
                return {"tier": self.tier, "fallback": self.fell, "iters": total, "cycles": cycles,
                        "residual": rel, "history": hist}
            self._escalate()
        # refinement stalled in low precision: complex128 LU
        X.copy_(B)
        self.F.solve(X)
        R = B.clone()
        E.cgemm(self.M, X, R, alpha=-1.0, beta=1.0, split=True)
        rel = (torch.linalg.norm(R, dim=0) / bnorm).max().item()
        return {"tier": "fp64", "fallback": True, "iters": total + 1, "cycles": 0, "residual": rel}


def solve_mixed(M, M32, B, X, prec="c64", tol=1e-12, max_iters=30, restart=12, pivot=False):
    """Solve M X = B to a relative FP64 residual <= tol (one MixedSolver,
    one solve).  Returns stats: tier ('nopiv', 'pivoted' or 'fp64'), iters
    (FP64 products), cycles (GMRES restarts + 1), residual, fallback, info."""
    ms = MixedSolver(M, M32, prec, tol=tol, max_iters=max_iters, restart=restart, pivot=pivot)
    st = ms.solve(B, X)
    st["info"] = ms.info
    return st
