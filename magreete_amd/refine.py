"""Mixed-precision solve: complex64 LU (FP32 or BF16x3 MFMA products) plus
FP64 GMRES iterative refinement, with fallbacks.

GMRES is right-preconditioned by the low-precision factors P = L U:
x = x0 + Z y with z_j = P^{-1} v_j, the Krylov basis built from A z_j in
FP64 (A = M, complex128, kept resident).  All T right-hand sides run
together: the matrix products are one (n x n) @ (n x T) FP64 GEMM and one
complex64 multi-RHS solve (GEMM-only, through stored diagonal-block
inverses) per iteration; the per-column Gram-Schmidt and Givens updates are
small (T x j) operations.  Convergence is declared on the true FP64 residual
||b - A x|| / ||b|| <= tol per column.

The first factorisation is unpivoted (no pivot search, no grid barrier, no
row swap); the refinement restores the accuracy.  If it breaks down (an exact
zero pivot) or the refinement stalls, the complex64 LU is redone with partial
pivoting, and if that stalls too the solve falls back to the complex128 LU.
The returned stats say which tier answered.
"""
import torch

from . import engine as E


def _gmres_ir(M, F, B, X, bnorm, tol, max_iters, restart):
    """Refine X (initially unset) for M X = B with preconditioner F.
    Returns (converged, iters, restarts, residual)."""
    dev = M.device
    n, T = B.shape
    lowp = torch.empty((n, T), dtype=torch.complex64, device=dev)

    def precond(v, out):
        E.convert(v, lowp)
        F.solve(lowp)
        E.convert(lowp, out)
        return out

    precond(B, X)                                          # x0 = P^{-1} b
    R = torch.empty_like(B)
    iters = restarts = 0
    prev = None
    while True:
        R.copy_(B)
        E.cgemm(M, X, R, alpha=-1.0, beta=1.0)              # r = b - A x  (FP64)
        iters += 1
        beta = torch.linalg.norm(R, dim=0)
        rel = (beta / bnorm).max().item()
        if rel <= tol:
            return True, iters, restarts, rel
        if not (rel == rel) or iters >= max_iters or (prev is not None and rel > 0.5 * prev):
            return False, iters, restarts, rel
        prev = rel
        restarts += 1
        m = min(restart, max_iters - iters)
        V = torch.empty((m + 1, n, T), dtype=torch.complex128, device=dev)
        Z = torch.empty((m, n, T), dtype=torch.complex128, device=dev)
        H = torch.zeros((m + 1, m, T), dtype=torch.complex128, device=dev)
        cs = torch.zeros((m, T), dtype=torch.complex128, device=dev)
        sn = torch.zeros((m, T), dtype=torch.complex128, device=dev)
        g = torch.zeros((m + 1, T), dtype=torch.complex128, device=dev)
        safe = torch.where(beta == 0, torch.ones_like(beta), beta)
        V[0] = R / safe
        g[0] = beta.to(torch.complex128)
        k = 0
        for j in range(m):
            precond(V[j], Z[j])
            w = V[j + 1]
            E.cgemm(M, Z[j], w, alpha=1.0, beta=0.0)          # w = A z_j (FP64)
            iters += 1
            for _ in range(2):                                # classical Gram-Schmidt, twice
                h = torch.einsum('int,nt->it', V[: j + 1].conj(), w)
                w -= torch.einsum('int,it->nt', V[: j + 1], h)
                H[: j + 1, j] += h
            hn = torch.linalg.norm(w, dim=0)
            H[j + 1, j] = hn.to(torch.complex128)
            w /= torch.where(hn == 0, torch.ones_like(hn), hn)
            for i in range(j):                                # apply earlier Givens rotations
                a, b = H[i, j].clone(), H[i + 1, j].clone()
                H[i, j] = cs[i].conj() * a + sn[i].conj() * b
                H[i + 1, j] = -sn[i] * a + cs[i] * b
            a, b = H[j, j], H[j + 1, j]
            den = torch.sqrt(a.abs() ** 2 + b.abs() ** 2)
            den = torch.where(den == 0, torch.ones_like(den), den)
            cs[j] = a / den
            sn[j] = b / den
            H[j, j] = den.to(torch.complex128)
            H[j + 1, j] = 0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j].conj() * g[j]
            k = j + 1
            est = (g[j + 1].abs() / bnorm).max().item()
            if est <= 0.1 * tol or iters >= max_iters:
                break
        Hk = H[:k, :k].permute(2, 0, 1).contiguous()
        gk = g[:k].T.contiguous().unsqueeze(-1)
        y = torch.linalg.solve_triangular(Hk, gk, upper=True).squeeze(-1).T   # (k, T)
        X += torch.einsum('knt,kt->nt', Z[:k], y)
        del V, Z, H


def solve_mixed(M, M32, B, X, prec="c64", tol=1e-12, max_iters=30, restart=12, pivot=False):
    """Solve M X = B to a relative FP64 residual <= tol.  M (complex128) is
    left intact; M32 holds M in complex64 on entry (overwritten by factors).
    Returns stats: tier ('nopiv', 'pivoted' or 'fp64'), iters (FP64
    products), restarts, residual, fallback, info."""
    bnorm = torch.linalg.norm(B, dim=0)
    bnorm = torch.where(bnorm == 0, torch.ones_like(bnorm), bnorm)
    tiers = [True] if pivot else [False, True]
    total = 0
    for idx, piv in enumerate(tiers):
        if idx > 0:
            E.convert(M, M32)
        F = E.Factor(M32, prec, pivot=piv)
        if not piv and int(F.info.item()) != 0:
            continue                                       # exact zero pivot: pivot this time
        ok, iters, restarts, rel = _gmres_ir(M, F, B, X, bnorm, tol, max_iters, restart)
        total += iters
        if ok:
            return {"tier": "pivoted" if piv else "nopiv", "fallback": idx > 0, "iters": total,
                    "restarts": restarts, "residual": rel, "info": F.info}
    # refinement stalled in low precision: complex128 LU
    F64 = E.Factor(M.clone(), "c128")
    X.copy_(B)
    F64.solve(X)
    R = B.clone()
    E.cgemm(M, X, R, alpha=-1.0, beta=1.0)
    rel = (torch.linalg.norm(R, dim=0) / bnorm).max().item()
    return {"tier": "fp64", "fallback": True, "iters": total + 1, "restarts": 0, "residual": rel, "info": F64.info}
