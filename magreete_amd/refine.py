"""Mixed-precision solve: complex64 LU (FP32 or BF16x3 MFMA products) plus
FP64 GMRES iterative refinement, with fallback to the FP64 LU.

GMRES is right-preconditioned by the low-precision factors P = L U:
x = x0 + Z y with z_j = P^{-1} v_j, the Krylov basis built from A z_j in
FP64 (A = M, complex128, kept resident).  All T right-hand sides run
together: the matrix products are one (n x n) @ (n x T) FP64 GEMM and one
complex64 multi-RHS triangular solve per iteration; the per-column
Gram-Schmidt and Givens updates are small (T x j) operations.  Convergence is
declared on the true FP64 residual ||b - A x|| / ||b|| <= tol per column.
If refinement stalls (too many iterations, or restarts that do not reduce
the residual), the solve falls back to the complex128 LU and says so.
"""
import torch

from . import engine as E


def _lp(v, buf):
    """complex128 block -> complex64 working buffer."""
    E.convert(v, buf)
    return buf


def solve_mixed(M, M32, B, X, prec="c64", tol=1e-12, max_iters=30, restart=12, M_keep=True):
    """Solve M X = B.  M (complex128) is left intact; M32 holds M in
    complex64 on entry and its LU factors on exit.  Returns stats:
    iters (FP64 products), restarts, residual (max over columns), fallback, info."""
    dev = M.device
    n, T = B.shape
    F = E.Factor(M32, prec)
    info = F.info
    lowp = torch.empty((n, T), dtype=torch.complex64, device=dev)
    bnorm = torch.linalg.norm(B, dim=0)
    bnorm = torch.where(bnorm == 0, torch.ones_like(bnorm), bnorm)

    def precond(v, out):
        E.convert(v, lowp)
        F.solve(lowp)
        E.convert(lowp, out)
        return out

    # x0 = P^{-1} b
    precond(B, X)
    R = torch.empty_like(B)
    iters = 0
    restarts = 0
    prev = None
    stats = {"fallback": False}
    while True:
        R.copy_(B)
        E.cgemm(M, X, R, alpha=-1.0, beta=1.0)              # r = b - A x  (FP64)
        iters += 1
        beta = torch.linalg.norm(R, dim=0)
        rel = (beta / bnorm).max().item()
        if rel <= tol:
            break
        if iters >= max_iters or (prev is not None and rel > 0.5 * prev):
            # stalled: complex128 LU
            stats["fallback"] = True
            F64 = E.Factor(M.clone(), "c128")
            X.copy_(B)
            F64.solve(X)
            info = F64.info
            R.copy_(B)
            E.cgemm(M, X, R, alpha=-1.0, beta=1.0)
            rel = (torch.linalg.norm(R, dim=0) / bnorm).max().item()
            del F64
            break
        prev = rel
        restarts += 1
        # one GMRES cycle on A d = r, right-preconditioned
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
            w.zero_()
            E.cgemm(M, Z[j], w, alpha=1.0, beta=0.0)          # w = A z_j (FP64)
            iters += 1
            # classical Gram-Schmidt, twice
            for _ in range(2):
                h = torch.einsum('int,nt->it', V[: j + 1].conj(), w)
                w -= torch.einsum('int,it->nt', V[: j + 1], h)
                H[: j + 1, j] += h
            hn = torch.linalg.norm(w, dim=0)
            H[j + 1, j] = hn.to(torch.complex128)
            w /= torch.where(hn == 0, torch.ones_like(hn), hn)
            # Givens rotations per column
            for i in range(j):
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
        # y = H[:k,:k]^{-1} g[:k] per column; x += Z y
        Hk = H[:k, :k].permute(2, 0, 1).contiguous()
        gk = g[:k].T.contiguous().unsqueeze(-1)
        y = torch.linalg.solve_triangular(Hk, gk, upper=True).squeeze(-1).T   # (k, T)
        X += torch.einsum('knt,kt->nt', Z[:k], y)
        del V, Z, H
    stats.update({"iters": iters, "restarts": restarts, "residual": rel, "info": info})
    return stats
