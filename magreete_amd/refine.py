"""Mixed-precision solve: complex64 LU (FP32 or BF16x3 MFMA products) plus
FP64 GMRES iterative refinement, with fallbacks.

GMRES is right-preconditioned by the low-precision factors P = L U and
flexible (the preconditioned directions are kept): x = Z y with
z_j = P^{-1} v_j, the Krylov basis built from A z_j in FP64 (A = M,
complex128, kept resident), starting from x = 0 so that the first direction
is P^{-1} b at no FP64 cost.  All T right-hand sides run together: the
matrix products are one (n x n) @ (n x T) FP64 GEMM and one complex64
multi-RHS solve (GEMM-only, through stored diagonal-block inverses) per
iteration; the per-column Gram-Schmidt (classical, twice), the Givens
updates and the triangular solve for the coefficients run as HIP kernels.
Convergence is declared on the least-squares residual |g_{k+1}| / ||b|| <=
tol of every column (the norm of V_{k+1} (beta e1 - H_k y), which equals
b - A x up to the rounding of the products already taken); a cycle that
ends unconverged forms that residual explicitly and restarts from it.  When
relaxed (split-bf16) products were taken, b - A x is formed once more with an
FP64 product before a solve is reported converged.

The first factorisation is unpivoted (no pivot search, no grid barrier, no
row swap); the refinement restores the accuracy.  If it breaks down (an exact
zero pivot) or the refinement stalls, the complex64 LU is redone with partial
pivoting, and if that stalls too the solve falls back to the complex128 LU.
The returned stats say which tier answered.
"""
import torch

from . import _lib as L
from . import engine as E

# Relaxed Krylov products (relax=True).  In a Krylov method an inexact
# product M z_j + E_j z_j moves the residual by E_j z_j y_j, and the
# coefficient y_j is of the size of the residual r at the time of the product
# (inexact Krylov methods: Simoncini and Szyld 2003, Bouras and Fraysse 2005),
# so once r is small a product may carry a relative error eps with
# eps r <= RELAX_SAFETY tol.  eps is measured per solve: the refinement's
# first product is FP64, and beside it the split-bf16 levels multiply a
# sample of RELAX_PROBE_COLS of its columns; a level's eps is RELAX_AMP times
# its largest relative deviation there.  RELAX_AMP covers directions whose
# product cancels more than the first's and the solution error such a
# residual perturbation causes beyond GMRES's own residual of the same size
# (the error passes through M^{-1}).  The residual itself is not left to the
# bound: a solve that took a relaxed product forms b - M x once more with an
# FP64 product and continues in FP64 if it misses tol.
RELAX_LEVELS = ("bf16x3", "bf16x6")      # cheapest first
RELAX_SAFETY = 0.5
RELAX_AMP = 16.0
RELAX_PROBE_COLS = 8


class _SplitA:
    """M split for the relaxed products (E.split_rows), made on a side stream
    beside the first preconditioner sweep (latency-bound, it leaves HBM
    bandwidth idle; beside the memory-bound start of the refinement the split
    slowed it); get() makes the caller's stream wait for it."""

    _streams = {}   # one side stream per device, kept for the process

    def __init__(self, M):
        self.M = M
        self.AS = None

    def launch(self):
        """Enqueues the split on the side stream after the work already on
        the caller's stream (once)."""
        if self.AS is not None:
            return
        M = self.M
        dev = M.device
        main = torch.cuda.current_stream(dev)
        idx = dev.index if dev.index is not None else torch.cuda.current_device()
        if idx not in _SplitA._streams:
            _SplitA._streams[idx] = torch.cuda.Stream(dev)
        side = _SplitA._streams[idx]
        # allocated from the main stream's pool (the blocks the previous
        # factor and split released), written on the side stream
        self.AS = torch.empty(E.split_rows_bytes(*M.shape), dtype=torch.uint8, device=dev)
        self.AS.record_stream(side)
        M.record_stream(side)                  # M's block outlives the split even if no product waits for it
        side.wait_stream(main)
        with torch.cuda.stream(side):
            E.split_rows(M, out=self.AS)
            self.ready = torch.cuda.Event()
            self.ready.record(side)

    def get(self):
        self.launch()
        torch.cuda.current_stream(self.AS.device).wait_event(self.ready)
        return self.AS


class _Krylov:
    """One flexible-GMRES refinement process over a block of columns (with
    its stream, optional factor workspace and Krylov buffers), as a state
    machine: start_cycle / precond / product / after_product.

    Solves M X = B from X = 0, right-preconditioned by F; the first Krylov
    direction is P^{-1} b, so the first cycle starts with no FP64 product.
    A cycle ends converged when every column's least-squares residual
    |g_{k+1}| / ||b|| meets the tolerance; otherwise (basis full) the residual
    b - M x is formed from the Arnoldi relation M Z_k = V_{k+1} H_k, i.e.
    r = V_{k+1} (beta e1 - H_k y) with the unrotated Hessenberg matrix, at
    the cost of one pass over the basis instead of another product, and the
    next cycle restarts from it."""

    def __init__(self, M, F, B, X, bnorm, tol, max_iters, restart, stream, handle, history, relax=None, gate=None):
        self.M, self.F, self.B, self.X = M, F, B, X
        self.gate = gate            # called once after the first host sync; False stops the process
        self.relax = relax          # _SplitA (relaxed products), or None
        self.products = []          # precision of every product taken
        self.tol, self.max_iters, self.restart = tol, max_iters, restart
        self.stream, self.handle, self.history = stream, handle, history
        dev = M.device
        n, T = B.shape
        with self._ctx():
            self.bnorm = bnorm.contiguous()
            self.lowp = torch.empty((n, T), dtype=torch.complex64, device=dev)
            self.fwork = torch.empty(F.work_elems(T), dtype=F.dtype, device=dev)
            m_max = max(1, min(restart, max_iters))
            self.work = torch.empty(L.load().mg_mdot_work_elems(m_max + 1, T), dtype=torch.complex128, device=dev)
            self.est_buf = torch.empty(1, dtype=torch.float64, device=dev)
            self.hn = torch.empty(T, dtype=torch.float64, device=dev)
            X.zero_()
            self.R = B.clone(memory_format=torch.contiguous_format)
            self.beta = E.colnorm(self.R, work=self.work)
        self.iters = self.cycles = 0
        self.prev = None
        self.ok = None
        self.rel = None
        self.j = self.k = 0
        self.m = 0
        self.eps = None             # level -> relaxed product error bound, measured at the first product
        self._eps_dev = None

    def _ctx(self):
        import contextlib
        cm = contextlib.ExitStack()
        cm.enter_context(torch.cuda.stream(self.stream))
        if self.handle is not None:
            cm.enter_context(self.handle)
        return cm

    def start_cycle(self):
        """Decides (host sync) whether another cycle runs; if so sets it up.
        False once the process has converged (ok) or failed."""
        if self.ok is not None:
            return False
        with self._ctx():
            self.rel = (self.beta / self.bnorm).max().item()
            rel, tol = self.rel, self.tol
            if rel <= tol:
                self.ok = True
                return False
            if not (rel == rel) or self.iters >= self.max_iters or (self.prev is not None and rel > 0.5 * self.prev):
                self.ok = False
                return False
            self.prev = rel
            self.res = rel
            self.cycles += 1
            n, T = self.B.shape
            dev = self.M.device
            m = self.m = max(1, min(self.restart, self.max_iters - self.iters))
            self.V = torch.empty((m + 1, n, T), dtype=torch.complex128, device=dev)
            self.Z = torch.empty((m, n, T), dtype=torch.complex128, device=dev)
            self.H = torch.zeros((m + 1, m, T), dtype=torch.complex128, device=dev)
            self.H0 = torch.zeros((m + 1, m, T), dtype=torch.complex128, device=dev)   # unrotated
            self.cs = torch.zeros((m, T), dtype=torch.complex128, device=dev)
            self.sn = torch.zeros((m, T), dtype=torch.complex128, device=dev)
            self.g = torch.zeros((m + 1, T), dtype=torch.complex128, device=dev)
            safe = torch.where(self.beta == 0, torch.ones_like(self.beta), self.beta)
            torch.div(self.R, safe, out=self.V[0])
            self.g[0] = self.beta.to(torch.complex128)
            self.j = 0
            return True

    def precond(self):
        """z_j = P^{-1} v_j (enqueued on the process's stream)."""
        with self._ctx():
            E.convert(self.V[self.j], self.lowp)
            if self.relax is not None:
                self.relax.launch()
            self.F.solve(self.lowp, work=self.fwork)
            E.convert(self.lowp, self.Z[self.j])

    def product_prec(self):
        """Precision of the next product: FP64, or with relaxed products the
        cheapest level whose measured error bound times the current residual
        estimate is within RELAX_SAFETY tol."""
        if self.relax is not None and self.eps:
            for prec in RELAX_LEVELS:
                if self.eps[prec] * self.res <= RELAX_SAFETY * self.tol:
                    return prec
        return "fp64"

    def _probe_levels(self, j, w):
        """Relative deviations of the relaxed levels from the FP64 product
        w = M z_j on a sample of its columns (device tensor, one per level)."""
        T = w.shape[1]
        cols = torch.arange(0, T, max(1, T // RELAX_PROBE_COLS), device=w.device)[:RELAX_PROBE_COLS]
        z = self.Z[j][:, cols].contiguous()
        ref = w[:, cols].contiguous()
        rn = E.colnorm(ref)
        rn = torch.where(rn == 0, torch.ones_like(rn), rn)
        AS = self.relax.get()
        errs = []
        for prec in RELAX_LEVELS:
            out = torch.empty_like(ref)
            E.relaxed_product(AS, self.M.shape[1], z, out, prec)
            out -= ref
            errs.append((E.colnorm(out) / rn).max())
        return torch.stack(errs)

    def product(self):
        """w = M z_j (FP64 or relaxed), classical Gram-Schmidt twice, the
        Arnoldi column and its Givens rotations (enqueued)."""
        j = self.j
        prec = self.product_prec()
        self.products.append(prec)
        with self._ctx():
            w = self.V[j + 1]
            if prec == "fp64":
                E.cgemm(self.M, self.Z[j], w, alpha=1.0, beta=0.0, split=True)
                if self.relax is not None and self.eps is None and self._eps_dev is None:
                    self._eps_dev = self._probe_levels(j, w)
            else:
                E.relaxed_product(self.relax.get(), self.M.shape[1], self.Z[j], w, prec)
            hn = self.hn
            Vj = self.V[: j + 1]
            h1 = E.mdot(Vj, w, work=self.work)
            if j + 1 <= 16:
                h2 = E.maxpy_mdot(Vj, h1, w, work=self.work)     # w -= V h1, then V^H w
            else:
                E.maxpy(Vj, h1, w, alpha=-1.0)
                h2 = E.mdot(Vj, w, work=self.work)
            E.maxpy_norm(Vj, h2, w, hn, alpha=-1.0, work=self.work)
            self.H[: j + 1, j] += h1 + h2
            w /= torch.where(hn == 0, torch.ones_like(hn), hn)
            E.givens(self.H, self.H0, self.cs, self.sn, self.g, hn, self.bnorm, j, self.est_buf)
        self.iters += 1

    def after_product(self):
        """Reads the least-squares residual estimate (host sync); ends the
        cycle (x += Z y) when it meets the tolerance (converged) or the basis
        is full (residual from the Arnoldi relation, for a restart).  True
        when the cycle ended."""
        with self._ctx():
            est = self.est_buf.item()   # max over columns of |g_{j+1}| / ||b||
            if self._eps_dev is not None:
                dev_eps = self._eps_dev.tolist()
                self._eps_dev = None
                self.eps = {prec: RELAX_AMP * e for prec, e in zip(RELAX_LEVELS, dev_eps)}
                self.measured = dict(zip(RELAX_LEVELS, dev_eps))
            if self.gate is not None:
                gate, self.gate = self.gate, None
                if not gate():
                    self.ok = False
                    return True
            self.res = est
            if self.history is not None:
                self.history.append(est)
            self.j += 1
            if est > self.tol and self.j < self.m:
                return False
            k = self.j
            y = E.hess_solve(self.H, self.g, k)                           # (k, T)
            E.maxpy(self.Z[:k], y, self.X, alpha=1.0)                  # x += Z y
            if est <= self.tol:
                # converged on the least-squares residual |g_{k+1}| / ||b||
                self.ok, self.rel = True, est
                del self.V, self.Z, self.H, self.H0
                return True
            c = -(self.H0[: k + 1, :k] * y.unsqueeze(0)).sum(1)          # beta e1 - H_k y
            c[0] += self.beta.to(torch.complex128)
            self.R.zero_()
            E.maxpy(self.V[: k + 1], c.contiguous(), self.R, alpha=1.0)   # r = V_{k+1} c
            del self.V, self.Z, self.H, self.H0
            self.beta = E.colnorm(self.R, work=self.work)
            return True


def _gmres_ir(M, F, B, X, bnorm, tol, max_iters, restart, history=None, relax=None, products=None, gate=None,
              probe=None):
    """Solve M X = B by flexible GMRES from X = 0, right-preconditioned by F,
    all columns as one batched process on the current stream.
    history: optional list that receives the largest per-column relative
    residual estimate after every product.  relax: a _SplitA of M (relaxed
    products), or None.  products: optional list that receives the
    precision of every product.  probe: optional dict that receives the
    relaxed levels' measured relative product errors.  gate: optional callable run once after the
    first host sync (the factorisation has completed by then); False stops
    the refinement unconverged.
    Returns (converged, products, cycles, residual).

    Measured and rejected: two half-width processes on two streams whose
    preconditioner sweeps alternate beside the other half's FP64 product.
    On the 12288-row, 256-column workload a half-width product (2.6 ms) and a
    half-width preconditioner (1.7 ms) took 3.4 ms together against 6.4 ms
    for the full-width pair in sequence, i.e. 6.8 ms per iteration: the
    sweeps' waiting workgroups hold CUs that the product's one-per-CU
    workgroups then cannot use."""
    K = _Krylov(M, F, B, X, bnorm, tol, max_iters, restart, torch.cuda.current_stream(M.device), None, history,
                relax, gate)
    while K.start_cycle():
        while True:
            K.precond()
            K.product()
            if K.after_product():
                break
    if products is not None:
        products.extend(K.products)
    if probe is not None and getattr(K, "measured", None):
        probe.update(K.measured)
    return K.ok, K.iters, K.cycles, K.rel


class MixedSolver:
    """A resident mixed-precision factorisation of M with FP64 GMRES-IR solves.

    M (complex128) stays intact and must outlive the solver; M32 holds M in
    complex64 on entry and is overwritten by the factors.  The tier starts
    at the unpivoted block LDU ('nopiv', or 'pivoted' with pivot=True); a
    solve whose refinement breaks down or stalls moves the solver to the
    next tier (complex64 LU with partial pivoting, then complex128 LU) for
    this and every later solve.

    relax=True: the refinement's late products run on split-bf16 MFMA when
    the residual allows (RELAX_LEVELS); M is split once, at the first solve,
    on a side stream (12 bytes per element, resident beside M).  The
    residual then holds to (1 + RELAX_SAFETY) tol while the levels' bounds
    hold."""

    TIERS = ("nopiv", "pivoted", "fp64")

    def __init__(self, M, M32, prec="c64", tol=1e-12, max_iters=30, restart=12, pivot=False, relax=False):
        self.M, self.M32, self.prec = M, M32, prec
        self.tol, self.max_iters, self.restart = tol, max_iters, restart
        self.tier = "pivoted" if pivot else "nopiv"
        self.fell = False
        self.F = None
        self.relax = relax
        self.AS = None           # _SplitA of M for the relaxed products
        self._gate = None        # deferred zero-pivot check of the unpivoted factorisation
        self._factor(fresh=True)

    def _factor(self, fresh=False):
        if self.tier == "fp64":
            self.M32 = self.AS = None
            self.F = E.Factor(self.M.clone(), "c128")
            return
        if self.M32 is None:
            self.M32 = torch.empty(self.M.shape, dtype=torch.complex64, device=self.M.device)
            fresh = False
        if not fresh:
            E.convert(self.M, self.M32)                    # the previous tier overwrote M32
        self.F = E.Factor(self.M32, self.prec, pivot=(self.tier == "pivoted"))
        if self.tier == "nopiv":
            # an exact zero pivot moves the solve to the pivoted tier; the
            # check waits for the refinement's first host sync instead of
            # stalling the stream here
            info = torch.empty(1, dtype=torch.int32, pin_memory=True)
            info.copy_(self.F.info, non_blocking=True)
            done = torch.cuda.Event()
            done.record(torch.cuda.current_stream(self.M.device))

            def gate():
                done.synchronize()
                return int(info[0]) == 0
            self._gate = gate
        if self.F.A is None:
            self.M32 = None                                # the split factor does not read it again

    def _escalate(self):
        self.tier = self.TIERS[self.TIERS.index(self.tier) + 1]
        self.fell = True
        self._factor()

    @property
    def info(self):
        return self.F.info

    def _true_residual(self, B, X, bnorm):
        """R = B - M X with an FP64 product; returns (R, max_t ||R_t|| / ||b_t||)."""
        R = B.clone(memory_format=torch.contiguous_format)
        E.cgemm(self.M, X, R, alpha=-1.0, beta=1.0, split=True)
        return R, (E.colnorm(R) / bnorm).max().item()

    def solve(self, B, X):
        """X <- M^{-1} B to a relative FP64 residual <= tol per column.
        Returns stats: tier, fallback, iters (products), cycles, residual,
        residual_kind, history (residual estimate after every product),
        products (the precision of every product).

        residual is max_t ||b_t - M x_t|| / ||b_t||.  With FP64 products only
        it is GMRES's least-squares residual (the Arnoldi relation holds to
        the rounding of FP64 products; residual_kind 'arnoldi').  When a
        relaxed product was taken the relation carries that product's error,
        so the residual is formed once more with an FP64 product (kind
        'fp64'); if it misses tol, GMRES continues from it with FP64 products
        only and the reported residual is that continuation's."""
        bnorm = E.colnorm(B.contiguous())
        bnorm = torch.where(bnorm == 0, torch.ones_like(bnorm), bnorm)
        total = 0
        hist = []
        prods = []
        probe = {}
        while self.tier != "fp64":
            if self.relax and self.AS is None:
                self.AS = _SplitA(self.M)
            gate, self._gate = self._gate, None
            ok, iters, cycles, rel = _gmres_ir(self.M, self.F, B, X, bnorm, self.tol, self.max_iters, self.restart,
                                               hist, self.AS, prods, gate, probe)
            total += iters
            kind = "arnoldi"
            if ok and any(p != "fp64" for p in prods):
                R, rel = self._true_residual(B, X, bnorm)
                prods.append("fp64-check")
                total += 1
                kind = "fp64"
                hist.append(rel)
                if rel > self.tol:
                    # the relaxed products' errors left the true residual above tol:
                    # solve for the correction from it, FP64 products only
                    D = torch.empty_like(X)
                    ok, iters, c2, rel = _gmres_ir(self.M, self.F, R, D, bnorm, self.tol, self.max_iters,
                                                   self.restart, hist, None, prods, None)
                    total += iters
                    cycles += c2
                    kind = "arnoldi"
                    X += D
            if ok:
                st = {"tier": self.tier, "fallback": self.fell, "iters": total, "cycles": cycles,
                      "residual": rel, "residual_kind": kind, "history": hist, "products": prods}
                if probe:
                    st["relax_probe"] = dict(probe)
                return st
            self._escalate()
        # refinement stalled in low precision: complex128 LU
        X.copy_(B)
        self.F.solve(X)
        R = B.clone()
        E.cgemm(self.M, X, R, alpha=-1.0, beta=1.0, split=True)
        rel = (E.colnorm(R) / bnorm).max().item()
        return {"tier": "fp64", "fallback": True, "iters": total + 1, "cycles": 0, "residual": rel,
                "residual_kind": "fp64"}


def solve_mixed(M, M32, B, X, prec="c64", tol=1e-12, max_iters=30, restart=12, pivot=False, relax=False):
    """Solve M X = B to a relative FP64 residual <= tol (one MixedSolver,
    one solve).  Returns stats: tier ('nopiv', 'pivoted' or 'fp64'), iters
    (products), cycles (GMRES restarts + 1), residual, fallback, info."""
    ms = MixedSolver(M, M32, prec, tol=tol, max_iters=max_iters, restart=restart, pivot=pivot, relax=relax)
    st = ms.solve(B, X)
    st["info"] = ms.info
    return st
