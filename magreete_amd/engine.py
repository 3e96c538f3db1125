"""Device-side operations of the coupled-dipole path, over libmagreete_hip.

Every function takes and returns torch tensors on a ROCm device and launches
HIP kernels on torch's current stream of that device.  Matrices are row-major
complex128 (or complex64 for the factor precisions), with the reference's
interleaved layout: row d*m + a, column d*j + b.
"""
import ctypes
import threading

import torch

from . import _lib as L

KINDS = {"2d_tm": L.MG_2D_TM, "2d_te": L.MG_2D_TE, "3d_vector": L.MG_3D_VECTOR, "3d_scalar": L.MG_3D_SCALAR}
DIM = {"2d_tm": 1, "2d_te": 2, "3d_vector": 3, "3d_scalar": 1}
SPACE = {"2d_tm": 2, "2d_te": 2, "3d_vector": 3, "3d_scalar": 3}
PRECS = {"c128": L.MG_C128, "c64": L.MG_C64, "bf16x3": L.MG_BF16X3, "bf16x6": L.MG_BF16X6}

_handles = {}


def device_of(t=None):
    """Target device: the tensor's if it is on a GPU, else cuda:0."""
    if t is not None and t.is_cuda:
        return t.device
    if not torch.cuda.is_available():
        raise RuntimeError("magreete_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


_tls = threading.local()


def handle(device):
    """The library handle (LU workspace, look-ahead stream) in effect for this
    thread: the one installed by a Workspace context, else the device's
    shared handle, created on first use."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    own = getattr(_tls, "handles", None)
    if own and idx in own:
        return own[idx]
    h = _handles.get(idx)
    if h is None:
        h = _create(idx)
        _handles[idx] = h
    return h


def _create(idx):
    lib = L.load()
    with torch.cuda.device(idx):
        hp = ctypes.c_void_p()
        L.check(lib.mg_create(ctypes.byref(hp)), "mg_create")
    return hp


class Workspace:
    """A private library handle for one device, for concurrent solves.

    Factorisations sequence their look-ahead through the handle's second
    stream and events, so two solves that run at the same time (different
    torch streams, usually different host threads) need different handles.
    Used as a context manager it becomes this thread's handle for `device`."""

    def __init__(self, device, leaf_blocks=0):
        self.idx = torch.device(device).index
        if self.idx is None:
            self.idx = torch.cuda.current_device()
        self.h = _create(self.idx)
        self._prev = None
        if leaf_blocks:
            # workgroups of one LU leaf launch: concurrent factorisations keep
            # their sum within the CU count (the leaf waits on a grid barrier)
            L.check(L.load().mg_set_leaf_blocks(self.h, int(leaf_blocks)), "mg_set_leaf_blocks")

    def __enter__(self):
        own = getattr(_tls, "handles", None)
        if own is None:
            own = _tls.handles = {}
        self._prev = own.get(self.idx)
        own[self.idx] = self.h
        return self

    def __exit__(self, *exc):
        own = _tls.handles
        if self._prev is None:
            own.pop(self.idx, None)
        else:
            own[self.idx] = self._prev
        return False

    def close(self):
        """Destroy the handle (waits for nothing: synchronise its streams first)."""
        if self.h is not None:
            with torch.cuda.device(self.idx):
                L.check(L.load().mg_destroy(self.h), "mg_destroy")
            self.h = None


def _pts(x, device):
    return x.to(device=device, dtype=torch.float64).contiguous()


def _ld(t):
    if t.dim() != 2 or (t.stride(1) != 1 and t.shape[1] > 1):
        raise ValueError("expected a row-major 2-D tensor")
    if t.shape[0] <= 1:
        return max(t.shape[1], 1)      # torch leaves the stride of a size-1 dimension arbitrary
    return t.stride(0)


# ---------------------------------------------------------------------------
def assemble(kind, rows, cols, k0, scale=1.0, flags=0, diag=0.0, fill=0.0, radius=0.0, out=None, out64=None,
             want64=False):
    """out[d i + a, d j + b] = scale * G_ab(rows[i] - cols[j]) with the MG_F_* fix-ups.

    Returns (out_c128 or None, out_c64 or None)."""
    lib = L.load()
    dev = device_of(rows)
    d = DIM[kind]
    rows = _pts(rows, dev)
    cols = _pts(cols, dev)
    nr, nc = rows.shape[0], cols.shape[0]
    if out is None and out64 is None:
        out = torch.empty((d * nr, d * nc), dtype=torch.complex128, device=dev)
    if want64 and out64 is None:
        out64 = torch.empty((d * nr, d * nc), dtype=torch.complex64, device=dev)
    scale, diag, fill = complex(scale), complex(diag), complex(fill)
    with torch.cuda.device(dev):
        st = lib.mg_assemble(KINDS[kind], _p(rows), nr, _p(cols), nc, float(k0), scale.real, scale.imag,
                             diag.real, diag.imag, fill.real, fill.imag, int(flags), float(radius),
                             _p(out), _ld(out) if out is not None else 0,
                             _p(out64), _ld(out64) if out64 is not None else 0, _stream(dev))
    L.check(st, "mg_assemble")
    return out, out64


def getrf(A, prec="c128"):
    """In-place LU with partial pivoting; returns (ipiv int32 (n,), info int32 (1,)) on the device."""
    lib = L.load()
    dev = A.device
    n = A.shape[0]
    ipiv = torch.empty(n, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_getrf(handle(dev), PRECS[prec], _p(A), n, _ld(A), _p(ipiv), _p(info), _stream(dev))
    L.check(st, "mg_getrf")
    return ipiv, info


def getrs(LU, ipiv, B, prec="c128"):
    """In place: B <- A^{-1} B from getrf's factors."""
    lib = L.load()
    dev = LU.device
    with torch.cuda.device(dev):
        st = lib.mg_getrs(handle(dev), PRECS[prec], _p(LU), LU.shape[0], _ld(LU), _p(ipiv), _p(B), _ld(B),
                          B.shape[1], _stream(dev))
    L.check(st, "mg_getrs")
    return B


class Factor:
    """A factorisation resident on the device, every later solve GEMM-only.

    pivot=True: LU with partial pivoting (getrf) plus the inverses of the
    256 x 256 diagonal blocks of L and U.  pivot=False (complex-float
    precisions only): block LDU without pivoting, A = Lh D Uh, with the
    inverses of the D blocks; the preconditioner of the mixed-precision
    refinement."""

    def __init__(self, A, prec="c128", pivot=True):
        lib = L.load()
        self.A, self.prec, self.n, self.pivot = A, prec, A.shape[0], pivot
        self.dtype, self.device = A.dtype, A.device
        dev = A.device
        if pivot:
            self.ipiv, self.info = getrf(A, prec)
            self.dinv = torch.empty(lib.mg_dinv_elems(self.n), dtype=A.dtype, device=dev)
            with torch.cuda.device(dev):
                st = lib.mg_getrf_dinv(handle(dev), PRECS[prec], _p(A), self.n, _ld(A), _p(self.dinv),
                                       _stream(dev))
            L.check(st, "mg_getrf_dinv")
        else:
            if A.dtype != torch.complex64:
                raise ValueError("the unpivoted block LDU factors complex64 matrices")
            self.ipiv = None
            self.info = torch.zeros(1, dtype=torch.int32, device=dev)
            # D inverses, then (split precisions) the pre-split panels
            self.dinv = torch.empty(lib.mg_ldu_factor_elems(PRECS[prec], self.n), dtype=A.dtype, device=dev)
            work = torch.empty(lib.mg_ldu_work_elems(self.n), dtype=A.dtype, device=dev)
            with torch.cuda.device(dev):
                st = lib.mg_getrf_ldu(handle(dev), PRECS[prec], _p(A), self.n, _ld(A), _p(self.dinv), _p(work),
                                      _p(self.info), _stream(dev))
            L.check(st, "mg_getrf_ldu")
            del work          # stream-ordered: the factorisation's last step waits on this stream
            if prec in ("bf16x3", "bf16x6"):
                # the solves read the pre-split panels and the D inverses only:
                # the factorised matrix's storage may go (the caller's copy is
                # the caller's to keep or free)
                self.A = None
        self._work = None

    consumed = False      # set when inverse_diag_blocks overwrote the factor

    def work_elems(self, nrhs):
        """Elements of the factor's dtype that a solve with nrhs columns needs as work."""
        if self.pivot:
            return self.n * nrhs
        return int(L.load().mg_ldu_solve_work_elems(PRECS[self.prec], self.n, nrhs))

    def solve(self, B, work=None):
        """In place: B <- A^{-1} B (B n x nrhs, the factor's dtype, row-major).
        work: work_elems(nrhs) elements of the factor's dtype owned by the
        caller (solves that run at the same time on one factor each bring
        their own); default: a buffer kept by the factor."""
        if self.consumed:
            raise RuntimeError("this factor was consumed by inverse_diag_blocks")
        lib = L.load()
        dev = self.device
        need = self.work_elems(B.shape[1])
        if work is None:
            if self._work is None or self._work.numel() < need:
                self._work = torch.empty(need, dtype=self.dtype, device=dev)
            work = self._work
        elif work.numel() < need or work.dtype != self.dtype:
            raise ValueError("work needs %d elements of %s" % (need, self.dtype))
        with torch.cuda.device(dev):
            if self.pivot:
                st = lib.mg_getrs_dinv(handle(dev), PRECS[self.prec], _p(self.A), self.n, _ld(self.A),
                                       _p(self.ipiv), _p(self.dinv), _p(B), _ld(B), B.shape[1], _p(work),
                                       _stream(dev))
            else:
                lda = _ld(self.A) if self.A is not None else self.n
                st = lib.mg_getrs_ldu(handle(dev), PRECS[self.prec], _p(self.A), self.n, lda,
                                      _p(self.dinv), _p(B), _ld(B), B.shape[1], _p(work), _stream(dev))
        L.check(st, "mg_getrs")
        return B


def colperm_from_ipiv(ipiv):
    """The position that getrf's row swaps (ipiv, applied in order) move each
    row to, as an int32 tensor on ipiv's device: with P M = L U, column c of
    L^{-1} P is column colperm[c] of L^{-1}."""
    ip = ipiv.cpu().tolist()
    n = len(ip)
    perm = list(range(n))
    for i, j in enumerate(ip):
        if j != i:
            perm[i], perm[j] = perm[j], perm[i]
    q = [0] * n
    for pos, row in enumerate(perm):
        q[row] = pos
    return torch.tensor(q, dtype=torch.int32, device=ipiv.device)


def inverse_diag_blocks(F, sites, d):
    """(len(sites), d, d) complex128: the diagonal blocks of M^{-1} at the
    given site indices, from a pivoted complex128 Factor of M, which this
    consumes (its storage is overwritten by the triangular inverses)."""
    if not F.pivot or F.dtype != torch.complex128 or F.A is None:
        raise ValueError("inverse_diag_blocks needs a resident complex128 LU with partial pivoting")
    lib = L.load()
    dev = F.device
    n = F.n
    sites = torch.as_tensor(sites, dtype=torch.int64)
    if sites.numel() and (int(sites.min()) < 0 or d * (int(sites.max()) + 1) > n):
        raise ValueError("site indices must lie in [0, n / d)")
    sites = sites.to(dev).contiguous()
    out = torch.empty((sites.shape[0], d, d), dtype=torch.complex128, device=dev)
    q = colperm_from_ipiv(F.ipiv)
    we = int(lib.mg_inverse_diag_work_elems(n))
    work = torch.empty(max(we, 1), dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_inverse_diag_blocks(handle(dev), L.MG_C128, _p(F.A), n, _ld(F.A), _p(F.dinv), _p(q), int(d),
                                        _p(sites), sites.shape[0], _p(out), _p(work), we, _stream(dev))
    L.check(st, "mg_inverse_diag_blocks")
    F.consumed = True
    return out


def profile_updates(device, on):
    """Enable/disable HIP-event timing of the LU trailing-update launches."""
    L.check(L.load().mg_profile_enable(handle(device), 1 if on else 0), "mg_profile_enable")


def read_update_profile(device):
    """(launches, algorithmic flops, milliseconds) of the trailing updates since the last read."""
    n = ctypes.c_longlong()
    f = ctypes.c_double()
    ms = ctypes.c_double()
    with torch.cuda.device(device):
        L.check(L.load().mg_profile_read(handle(device), _stream(device), ctypes.byref(n), ctypes.byref(f),
                                         ctypes.byref(ms)), "mg_profile_read")
    return n.value, f.value, ms.value


def read_update_profile_classes(device):
    """As read_update_profile, per kernel class: {"reg": (launches, flops, ms),
    "dma": (...)} -- operands staged through registers, or DMA'd to LDS."""
    n = (ctypes.c_longlong * 2)()
    f = (ctypes.c_double * 2)()
    ms = (ctypes.c_double * 2)()
    with torch.cuda.device(device):
        L.check(L.load().mg_profile_read_classes(handle(device), _stream(device), n, f, ms),
                "mg_profile_read_classes")
    return {"reg": (n[0], f[0], ms[0]), "dma": (n[1], f[1], ms[1])}


def check_barriers(device):
    """Synchronises the current stream and raises if a grid barrier of the
    LU leaf on this thread's handle timed out since the last check (the
    report is cleared by the check)."""
    lib = L.load()
    with torch.cuda.device(device):
        L.check(lib.mg_check(handle(device), _stream(device)), "mg_check")


def check_info(info, pivoted=True):
    """Raises on a factorisation status read from the device: negative is a
    failed leaf (its panel was written as NaN), positive a zero pivot
    (singular M) when the factor is the pivoted LU."""
    i = int(info.item()) if torch.is_tensor(info) else int(info)
    if i < 0:
        raise L.MGError(i, "LU factorisation (a grid barrier of its leaf timed out)")
    if i > 0 and pivoted:
        raise torch.linalg.LinAlgError("the system matrix is singular (U[%d,%d] is exactly zero)" % (i, i))
    return i


def cgemm(A, B, C, alpha=1.0, beta=0.0, prec="c128", split=False):
    """C <- alpha A B + beta C (row-major).  split: a tall-skinny product may
    split K over a workspace (torch's caching allocator) so that its tile
    grid fills the CUs; the partial sums are reduced in a fixed order."""
    lib = L.load()
    dev = C.device
    alpha, beta = complex(alpha), complex(beta)
    m, n, k = C.shape[0], C.shape[1], A.shape[1]
    with torch.cuda.device(dev):
        if split:
            elems = lib.mg_cgemm_work_elems(PRECS[prec], m, n, k)
            work = torch.empty(max(int(elems), 1), dtype=C.dtype, device=dev)
            st = lib.mg_cgemm_ws(PRECS[prec], m, n, k, alpha.real, alpha.imag, _p(A), _ld(A), _p(B), _ld(B),
                                 beta.real, beta.imag, _p(C), _ld(C), _p(work), int(elems), _stream(dev))
        else:
            st = lib.mg_cgemm(PRECS[prec], m, n, k, alpha.real, alpha.imag, _p(A), _ld(A), _p(B), _ld(B),
                              beta.real, beta.imag, _p(C), _ld(C), _stream(dev))
    L.check(st, "mg_cgemm")
    return C


def split_rows_bytes(rows, cols):
    return int(L.load().mg_split_rows_bytes(rows, cols))


def split_rows(A, out=None):
    """A (complex128 or complex64, row-major) split once for relaxed_product:
    every element rounded to complex float and held as three bf16 parts of
    its real and imaginary part (a uint8 buffer of split_rows_bytes, 12 bytes
    per element; out: the caller's)."""
    lib = L.load()
    dev = A.device
    rows, cols = A.shape
    prec = L.MG_C128 if A.dtype == torch.complex128 else L.MG_C64
    nb = split_rows_bytes(rows, cols)
    if out is None:
        out = torch.empty(nb, dtype=torch.uint8, device=dev)
    elif out.numel() < nb or out.dtype != torch.uint8:
        raise ValueError("split_rows needs %d bytes" % nb)
    AS = out
    with torch.cuda.device(dev):
        st = lib.mg_split_rows(prec, _p(A), _ld(A), rows, cols, _p(AS), _stream(dev))
    L.check(st, "mg_split_rows")
    return AS


def relaxed_product(AS, k, B, C, prec):
    """C <- A B (complex128) with A (C.shape[0] x k) given by split_rows and
    the products on split-bf16 MFMA (prec 'bf16x3': relative error ~1e-5,
    'bf16x6': ~1e-7); the Krylov products that may be inexact."""
    lib = L.load()
    dev = C.device
    m, n = C.shape
    nb = int(lib.mg_relaxed_product_work_bytes(PRECS[prec], m, n, k))
    if nb < 0:
        raise ValueError("relaxed products are bf16x3 or bf16x6, not %r" % prec)
    work = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_relaxed_product(PRECS[prec], m, n, k, _p(AS), _p(B), _ld(B), _p(C), _ld(C), _p(work), nb,
                                    _stream(dev))
    L.check(st, "mg_relaxed_product")
    return C


def convert(src, dst):
    lib = L.load()
    dev = src.device
    direction = 0 if src.dtype == torch.complex128 else 1
    with torch.cuda.device(dev):
        st = lib.mg_convert(direction, _p(src), _ld(src), _p(dst), _ld(dst), src.shape[0], src.shape[1],
                            _stream(dev))
    L.check(st, "mg_convert")
    return dst


def propagate(kind, meas, scat, k0, alpha, Ek, out, flags=0, radius=0.0, beta=1.0):
    """out <- alpha * G(meas, scat) @ Ek + beta * out, G evaluated in-kernel
    (the scatterers split over a workspace from torch's allocator when the
    output tiles alone would leave CUs idle)."""
    lib = L.load()
    dev = out.device
    meas = _pts(meas, dev)
    scat = _pts(scat, dev)
    alpha, beta = complex(alpha), complex(beta)
    nb = int(lib.mg_propagate_work_bytes(KINDS[kind], meas.shape[0], scat.shape[0], Ek.shape[1]))
    work = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_propagate_ws(KINDS[kind], _p(meas), meas.shape[0], _p(scat), scat.shape[0], float(k0),
                                 alpha.real, alpha.imag, int(flags), float(radius), _p(Ek), _ld(Ek), Ek.shape[1],
                                 beta.real, beta.imag, _p(out), _ld(out), _p(work), nb, _stream(dev))
    L.check(st, "mg_propagate")
    return out


def propagate_src(kind, meas, scat, k0, alpha, Ek, out, beam, dirs, pol, w, flags=0, radius=0.0):
    """out <- alpha * G(meas, scat) @ Ek + E0(meas), the incident plane waves
    (beam=False) or Gaussian beams evaluated in the propagation's final pass
    (dirs, pol as for sources; out (d M, T) is write-only)."""
    lib = L.load()
    dev = out.device
    meas = _pts(meas, dev)
    scat = _pts(scat, dev)
    alpha = complex(alpha)
    dirs = dirs.to(device=dev, dtype=torch.float64).contiguous()
    if dirs.shape[0] != Ek.shape[1] or out.shape != (DIM[kind] * meas.shape[0], Ek.shape[1]):
        raise ValueError("propagate_src: dirs (T, S), Ek (d n, T) and out (d M, T) must agree")
    if pol is not None:
        pol = pol.to(device=dev, dtype=torch.float64).contiguous()
        if pol.shape != (dirs.shape[0], DIM[kind]):
            raise ValueError("propagate_src: pol must be (T, d)")
    nb = int(lib.mg_propagate_work_bytes(KINDS[kind], meas.shape[0], scat.shape[0], Ek.shape[1]))
    work = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_propagate_src(KINDS[kind], _p(meas), meas.shape[0], _p(scat), scat.shape[0], float(k0),
                                  alpha.real, alpha.imag, int(flags), float(radius), _p(Ek), _ld(Ek), Ek.shape[1],
                                  1 if beam else 0, _p(dirs), _p(pol), float(w), _p(out), _ld(out), _p(work), nb,
                                  _stream(dev))
    L.check(st, "mg_propagate_src")
    return out


class GroupLU:
    """Grouped LU with partial pivoting of up to 256 independent square
    matrices of uneven order, factored in lockstep (one launch per step of
    the blocked algorithm for every problem with work at that column), and
    the grouped solve of their right-hand sides.

    mats: device tensors (n_i, n_i), complex128 (prec 'c128') or complex64
    ('c64', 'bf16x3': trailing updates on BF16x3 products), row-major,
    factored in place; rhs: optional (n_i, nrhs) tensors of the same dtype
    (one nrhs for all), solved in place by solve().  The factors, pivots and
    info per problem follow getrf; pivot ('bf16x3' only): 'tournament' (the
    default) picks each leaf's pivots by partial pivoting among the rows
    that chunks of 256 rows offer, 'partial' as getrf."""

    MAX = 256
    PIVOTS = {"partial": 0, "tournament": 1}

    def __init__(self, mats, rhs=None, prec="c128", nb=256, leaf=16, tile=0, leaf_threads=1024,
                 pivot="tournament"):
        lib = L.load()
        if not 0 < len(mats) <= self.MAX:
            raise ValueError("a group holds 1 to %d problems" % self.MAX)
        dt = torch.complex128 if prec == "c128" else torch.complex64
        dev = mats[0].device
        for A in mats:
            if A.dtype != dt or A.device != dev or A.dim() != 2 or A.shape[0] != A.shape[1] or A.stride(1) != 1:
                raise ValueError("group matrices: square, row-major, %s on one device" % dt)
        if rhs is not None:
            if len(rhs) != len(mats):
                raise ValueError("one right-hand side block per matrix")
            nrhs = rhs[0].shape[1]
            for A, B in zip(mats, rhs):
                if B.dtype != dt or B.device != dev or B.shape != (A.shape[0], nrhs) or B.stride(1) != 1:
                    raise ValueError("right-hand sides: (n_i, nrhs) row-major, the matrices' dtype")
        self.mats, self.rhs, self.prec, self.device = list(mats), rhs, prec, dev
        P = len(mats)
        ns = [A.shape[0] for A in mats]
        self.off = [0]
        for n in ns:
            self.off.append(self.off[-1] + n)
        self.ipiv_all = torch.empty(self.off[-1], dtype=torch.int32, device=dev)
        self.info = torch.zeros(P, dtype=torch.int32, device=dev)
        arr = lambda ct, xs: (ct * P)(*xs)
        vp = ctypes.c_void_p
        self._keep = (
            arr(vp, [_p(A) for A in mats]), arr(ctypes.c_int64, ns), arr(ctypes.c_int64, [_ld(A) for A in mats]),
            arr(vp, [self.ipiv_all.data_ptr() + 4 * o for o in self.off[:-1]]),
            arr(vp, [self.info.data_ptr() + 4 * p for p in range(P)]),
            arr(vp, [_p(B) for B in rhs]) if rhs is not None else None,
            arr(ctypes.c_int64, [_ld(B) for B in rhs]) if rhs is not None else None)
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            L.check(lib.mg_group_create(PRECS[prec], P, *self._keep, nrhs if rhs is not None else 0,
                                        ctypes.byref(h)), "mg_group_create")
        self.h = h
        self._captured = set()
        if pivot not in self.PIVOTS:
            raise ValueError("pivot: 'partial' or 'tournament'")
        L.check(lib.mg_group_configure(self.h, int(nb), int(leaf), int(tile), int(leaf_threads),
                                       self.PIVOTS[pivot]), "mg_group_configure")

    def ipiv(self, i):
        return self.ipiv_all[self.off[i]:self.off[i + 1]]

    def tune(self):
        """Pick the tunables by timing candidates on a sample of this group's
        matrices (copies; call before factor) unless a group of the same size
        class was tuned before in this process.  Returns ((nb, leaf, tile,
        leaf_threads), best ms or 0 for a cached choice)."""
        cfg = (ctypes.c_int * 4)()
        ms = ctypes.c_float()
        with torch.cuda.device(self.device):
            L.check(L.load().mg_group_tune(self.h, cfg, ctypes.byref(ms), _stream(self.device)), "mg_group_tune")
        self._captured.clear()      # sequences captured before hold the previous tunables
        return tuple(cfg), ms.value

    def factor(self, graph=False):
        """Factor in place (graph: replay the captured factor sequence,
        capturing it on first use)."""
        with torch.cuda.device(self.device):
            if graph:
                self._replay(0)
            else:
                L.check(L.load().mg_group_getrf(self.h, _stream(self.device)), "mg_group_getrf")
        return self

    def solve(self, graph=False):
        """Solve the right-hand sides in place from the factors (graph: replay
        the captured solve sequence, capturing it on first use)."""
        if self.rhs is None:
            raise ValueError("this group has no right-hand sides")
        with torch.cuda.device(self.device):
            if graph:
                self._replay(1)
            else:
                L.check(L.load().mg_group_getrs(self.h, _stream(self.device)), "mg_group_getrs")
        return self.rhs

    def set_aux(self, X):
        """Register one (n_i, nrhs) block per problem (the matrices' dtype,
        row-major) as the X of residual()."""
        if self.rhs is None or len(X) != len(self.mats):
            raise ValueError("one auxiliary block per problem of a group with right-hand sides")
        for A, B, x in zip(self.mats, self.rhs, X):
            if x.dtype != A.dtype or x.device != self.device or x.shape != B.shape or x.stride(1) != 1:
                raise ValueError("auxiliary blocks: the right-hand sides' shape, the matrices' dtype")
        P = len(X)
        self._aux = ((ctypes.c_void_p * P)(*[_p(x) for x in X]), (ctypes.c_int64 * P)(*[_ld(x) for x in X]))
        self._aux_t = list(X)
        L.check(L.load().mg_group_set_aux(self.h, *self._aux), "mg_group_set_aux")

    def residual(self):
        """B_i -= A_i X_i for every problem in one launch (A_i unfactored)."""
        with torch.cuda.device(self.device):
            L.check(L.load().mg_group_residual(self.h, _stream(self.device)), "mg_group_residual")

    def colnorms(self, out=None):
        """(P, nrhs) float64: the column norms of every problem's right-hand sides."""
        if out is None:
            out = torch.empty((len(self.mats), self.rhs[0].shape[1]), dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            L.check(L.load().mg_group_colnorm(self.h, _p(out), _stream(self.device)), "mg_group_colnorm")
        return out

    def _replay(self, op):
        lib = L.load()
        if op not in self._captured:
            L.check(lib.mg_group_capture(self.h, op), "mg_group_capture")
            self._captured.add(op)
        L.check(lib.mg_group_replay(self.h, op, _stream(self.device)), "mg_group_replay")

    @property
    def launches(self):
        return int(L.load().mg_group_launches(self.h))

    def close(self):
        if getattr(self, "h", None):
            L.load().mg_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def coldot(Y, X):
    """out[c] = sum_r Y[r, c] X[r, c] (no conjugation), complex128."""
    lib = L.load()
    dev = Y.device
    out = torch.empty(Y.shape[1], dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_coldot(_p(Y), _ld(Y), _p(X), _ld(X), Y.shape[0], Y.shape[1], _p(out), _stream(dev))
    L.check(st, "mg_coldot")
    return out


def _vecs(V):
    """(pointer, vector stride, row stride) of a stack V (k, n, T) of row-major matrices."""
    k, n, T = V.shape
    if V.dim() != 3 or (V.stride(2) != 1 and T > 1):
        raise ValueError("Krylov vectors: (k, n, T) with unit column stride")
    ld = V.stride(1) if n > 1 else max(T, 1)           # size-1 dimensions carry arbitrary strides
    sv = V.stride(0) if k > 1 else n * ld
    return _p(V), sv, ld


def mdot(V, w, out=None, work=None):
    """h[i, t] = sum_r conj(V[i, r, t]) w[r, t] (per-column dot products, fixed summation order)."""
    lib = L.load()
    dev = w.device
    k, n, T = V.shape
    if out is None:
        out = torch.empty((k, T), dtype=torch.complex128, device=dev)
    need = lib.mg_mdot_work_elems(k, T)
    if work is None or work.numel() < need:
        work = torch.empty(max(need, 1), dtype=torch.complex128, device=dev)
    pv, sv, ldv = _vecs(V)
    with torch.cuda.device(dev):
        st = lib.mg_mdot(pv, sv, ldv, k, _p(w), _ld(w), n, T, _p(work), _p(out), _ld(out), _stream(dev))
    L.check(st, "mg_mdot")
    return out


def maxpy(V, h, w, alpha=1.0):
    """w[r, t] += alpha sum_i V[i, r, t] h[i, t] (in place)."""
    lib = L.load()
    dev = w.device
    k, n, T = V.shape
    pv, sv, ldv = _vecs(V)
    with torch.cuda.device(dev):
        st = lib.mg_maxpy(pv, sv, ldv, k, _p(h), _ld(h), float(alpha), _p(w), _ld(w), n, T, _stream(dev))
    L.check(st, "mg_maxpy")
    return w


def maxpy_mdot(V, h, w, out=None, work=None):
    """w <- w - V h (per column), then out[i, t] = sum_r conj(V[i, r, t]) w[r, t]
    with the updated w, in one read of V (k <= 16 vectors)."""
    lib = L.load()
    dev = w.device
    k, n, T = V.shape
    if out is None:
        out = torch.empty((k, T), dtype=torch.complex128, device=dev)
    need = lib.mg_mdot_work_elems(k, T)
    if work is None or work.numel() < need:
        work = torch.empty(max(need, 1), dtype=torch.complex128, device=dev)
    pv, sv, ldv = _vecs(V)
    with torch.cuda.device(dev):
        st = lib.mg_maxpy_mdot(pv, sv, ldv, k, _p(h), _ld(h), _p(w), _ld(w), n, T, _p(work), _p(out), _ld(out),
                               _stream(dev))
    L.check(st, "mg_maxpy_mdot")
    return out


def maxpy_norm(V, h, w, norms, alpha=1.0, work=None):
    """maxpy, then norms[t] = ||w[:, t]|| of the updated w (fixed summation
    order); work: optional complex128 buffer of mg_mdot_work_elems(1, T)."""
    lib = L.load()
    dev = w.device
    k, n, T = V.shape
    need = lib.mg_mdot_work_elems(1, T)
    if work is None or work.numel() < need:
        work = torch.empty(max(need, 1), dtype=torch.complex128, device=dev)
    pv, sv, ldv = _vecs(V)
    with torch.cuda.device(dev):
        st = lib.mg_maxpy_norm(pv, sv, ldv, k, _p(h), _ld(h), float(alpha), _p(w), _ld(w), n, T, _p(work),
                               _p(norms), _stream(dev))
    L.check(st, "mg_maxpy_norm")
    return norms


def colnorm(X, out=None, work=None):
    """out[t] = ||X[:, t]|| (complex128 X, float64 out; fixed summation order)."""
    lib = L.load()
    dev = X.device
    n, T = X.shape
    if out is None:
        out = torch.empty(T, dtype=torch.float64, device=dev)
    need = lib.mg_mdot_work_elems(1, T)
    if work is None or work.numel() < need:
        work = torch.empty(max(need, 1), dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_colnorm(_p(X), _ld(X), n, T, _p(work), _p(out), _stream(dev))
    L.check(st, "mg_colnorm")
    return out


def hess_solve(H, g, k, out=None):
    """y (k x T): back substitution in the first k x k of every column's
    rotated Hessenberg matrix H ((m+1) x m x T) against g ((m+1) x T)."""
    lib = L.load()
    dev = H.device
    m, T = H.shape[1], H.shape[2]
    if out is None:
        out = torch.empty((k, T), dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_hess_solve(_p(H), _p(g), m, k, T, _p(out), _stream(dev))
    L.check(st, "mg_hess_solve")
    return out


def givens(H, H0, cs, sn, g, hn, bnorm, j, est):
    """Column j of the batched GMRES Hessenberg matrices (H, H0: (m+1, m, T);
    cs, sn: (m, T); g: (m+1, T), complex128; hn, bnorm: (T,) float64): sets
    h_{j+1,j} = hn, copies the column to H0, applies the earlier Givens
    rotations and a new one, rotates g; est (1,) float64 receives
    max_t |g_{j+1,t}| / bnorm[t]."""
    lib = L.load()
    dev = H.device
    m, T = H.shape[1], H.shape[2]
    for t in (H, H0, cs, sn, g, hn, bnorm, est):
        if not t.is_contiguous():
            raise ValueError("givens: contiguous tensors expected")
    with torch.cuda.device(dev):
        st = lib.mg_givens(_p(H), _p(H0), _p(cs), _p(sn), _p(g), _p(hn), _p(bnorm), m, j, T, _p(est), _stream(dev))
    L.check(st, "mg_givens")
    return est


def hankel1(nu, x, exact=False):
    """H_nu^(1)(x); exact=False gives the reference's torch values (see bessel_ref.h)."""
    lib = L.load()
    dev = device_of(x)
    xd = x.to(device=dev, dtype=torch.float64).contiguous()
    out = torch.empty(xd.shape, dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_hankel1(int(nu), 1 if exact else 0, _p(xd), xd.numel(), _p(out), _stream(dev))
    L.check(st, "mg_hankel1")
    return out


def sources(space, beam, points, dirs, pol, k0, w):
    """Incident plane waves (beam=False) or Gaussian beams at points.

    dirs: (T, 2) of (cos(-theta), sin(-theta)) in 2D, (T, 3) directions in 3D;
    pol: None for scalar fields, else (T, d) polarisation components.
    Returns (M, T) or (M, d, T) complex128."""
    lib = L.load()
    dev = device_of(points)
    pts = _pts(points, dev)
    dirs = dirs.to(device=dev, dtype=torch.float64).contiguous()
    T = dirs.shape[0]
    d = 1 if pol is None else pol.shape[1]
    if pol is not None:
        pol = pol.to(device=dev, dtype=torch.float64).contiguous()
    shape = (pts.shape[0], T) if pol is None else (pts.shape[0], d, T)
    out = torch.empty(shape, dtype=torch.complex128, device=dev)
    with torch.cuda.device(dev):
        st = lib.mg_sources(space, 1 if beam else 0, _p(pts), pts.shape[0], _p(dirs), T, _p(pol), d, float(k0),
                            float(w), _p(out), _stream(dev))
    L.check(st, "mg_sources")
    return out
