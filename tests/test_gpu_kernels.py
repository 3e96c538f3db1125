"""HIP kernels through the C ABI against the oracle / CPU references."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from magreete_amd import _lib as L
from magreete_amd import engine as E
from magreete_amd import workloads as W
from oracle import transmission as O

DEV = "cuda:0"


def say(*a):
    print(*a, flush=True)


def crand(shape, rng, dtype=np.complex128):
    return (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(dtype)


@pytest.mark.parametrize("prec", ["c128", "c64", "bf16x3", "bf16x6"])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (17, 33, 9), (64, 64, 8), (130, 70, 257), (300, 200, 100), (5, 700, 3),
                                 (4200, 4300, 70)])
def test_cgemm(prec, mnk):
    m, n, k = mnk
    say("cgemm", prec, mnk)
    rng = np.random.default_rng(sum(mnk))
    dt = np.complex128 if prec == "c128" else np.complex64
    A, B, C = crand((m, k), rng, dt), crand((k, n), rng, dt), crand((m, n), rng, dt)
    alpha, beta = 0.7 - 0.3j, -1.1 + 0.4j
    ref = alpha * (A.astype(np.complex128) @ B.astype(np.complex128)) + beta * C.astype(np.complex128)
    Cd = torch.from_numpy(C).to(DEV)
    E.cgemm(torch.from_numpy(A).to(DEV), torch.from_numpy(B).to(DEV), Cd, alpha, beta, prec)
    got = Cd.cpu().numpy().astype(np.complex128)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    # BF16x3 keeps ~16 significant bits per product, BF16x6 ~24 (FP32-level)
    assert err < {"c128": 1e-14, "c64": 2e-6, "bf16x3": 5e-5, "bf16x6": 2e-6}[prec], err


@pytest.mark.parametrize("mnk", [(4200, 4300, 96), (4099, 4113, 64), (4099, 4113, 48), (4200, 4300, 32),
                                 (12288, 256, 512), (2100, 4113, 1024)])
@pytest.mark.parametrize("ld_pad", [0, 24])
def test_cgemm_fp64_dma_shapes(mnk, ld_pad):
    """Shapes that take the FP64 kernel with DMA-staged operands (tiles for
    every CU, depth a multiple of 16), with ragged edges and, with ld_pad,
    operands and C as device views whose leading dimension exceeds their
    columns."""
    m, n, k = mnk
    say("cgemm dma", mnk, ld_pad)
    rng = np.random.default_rng(sum(mnk) + ld_pad)
    Af = torch.from_numpy(crand((m, k + ld_pad), rng)).to(DEV)
    Bf = torch.from_numpy(crand((k, n + ld_pad), rng)).to(DEV)
    Cf = torch.from_numpy(crand((m, n + ld_pad), rng)).to(DEV)
    o = ld_pad // 2
    Ad, Bd, Cd = Af[:, o:o + k], Bf[:, :n], Cf[:, ld_pad:]
    A, B, C = Ad.cpu().numpy(), Bd.cpu().numpy(), Cd.cpu().numpy()
    outside = Cf[:, :ld_pad].clone()
    alpha, beta = 0.7 - 0.3j, -1.1 + 0.4j
    ref = alpha * (A @ B) + beta * C
    E.cgemm(Ad, Bd, Cd, alpha, beta, "c128")
    got = Cd.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-14, err
    assert torch.equal(Cf[:, :ld_pad], outside)   # nothing written outside the view


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("mnk", [(3000, 40, 4096), (2000, 70, 1500), (12288 // 4, 256, 3072)])
def test_cgemm_split_k(prec, mnk):
    """Split-K tall-skinny products: partials over K in a workspace, summed in
    a fixed order; the same result as the unsplit GEMM to rounding, and
    repeatable bit for bit."""
    m, n, k = mnk
    say("cgemm split-K", prec, mnk)
    lib = L.load()
    elems = lib.mg_cgemm_work_elems(E.PRECS[prec], m, n, k)
    assert elems >= 2 * m * n                     # these shapes split
    assert lib.mg_cgemm_work_elems(E.PRECS[prec], 128 * 40, 64 * 40, k) == 0
    rng = np.random.default_rng(m + n + k)
    dt = np.complex128 if prec == "c128" else np.complex64
    # A as a strided view (lda > k), C with ldc > n
    Abig = torch.from_numpy(crand((m, k + 3), rng, dt)).to(DEV)
    A = Abig[:, 1:k + 1]
    B = torch.from_numpy(crand((k, n), rng, dt)).to(DEV)
    Cbig = torch.from_numpy(crand((m, n + 5), rng, dt)).to(DEV)
    alpha, beta = -1.0 + 0.25j, 0.5 - 1.0j
    C1 = Cbig.clone()
    E.cgemm(A, B, C1[:, 2:n + 2], alpha, beta, prec, split=True)
    C2 = Cbig.clone()
    E.cgemm(A, B, C2[:, 2:n + 2], alpha, beta, prec, split=False)
    C3 = Cbig.clone()
    E.cgemm(A, B, C3[:, 2:n + 2], alpha, beta, prec, split=True)
    assert torch.equal(C1, C3)                    # deterministic
    assert torch.equal(C1[:, :2], Cbig[:, :2]) and torch.equal(C1[:, n + 2:], Cbig[:, n + 2:])
    ref = alpha * (A.cpu().numpy().astype(np.complex128) @ B.cpu().numpy().astype(np.complex128)) \
        + beta * Cbig[:, 2:n + 2].cpu().numpy().astype(np.complex128)
    # FP32 accumulation over K terms: rounding grows like sqrt(K)
    tol = 1e-14 if prec == "c128" else 1e-7 * np.sqrt(k) * 4
    for C in (C1, C2):
        got = C[:, 2:n + 2].cpu().numpy().astype(np.complex128)
        assert np.abs(got - ref).max() / np.abs(ref).max() < tol


def test_cgemm_strided_views():
    say("cgemm strided")
    rng = np.random.default_rng(7)
    big = torch.from_numpy(crand((200, 300), rng)).to(DEV)
    A = big[10:90, 20:60]
    B = big[100:140, 5:105]
    C = torch.zeros((80, 100), dtype=torch.complex128, device=DEV)
    E.cgemm(A, B, C)
    ref = A.cpu().numpy() @ B.cpu().numpy()
    assert np.abs(C.cpu().numpy() - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("kind", ["2d_tm", "2d_te", "3d_vector", "3d_scalar"])
@pytest.mark.parametrize("exact", [False, True])
def test_assemble_vs_oracle(kind, exact):
    say("assemble", kind, exact)
    rng = np.random.default_rng(1)
    sp = 2 if kind.startswith("2d") else 3
    rows = rng.uniform(-1, 1, size=(37, sp))
    cols = rng.uniform(-1, 1, size=(53, sp))
    cols[5] = rows[3]  # one coincident pair -> NaN block
    k0 = 11.3
    rt, ct = torch.from_numpy(rows), torch.from_numpy(cols)
    flags = L.MG_F_BESSEL_EXACT if exact else 0
    G, G64 = E.assemble(kind, rt.to(DEV), ct.to(DEV), k0, flags=flags, want64=True)
    cls = O.CLASSES[kind](ct, source='plane')
    if kind == "2d_tm":
        ref = cls.G0(rt, k0)
    elif kind == "2d_te":
        ref = cls.G0_vector(rt, k0)
    else:
        ref = cls.G0(rt, k0)
    got = G.cpu()
    assert torch.equal(torch.isnan(got), torch.isnan(ref))
    m = ~torch.isnan(ref)
    diff = (got[m] - ref[m]).abs()
    scale = ref[m].abs()
    tol = 3e-6 if (exact and kind.startswith("2d")) else 1e-12   # exact mode departs from torch's Bessel values
    assert (diff <= tol * scale + 1e-15).all(), diff.max()
    g64 = G64.cpu().to(torch.complex128)
    assert ((g64[m] - got[m]).abs() <= 1e-6 * scale + 1e-30).all()


@pytest.mark.parametrize("kind", ["2d_tm", "2d_te", "3d_vector", "3d_scalar"])
def test_system_matrix_vs_oracle(kind):
    """M = -alpha k0^2 G0(None), unit diagonal, self-interaction shift (drop-in _system)."""
    say("system", kind)
    from magreete_amd import Transmission2D as T2, Transmission3D as T3
    rng = np.random.default_rng(2)
    sp = 2 if kind.startswith("2d") else 3
    pts = torch.from_numpy(rng.uniform(-1, 1, size=(40, sp)))
    k0, radius = 9.0, 0.02
    alpha = W.alpha_3d(k0) if kind.startswith("3d") else W.alpha_2d_tm(k0)
    mod = {"2d_tm": T2.Transmission2D_scalar, "2d_te": T2.Transmission2D_vector,
           "3d_vector": T3.Transmission3D_vector, "3d_scalar": T3.Transmission3D_scalar}[kind]
    dut = mod(pts.to(DEV), source='plane')
    ora = O.CLASSES[kind](pts, source='plane')
    for si in (False, True):
        M, _ = dut._system(k0, alpha, radius, si, "Rayleigh")
        if kind == "2d_tm":
            G = ora.G0(ora.r, k0)
        elif kind == "2d_te":
            G = ora.G0_vector(None, k0)
        else:
            G = ora.G0(None, k0)
        S = ora.__class__  # noqa
        ref = O._system_matrix(G.clone(), alpha, k0,
                               {"2d_tm": O.self_interaction_integral_scalar_2d,
                                "2d_te": O.self_interaction_integral_vector_2d,
                                "3d_vector": O.self_interaction_integral_vector_3d,
                                "3d_scalar": O.self_interaction_integral_scalar_3d}[kind](k0, radius) if si else 0,
                               (np.pi * radius ** 2) if kind.startswith("2d") else 4 * np.pi * radius ** 3 / 3,
                               si, order2d=kind.startswith("2d"))
        got = M.cpu()
        assert ((got - ref).abs() <= 1e-12 * ref.abs() + 1e-15).all()


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("n", [1, 7, 33, 100, 257, 600, 1100, 4100])
def test_getrf_getrs(prec, n):
    say("lu", prec, n)
    rng = np.random.default_rng(n)
    dt = np.complex128 if prec == "c128" else np.complex64
    # spectrum in a disk of radius ~1 around 2: condition number ~3 at every n
    A = crand((n, n), rng) / np.sqrt(2 * n) + 2.0 * np.eye(n)
    B = crand((n, 5), rng)
    Ad = torch.from_numpy(A.astype(dt)).to(DEV)
    Bd = torch.from_numpy(B.astype(dt)).to(DEV)
    ipiv, info = E.getrf(Ad, prec)
    E.getrs(Ad, ipiv, Bd, prec)
    E.check_barriers(Ad.device)
    assert info.item() == 0
    x = Bd.cpu().numpy().astype(np.complex128)
    ref = np.linalg.solve(A, B)
    err = np.linalg.norm(x - ref, axis=0) / np.linalg.norm(ref, axis=0)
    assert err.max() < (1e-12 if prec == "c128" else 1e-4), err.max()
    # pivots are a valid LAPACK sequence
    p = ipiv.cpu().numpy()
    assert np.all(p >= np.arange(n)) and np.all(p < n)


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("n", [5, 255, 256, 257, 700, 1100])
def test_factor_dinv_solve(prec, n):
    """Solves through stored diagonal-block inverses (GEMM-only), repeated on one factor."""
    say("factor dinv", prec, n)
    rng = np.random.default_rng(n + 17)
    dt = np.complex128 if prec == "c128" else np.complex64
    A = crand((n, n), rng) / np.sqrt(2 * n) + 2.0 * np.eye(n)
    F = E.Factor(torch.from_numpy(A.astype(dt)).to(DEV), prec)
    for T in (1, 37):
        B = crand((n, T), rng)
        Bd = torch.from_numpy(B.astype(dt)).to(DEV)
        F.solve(Bd)
        x = Bd.cpu().numpy().astype(np.complex128)
        ref = np.linalg.solve(A, B)
        err = (np.linalg.norm(x - ref, axis=0) / np.linalg.norm(ref, axis=0)).max()
        assert err < (1e-13 if prec == "c128" else 1e-5), (T, err)
    E.check_barriers(F.A.device)


def test_getrf_matches_lapack_pivots():
    say("lu pivots")
    rng = np.random.default_rng(11)
    n = 300
    A = crand((n, n), rng)
    Ad = torch.from_numpy(A).to(DEV)
    ipiv, info = E.getrf(Ad)
    import scipy.linalg
    lu, piv = scipy.linalg.lu_factor(A)
    assert np.array_equal(ipiv.cpu().numpy(), piv)
    assert np.abs(Ad.cpu().numpy() - lu).max() < 1e-11 * np.abs(lu).max()


def test_getrf_singular_info():
    say("lu singular")
    rng = np.random.default_rng(3)
    n = 70
    A = crand((n, n), rng)
    A[:, 40] = 0.0
    Ad = torch.from_numpy(A).to(DEV)
    ipiv, info = E.getrf(Ad)
    E.check_barriers(Ad.device)
    assert info.item() == 41


@pytest.mark.parametrize("kind", ["2d_tm", "2d_te", "3d_vector", "3d_scalar"])
@pytest.mark.parametrize("nscat", [45, 1601])
def test_propagate_kernel_vs_oracle(kind, nscat):
    """Fused propagation against the oracle's G @ Ek; with 1601 scatterers the
    scatterers are split over a workspace, whose parts must add up to the
    unsplit launch."""
    say("propagate", kind, nscat)
    rng = np.random.default_rng(4)
    sp = 2 if kind.startswith("2d") else 3
    d = {"2d_tm": 1, "2d_te": 2, "3d_vector": 3, "3d_scalar": 1}[kind]
    scat = torch.from_numpy(rng.uniform(-1, 1, size=(nscat, sp)))
    meas = torch.from_numpy(rng.uniform(-2, 2, size=(29, sp)))
    k0 = 7.5
    T = 70
    Ek = torch.from_numpy(crand((d * nscat, T), rng))
    E0 = torch.from_numpy(crand((d * 29, T), rng))
    alpha = 0.3 - 0.1j
    out = E0.clone().to(DEV)
    E.propagate(kind, meas.to(DEV), scat.to(DEV), k0, alpha, Ek.to(DEV), out)
    lib = L.load()
    nb = lib.mg_propagate_work_bytes(E.KINDS[kind], 29, nscat, T)
    assert (nb > 0) == (nscat > 1000), nb
    whole = E0.clone().to(DEV)
    md, sd, ekd = meas.to(DEV).contiguous(), scat.to(DEV).contiguous(), Ek.to(DEV)
    st = lib.mg_propagate(E.KINDS[kind], md.data_ptr(), 29, sd.data_ptr(), nscat, k0, alpha.real, alpha.imag, 0, 0.0,
                          ekd.data_ptr(), T, T, 1.0, 0.0, whole.data_ptr(), T,
                          torch.cuda.current_stream().cuda_stream)
    assert st == 0
    assert ((out - whole).abs().max() / whole.abs().max()).item() < 1e-14
    cls = O.CLASSES[kind](scat, source='plane')
    if kind == "2d_te":
        G = cls.G0_vector(meas, k0)
    else:
        G = cls.G0(meas, k0)
    ref = alpha * (G @ Ek) + E0
    err = (out.cpu() - ref).abs().max() / ref.abs().max()
    assert err < 1e-13, err


@pytest.mark.parametrize("k,n,T", [(1, 1, 1), (3, 100, 70), (9, 777, 256), (13, 2000, 65)])
def test_krylov_mdot_maxpy(k, n, T):
    """Per-column Gram-Schmidt passes against numpy; the dot products repeat bit for bit."""
    say("krylov", k, n, T)
    rng = np.random.default_rng(k * n + T)
    V = crand((k + 1, n, T), rng)
    w = crand((n, T), rng)
    Vd = torch.from_numpy(V).to(DEV)[:k]              # a view with the stack's strides
    wd = torch.from_numpy(w).to(DEV)
    h = E.mdot(Vd, wd)
    ref = np.einsum("int,nt->it", V[:k].conj(), w)
    assert np.abs(h.cpu().numpy() - ref).max() < 1e-12 * np.abs(ref).max() + 1e-13
    assert torch.equal(h, E.mdot(Vd, wd))
    w2 = wd.clone()
    E.maxpy(Vd, h, wd, alpha=-0.5)
    ref_w = w - 0.5 * np.einsum("int,it->nt", V[:k], ref)
    assert np.abs(wd.cpu().numpy() - ref_w).max() < 1e-12 * np.abs(ref_w).max()
    # the fused form: the same update bit for bit, and the column norms
    norms = torch.empty(T, dtype=torch.float64, device=DEV)
    E.maxpy_norm(Vd, h, w2, norms, alpha=-0.5)
    assert torch.equal(w2, wd)
    ref_n = np.linalg.norm(ref_w, axis=0)
    assert np.abs(norms.cpu().numpy() - ref_n).max() < 1e-13 * ref_n.max()
    # read-only column norms, the same fixed summation order
    assert torch.equal(E.colnorm(w2), norms)
    # the fused update-and-dots pass: the update of maxpy (alpha -1) bit for bit, the dots of mdot
    w3 = torch.from_numpy(w).to(DEV)
    w4 = w3.clone()
    h3 = E.maxpy_mdot(Vd, h, w3)
    E.maxpy(Vd, h, w4, alpha=-1.0)
    assert torch.equal(w3, w4)
    assert torch.equal(h3, E.mdot(Vd, w4))
