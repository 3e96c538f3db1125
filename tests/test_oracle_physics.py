"""The oracle's solve / propagate / LDOS / sources against closed forms,
identities and independent libraries (numpy's OpenBLAS LAPACK, mpmath)."""
import cmath
import math

import mpmath as mp
import numpy as np
import pytest
import torch

from oracle import transmission as O
from magreete_amd import workloads as W


def textbook_g3(r, k):
    R = np.linalg.norm(r)
    kr = k * R
    a = 1 + 1j / kr - 1 / kr ** 2
    b = -1 - 3j / kr + 3 / kr ** 2
    return np.exp(1j * kr) / (4 * np.pi * R) * (a * np.eye(3) + b * np.outer(r, r) / R ** 2)


def small_case(kind, n, seed=0):
    rng = np.random.default_rng(seed)
    if kind.startswith("2d"):
        pts = W.unit_disk(n, rng)
    else:
        pts = W.unit_ball(n, rng)
    k0 = 1.0 / W.mean_nn_spacing(pts)
    alpha = W.alpha_3d(k0) if kind.startswith("3d") else (W.alpha_2d_tm(k0) if kind == "2d_tm" else W.alpha_2d_te(k0))
    return torch.from_numpy(pts), k0, alpha


def plane_rhs(cls, kind, pts, k0, T=4):
    if kind.startswith("2d"):
        th = 2 * np.pi * np.arange(T) / T
        return cls.generate_source(pts, k0, th, 0.2)
    u = torch.from_numpy(W.fibonacci_sphere(T))
    if kind == "3d_scalar":
        return cls.generate_source(pts, k0, u, 0.2)
    p1, _ = W.polarisation_pairs(u.numpy())
    return cls.generate_source(pts, k0, u, torch.from_numpy(p1), 0.2)


# ---------------------------------------------------------------------------
def test_solve_single_scatterer_closed_form():
    k0, radius = 6.0, 0.03
    alpha = W.alpha_3d(k0)
    s = O.Transmission3D_vector(torch.zeros((1, 3), dtype=torch.float64), source='plane')
    E0 = torch.tensor([[1.0 + 0.5j], [0.2 - 0.1j], [-0.3j]], dtype=torch.complex128)
    assert torch.allclose(s.solve(k0, alpha, radius, E0, self_interaction=False), E0, rtol=0, atol=0)
    S = O.self_interaction_integral_vector_3d(k0, radius)
    V = 4 * math.pi * radius ** 3 / 3
    c = k0 ** 2 * alpha * S / V
    got = s.solve(k0, alpha, radius, E0, self_interaction=True)
    # the reference forms c * torch.eye(n) with eye's default float32 dtype, so
    # the shift on the diagonal is c rounded to complex64
    ref = E0 / (1 - complex(np.complex64(c)))
    assert torch.allclose(got, ref, rtol=1e-14, atol=0)
    exact = E0 / (1 - c)
    dev = ((got - exact).abs() / exact.abs()).max().item()
    assert 1e-12 < dev < 1e-7
    # 2D TM
    t = O.Transmission2D_scalar(torch.zeros((1, 2), dtype=torch.float64), source='plane')
    a2 = W.alpha_2d_tm(k0)
    S2 = O.self_interaction_integral_scalar_2d(k0, radius)
    c2 = a2 * k0 * k0 * S2 / (math.pi * radius ** 2)
    ref2 = E0[:1] / (1 - complex(np.complex64(c2)))
    assert torch.allclose(t.solve(k0, a2, radius, E0[:1], self_interaction=True), ref2, rtol=1e-14, atol=0)


def test_solve_two_scatterers_by_hand():
    k0 = 5.0
    alpha = W.alpha_3d(k0)
    r1, r2 = np.array([0.1, -0.2, 0.05]), np.array([-0.15, 0.1, 0.2])
    s = O.Transmission3D_vector(torch.from_numpy(np.stack([r1, r2])), source='plane')
    rng = np.random.default_rng(1)
    E0 = rng.normal(size=(6, 3)) + 1j * rng.normal(size=(6, 3))
    c = alpha * k0 ** 2
    G = textbook_g3(r1 - r2, k0)
    # [[I, -cG], [-cG, I]] [E1; E2] = [b1; b2]  =>  (I - c^2 G G) E1 = b1 + c G b2
    b1, b2 = E0[:3], E0[3:]
    E1 = np.linalg.solve(np.eye(3) - c * c * G @ G, b1 + c * G @ b2)
    E2 = b2 + c * G @ E1
    got = s.solve(k0, alpha, 0.0, torch.from_numpy(E0), self_interaction=False).numpy()
    ref = np.concatenate([E1, E2])
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("kind", ["2d_tm", "2d_te", "3d_vector", "3d_scalar"])
def test_solve_vs_numpy_lapack(kind):
    pts, k0, alpha = small_case(kind, 40)
    cls = O.CLASSES[kind](pts, source='plane')
    E0 = plane_rhs(cls, kind, pts, k0)
    got = cls.solve(k0, alpha, 0.01, E0, self_interaction=True)
    # rebuild M independently in numpy: -alpha k0^2 G, unit diagonal, shift
    if kind == "2d_tm":
        G = cls.G0(cls.r, k0)
    elif kind == "2d_te":
        G = cls.G0_vector(None, k0)
    else:
        G = cls.G0(None, k0)
    G = G.numpy().copy()
    n = G.shape[0]
    M = -alpha * k0 * k0 * G
    np.fill_diagonal(M, 1.0)
    if kind == "2d_tm":
        c = alpha * k0 * k0 * O.self_interaction_integral_scalar_2d(k0, 0.01) / (math.pi * 0.01 ** 2)
    elif kind == "2d_te":
        c = alpha * k0 * k0 * O.self_interaction_integral_vector_2d(k0, 0.01) / (math.pi * 0.01 ** 2)
    elif kind == "3d_vector":
        c = k0 ** 2 * alpha * O.self_interaction_integral_vector_3d(k0, 0.01) / (4 * math.pi * 0.01 ** 3 / 3)
    else:
        c = k0 ** 2 * alpha * O.self_interaction_integral_scalar_3d(k0, 0.01) / (4 * math.pi * 0.01 ** 3 / 3)
    M[np.diag_indices(n)] -= complex(np.complex64(c))   # c * float32 eye, as the reference
    ref = np.linalg.solve(M, E0.reshape(n, -1).numpy())
    assert np.max(np.abs(got.numpy() - ref)) / np.max(np.abs(ref)) < 1e-11


def test_solve_vs_mpmath_lu():
    pts, k0, alpha = small_case("3d_vector", 4, seed=2)
    cls = O.Transmission3D_vector(pts, source='plane')
    E0 = plane_rhs(cls, "3d_vector", pts, k0, T=2)[..., :1]
    got = cls.solve(k0, alpha, 0.0, E0, self_interaction=False).numpy()[:, 0]
    G = cls.G0(None, k0).numpy()
    n = G.shape[0]
    mp.mp.dps = 30
    M = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            M[i, j] = 1 if i == j else -alpha * k0 * k0 * complex(G[i, j])
    b = mp.matrix([complex(v) for v in E0.reshape(n).numpy()])
    x = mp.lu_solve(M, b)
    ref = np.array([complex(v) for v in x])
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12


def test_propagate_single_scatterer_closed_form():
    k0 = 4.0
    alpha = W.alpha_3d(k0)
    s = O.Transmission3D_vector(torch.zeros((1, 3), dtype=torch.float64), source='plane')
    Ek = torch.tensor([[1.0 + 0.2j], [0.3], [-0.5j]], dtype=torch.complex128)
    pts = torch.tensor([[1.0, 0.5, -0.3], [0.0, 2.0, 0.0]], dtype=torch.float64)
    E0 = torch.zeros((2, 3, 1), dtype=torch.complex128)
    got = s.propagate(pts, Ek, k0, alpha, E0)
    for m in range(2):
        ref = alpha * k0 ** 2 * textbook_g3(pts[m].numpy(), k0) @ Ek.numpy()[:, 0]
        assert np.max(np.abs(got[m, :, 0].numpy() - ref)) < 1e-13 * np.max(np.abs(ref))


def test_propagate_linear_in_columns():
    pts, k0, alpha = small_case("3d_vector", 30)
    cls = O.Transmission3D_vector(pts, source='plane')
    meas = torch.from_numpy(2.0 * W.fibonacci_sphere(16))
    rng = np.random.default_rng(0)
    Ek = torch.from_numpy(rng.normal(size=(90, 3)) + 1j * rng.normal(size=(90, 3)))
    E0 = torch.zeros((16, 3, 3), dtype=torch.complex128)
    full = cls.propagate(meas, Ek, k0, alpha, E0)
    comb = cls.propagate(meas, Ek[:, :1] * 2 - 3j * Ek[:, 1:2], k0, alpha, E0[..., :1])
    assert torch.allclose(comb[..., 0], 2 * full[..., 0] - 3j * full[..., 1], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("kind", ["2d_tm", "2d_te", "3d_vector", "3d_scalar"])
def test_ldos_site_identity(kind):
    """self_interaction=False, points = scatterers: G_sites = (I - M)/(k0^2 alpha),
    so diag(G W G^T) = (W_rr - 1)/(k0^2 alpha)^2 (M complex symmetric)."""
    pts, k0, alpha = small_case(kind, 25)
    cls = O.CLASSES[kind](pts, source='plane')
    got = cls.LDOS_measurements(pts, k0, alpha, 0.01, self_interaction=False)
    W_ = cls._inverse(k0, alpha, 0.01, False, "Rayleigh").numpy()
    c = k0 * k0 * alpha
    pref = {"2d_tm": 4 * k0 * k0, "2d_te": 4 * k0 * k0, "3d_vector": 2 * np.pi * k0,
            "3d_scalar": 4 * np.pi * k0}[kind]
    rows = np.imag(pref * alpha * (np.diag(W_) - 1) / c ** 2)
    d = cls.d
    ref = rows.reshape(-1, d).sum(1)
    assert np.max(np.abs(got.numpy().reshape(-1) - ref)) < 1e-10 * np.max(np.abs(ref))


def test_ldos_self_block_quirk_single_scatterer():
    """One scatterer probed at its own position with self-interaction: every
    entry of the coincident 3x3 block becomes S/V, so
    LDOS = Im(2 pi k0 alpha d^2 (S/V)^2 / (1 - k0^2 alpha S/V))."""
    k0, radius = 5.0, 0.02
    alpha = W.alpha_3d(k0)
    pt = torch.zeros((1, 3), dtype=torch.float64)
    s = O.Transmission3D_vector(pt, source='plane')
    got = s.LDOS_measurements(pt, k0, alpha, radius, self_interaction=True).item()
    sv = O.self_interaction_integral_vector_3d(k0, radius) / (4 * np.pi * radius ** 3 / 3)
    w = 1 / (1 - complex(np.complex64(k0 ** 2 * alpha * sv)))   # shift rounded as the reference's
    ref = np.imag(2 * np.pi * k0 * alpha * 9 * sv * sv * w)
    # Im() of a value ~2e4 times larger: compare at that scale; d^2 = 9 (not d = 3)
    assert abs(got - ref) < 1e-9 * abs(ref)


@pytest.mark.parametrize("kind", ["2d_tm", "2d_te", "3d_vector", "3d_scalar"])
def test_mean_dos_is_mean_ldos_and_batches(kind):
    pts, k0, alpha = small_case(kind, 20)
    cls = O.CLASSES[kind](pts, source='plane')
    rng = np.random.default_rng(9)
    probes = torch.from_numpy(0.5 * (W.unit_disk(7, rng) if kind.startswith("2d") else W.unit_ball(7, rng)))
    probes = torch.cat([probes, pts[:3]])
    ld = cls.LDOS_measurements(probes, k0, alpha, 0.01)
    md = cls.mean_DOS_measurements(probes, k0, alpha, 0.01)
    # both are Im() of sums much larger than the result: agree to ~1e-9 of max|LDOS|
    assert abs(md.item() - ld.mean().item()) < 1e-7 * abs(ld).max().item()
    bl = cls.batch_LDOS_measurements(torch.split(probes, 4), k0, alpha, 0.01)
    assert torch.allclose(bl, ld, rtol=1e-12, atol=0)


def test_sources_closed_forms(golden):
    for case in golden["sources"]["2d"]:
        pts = torch.tensor([case["x"]], dtype=torch.float64)
        for src in ("plane", "beam"):
            tm = O.Transmission2D_scalar(pts, source=src)
            te = O.Transmission2D_vector(pts, source=src)
            e = tm.generate_source(pts, case["k0"], np.array([case["theta"]]), case["w"])[0, 0].item()
            ref = complex(*case[src])
            assert abs(e - ref) < 1e-13
            ev = te.generate_source(pts, case["k0"], np.array([case["theta"]]), case["w"])[0, :, 0].numpy()
            th = case["theta"]
            # TE polarisation (-sin, cos), held by the reference in a float32 tensor
            u = np.array([-np.sin(th), np.cos(th)]).astype(np.float32).astype(np.float64)
            assert np.allclose(ev, ref * u, rtol=0, atol=1e-13)
    for case in golden["sources"]["3d"]:
        pts = torch.tensor([case["x"]], dtype=torch.float64)
        u = torch.tensor([case["u"]], dtype=torch.float64)
        p1, _ = W.polarisation_pairs(u.numpy())
        for src in ("plane", "beam"):
            sc = O.Transmission3D_scalar(pts, source=src)
            vc = O.Transmission3D_vector(pts, source=src)
            ref = complex(*case[src])
            assert abs(sc.generate_source(pts, case["k0"], u, case["w"])[0, 0].item() - ref) < 1e-13
            ev = vc.generate_source(pts, case["k0"], u, torch.from_numpy(p1), case["w"])[0, :, 0].numpy()
            assert np.allclose(ev, ref * p1[0], rtol=0, atol=1e-13)
