"""Green's functions: oracle (torch restatement of the reference) and the
kernels' green.h (host build) against independently written textbook forms
(tests/golden), symmetry, small-R limits, NaN placement and the regularise
quirks of each kind."""
import ctypes
import math

import numpy as np
import pytest
import torch

from oracle import transmission as O

KIND = {"2d_tm": 0, "2d_te": 1, "3d_vector": 2, "3d_scalar": 3}
DIM = {"2d_tm": 1, "2d_te": 2, "3d_vector": 3, "3d_scalar": 1}
EXACT = 128
REG = 1


def host_green(lib, kind, r, k0, flags=0, radius=0.0):
    d = DIM[kind]
    rr = np.zeros(3)
    rr[:len(r)] = r
    g = np.zeros(2 * d * d)
    lib.shim_green(KIND[kind], rr.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(k0), flags,
                   ctypes.c_double(radius), g.ctypes.data_as(ctypes.c_void_p))
    return (g[0::2] + 1j * g[1::2]).reshape(d, d)


def oracle_green(kind, r, k0, regularize=False, radius=0.0):
    t = torch.tensor(r, dtype=torch.float64).reshape(1, 1, -1)
    f = {"2d_tm": O.greens_2d_tm, "2d_te": O.greens_2d_te, "3d_vector": O.greens_3d_vector,
         "3d_scalar": O.greens_3d_scalar}[kind]
    return f(t, k0, regularize=regularize, radius=radius).reshape(DIM[kind], DIM[kind]).numpy()


def golden_value(kind, case):
    g = np.array(case["G"])
    if kind in ("2d_tm", "3d_scalar"):
        return np.array([[g[0] + 1j * g[1]]])
    return g[..., 0] + 1j * g[..., 1]


def _kr(case):
    return case["k0"] * math.sqrt(sum(v * v for v in case["r"]))


@pytest.mark.parametrize("kind", list(KIND))
def test_oracle_vs_textbook(golden, kind):
    for case in golden["green"][kind]:
        ref = golden_value(kind, case)
        got = oracle_green(kind, case["r"], case["k0"])
        # 2D values carry torch's Bessel deviation (up to ~3e-6 for kR in [1.5, 30])
        tol = 3e-6 if kind.startswith("2d") and 1.5 <= _kr(case) <= 30 else 1e-13
        assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref)), (kind, case["r"], case["k0"])


@pytest.mark.parametrize("kind", list(KIND))
def test_kernel_exact_vs_textbook(golden, host_kernels, kind):
    for case in golden["green"][kind]:
        ref = golden_value(kind, case)
        got = host_green(host_kernels, kind, case["r"], case["k0"], EXACT)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), (kind, case["r"], case["k0"])


@pytest.mark.parametrize("kind", list(KIND))
def test_kernel_reference_mode_vs_oracle(kind, host_kernels):
    rng = np.random.default_rng(3)
    sp = 2 if kind.startswith("2d") else 3
    for _ in range(200):
        r = rng.normal(size=sp) * rng.choice([1e-3, 0.05, 0.3, 1.0])
        k0 = rng.uniform(1.0, 40.0)
        got = host_green(host_kernels, kind, r, k0)
        ref = oracle_green(kind, r, k0)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("kind", list(KIND))
def test_symmetry(kind, host_kernels):
    rng = np.random.default_rng(5)
    sp = 2 if kind.startswith("2d") else 3
    for _ in range(50):
        r = rng.normal(size=sp)
        k0 = rng.uniform(1, 20)
        g = oracle_green(kind, r, k0)
        assert np.array_equal(g, oracle_green(kind, -r, k0))
        assert np.array_equal(g, g.T)
        h = host_green(host_kernels, kind, r, k0)
        assert np.array_equal(h, host_green(host_kernels, kind, -r, k0))
        assert np.array_equal(h, h.T)


def test_small_r_limits(host_kernels):
    """Im G -> k/(6 pi) I (3D vector), k/(4 pi) (3D scalar), 1/4 (TM), I/8 (TE)."""
    k0 = 7.0
    for kind, lim in [("3d_vector", k0 / (6 * math.pi) * np.eye(3)), ("3d_scalar", np.array([[k0 / (4 * math.pi)]])),
                      ("2d_tm", np.array([[0.25]])), ("2d_te", np.eye(2) / 8)]:
        r = np.array([1e-4, -2e-4, 1.5e-4][: (2 if kind.startswith("2d") else 3)])
        g = host_green(host_kernels, kind, r, k0, EXACT)
        tol = 1e-4 if kind == "3d_vector" else 1e-6
        assert np.max(np.abs(g.imag - lim)) < tol * np.max(np.abs(lim)) + 1e-9, kind
        go = oracle_green(kind, r, k0)
        assert np.max(np.abs(go.imag - lim)) < tol * np.max(np.abs(lim)) + 1e-9, kind


def test_far_field_transverse(host_kernels):
    # 3D dyadic far field: G R̂ -> O(1/(kR)^2) relative to |G|
    r = np.array([30.0, -40.0, 20.0])
    g = host_green(host_kernels, "3d_vector", r, 5.0, EXACT)
    rh = r / np.linalg.norm(r)
    assert np.linalg.norm(g @ rh) < 2e-2 * np.linalg.norm(g)


@pytest.mark.parametrize("kind", list(KIND))
def test_coincident_is_nan(kind, host_kernels):
    sp = 2 if kind.startswith("2d") else 3
    g = host_green(host_kernels, kind, np.zeros(sp), 3.0)
    assert np.all(np.isnan(g.real) | np.isnan(g.imag))
    go = oracle_green(kind, np.zeros(sp), 3.0)
    assert np.all(np.isnan(go))


def test_regularize_quirks(host_kernels):
    # TM and 3D compare R with radius; TE compares the scaled k0 R
    k0, rad = 10.0, 0.05
    r = np.array([0.02, 0.0])
    for kind in ("2d_tm", "2d_te"):
        go = oracle_green(kind, r, k0, regularize=True, radius=rad)
        gk = host_green(host_kernels, kind, r, k0, REG, rad)
        assert np.isnan(go).all() == np.isnan(gk).all()
    # TE: k0 R = 0.2 > radius -> finite; TM: R = 0.02 < radius -> NaN
    assert np.isnan(oracle_green("2d_tm", r, k0, True, rad)).all()
    assert not np.isnan(oracle_green("2d_te", r, k0, True, rad)).any()
    assert not np.isnan(host_green(host_kernels, "2d_te", r, k0, REG, rad)).any()
    r3 = np.array([0.02, 0.0, 0.0])
    assert np.isnan(oracle_green("3d_vector", r3, k0, True, rad)).all()
    assert np.isnan(host_green(host_kernels, "3d_vector", r3, k0, REG, rad)).all()
