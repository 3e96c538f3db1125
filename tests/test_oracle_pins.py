"""Pins of the oracle functions that the other CPU tests leave open, each
against what the mathematics fixes rather than a restatement of the same
formula:

* self-interaction integrals (Transmission2D.py:12-36, Transmission3D.py:13-37):
  Im(S)/V equals the R -> 0 limit of Im G; the "full" forms are the volume
  integrals of the Green's functions over the particle (scalar: mpmath
  quadrature of the Green's function; vector: the depolarisation identity
  int_ball G = (1 - 1/d) int_ball g I - I / (d k^2) of G = (I + grad grad / k^2) g);
  the "Rayleigh" forms are their small-k0 a series;
* propagate_ss (Transmission3D.py:208-238 and siblings): the single-scatterer
  Born closed form and the coincident-point fix-up, which indexes E0_meas by
  the scatterer's index (a reference quirk, kept);
* the periodic wrap of the Green's functions (Transmission2D.py:74-81,
  Transmission3D.py:58-69) and the kernels' MG_F_PERIODIC_* flags: a shift by
  a unit lattice vector along a periodic axis leaves G unchanged;
* the regularised NaN fix-up with several scatterers inside the radius
  (Transmission3D.py:166-177, Transmission2D.py:193-203): brute force on tiny
  inputs, including the 2D TE fix-up that copies a flat row of Ek;
* the 3D 'point' sources (deliberate deviation, the reference's branch fails):
  the dipole field sqrt(I) G(r - r_s) p against the textbook dyadic.
"""
import ctypes
import math

import mpmath as mp
import numpy as np
import pytest
import torch

from oracle import transmission as O
from magreete_amd import Transmission2D as D2, Transmission3D as D3

SI = {  # kind -> (oracle S function, volume of the particle, Im G(0))
    "2d_tm": (O.self_interaction_integral_scalar_2d, lambda a: math.pi * a * a, lambda k: 0.25),
    "2d_te": (O.self_interaction_integral_vector_2d, lambda a: math.pi * a * a, lambda k: 0.125),
    "3d_vector": (O.self_interaction_integral_vector_3d, lambda a: 4 * math.pi * a ** 3 / 3,
                  lambda k: k / (6 * math.pi)),
    "3d_scalar": (O.self_interaction_integral_scalar_3d, lambda a: 4 * math.pi * a ** 3 / 3,
                  lambda k: k / (4 * math.pi)),
}
GREEN = {"2d_tm": O.greens_2d_tm, "2d_te": O.greens_2d_te, "3d_vector": O.greens_3d_vector,
         "3d_scalar": O.greens_3d_scalar}
DIM = {"2d_tm": 1, "2d_te": 2, "3d_vector": 3, "3d_scalar": 1}


def _g(kind, r, k0, **kw):
    t = torch.tensor(r, dtype=torch.float64).reshape(1, 1, -1)
    d = DIM[kind]
    return GREEN[kind](t, k0, **kw).reshape(d, d).numpy()


# ---------------------------------------------------------------------------
# self-interaction integrals
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", list(SI))
def test_rayleigh_imag_is_green_limit(kind):
    """Im S_Rayleigh / V = lim_{R->0} Im G (1/4, 1/8, k/6pi, k/4pi), the
    radiative damping of a point scatterer; the limit read off the oracle's
    own Green's function at k0 R = 1e-3 along a skew axis (the next terms
    of Im G are O((k0 R)^2); the 3D dyadic's real part, ~(k0 R)^-3 larger,
    costs ~1e-7 of cancellation)."""
    f, vol, img0 = SI[kind]
    for k0 in (0.7, 5.0, 31.0):
        for a in (1e-3, 0.02):
            S = f(k0, a, "Rayleigh")
            assert abs(S.imag / vol(a) - img0(k0)) <= 1e-13 * img0(k0), (kind, k0, a)
        r = 1e-3 / k0 * np.array([0.6, 0.8, 0.0][:2] if kind.startswith("2d") else [0.48, 0.6, 0.64])
        G = _g(kind, r, k0)
        assert np.allclose(np.imag(G), img0(k0) * np.eye(DIM[kind]), rtol=0, atol=1e-5 * img0(k0)), \
            (kind, np.imag(G))


def _quad_disk_tm(k0, a):
    """int_{|x|<a} (i/4) H0(k0 |x|) dx by mpmath quadrature (log singularity at 0)."""
    mp.mp.dps = 30
    f = lambda r: 0.25j * mp.hankel1(0, k0 * r) * 2 * mp.pi * r
    return complex(mp.quad(f, [0, a / 4, a]))


def _quad_ball_scalar(k0, a):
    """int_{|x|<a} exp(i k0 |x|) / (4 pi |x|) dx by mpmath quadrature."""
    mp.mp.dps = 30
    return complex(mp.quad(lambda r: r * mp.exp(1j * k0 * r), [0, a]))


@pytest.mark.parametrize("k0,a", [(1.0, 0.3), (7.5, 0.05), (30.0, 0.011), (2.0, 1.2)])
def test_full_scalar_is_volume_integral(k0, a):
    """The "full" scalar forms are the Green's function integrated over the
    disk / ball (quadrature with mpmath's Hankel and exponential, not the
    oracle's scipy call nor its closed form)."""
    s2 = O.self_interaction_integral_scalar_2d(k0, a, "full")
    q2 = _quad_disk_tm(k0, a)
    assert abs(s2 - q2) <= 1e-12 * abs(q2)
    s3 = O.self_interaction_integral_scalar_3d(k0, a, "full")
    q3 = _quad_ball_scalar(k0, a)
    assert abs(s3 - q3) <= 1e-12 * abs(q3)


@pytest.mark.parametrize("k0,a", [(1.0, 0.3), (7.5, 0.05), (30.0, 0.011), (2.0, 1.2)])
def test_full_vector_depolarisation_identity(k0, a):
    """G = (I + grad grad / k^2) g, so over a ball in d dimensions
    int G = (1 - 1/d) int g I - I / (d k^2): the vector "full" forms follow
    from the scalar integrals (checked by quadrature above)."""
    q2 = _quad_disk_tm(k0, a)
    v2 = O.self_interaction_integral_vector_2d(k0, a, "full")
    assert abs(v2 - (0.5 * q2 - 1 / (2 * k0 * k0))) <= 1e-12 * abs(v2)
    q3 = _quad_ball_scalar(k0, a)
    v3 = O.self_interaction_integral_vector_3d(k0, a, "full")
    assert abs(v3 - (2.0 / 3.0 * q3 - 1 / (3 * k0 * k0))) <= 1e-12 * abs(v3)


@pytest.mark.parametrize("kind", list(SI))
def test_rayleigh_is_small_particle_series_of_full(kind):
    """Rayleigh = the full form's expansion in k0 a through O(a^2) in the real
    part and O(a^3) in the imaginary part (2D: with the log term of Y1): the
    difference shrinks like a^4 (up to a log), so halving a divides it by
    about 16."""
    f, vol, _ = SI[kind]
    k0 = 3.0
    diffs = [abs(f(k0, a, "full") - f(k0, a, "Rayleigh")) for a in (0.02, 0.01, 0.005)]
    for big, small in zip(diffs, diffs[1:]):
        assert 10.0 < big / small < 20.0, (kind, diffs)
    # at small k0 a the remainder is a small part of the a^2 terms
    a = 0.005
    assert diffs[-1] <= 1e-3 * a * a, (kind, diffs)


@pytest.mark.parametrize("kind", list(SI))
def test_unknown_self_interaction_type(kind):
    f = SI[kind][0]
    with pytest.raises(NotImplementedError):
        f(1.0, 0.1, "Mie")


def test_dropin_module_functions_match_oracle():
    """The drop-in's module-level self_interaction_integral_* (utils.py:21-23
    imports them) equal the oracle's for both types."""
    for k0, a in ((1.0, 0.3), (12.0, 0.02)):
        for t in ("Rayleigh", "full"):
            assert D2.self_interaction_integral_scalar(k0, a, t) == O.self_interaction_integral_scalar_2d(k0, a, t)
            assert D2.self_interaction_integral_vector(k0, a, t) == O.self_interaction_integral_vector_2d(k0, a, t)
            assert D3.self_interaction_integral_vector(k0, a, t) == O.self_interaction_integral_vector_3d(k0, a, t)
            assert D3.self_interaction_integral_scalar(k0, a, t) == O.self_interaction_integral_scalar_3d(k0, a, t)


# ---------------------------------------------------------------------------
# propagate_ss
# ---------------------------------------------------------------------------
def _textbook_g3(r, k):
    R = np.linalg.norm(r)
    kr = k * R
    a = 1 + 1j / kr - 1 / kr ** 2
    b = -1 - 3j / kr + 3 / kr ** 2
    return np.exp(1j * kr) / (4 * np.pi * R) * (a * np.eye(3) + b * np.outer(r, r) / R ** 2)


def test_propagate_ss_single_scatterer_closed_form():
    """N = 1: E(r) = E0_meas(r) + alpha k0^2 G(r - r1) E0_scat (Born), for
    the 3D vector and 2D TM classes; the dyadic from the textbook form, the
    TM value from mpmath's Hankel function."""
    k0 = 4.0
    alpha = 0.01 - 0.002j
    r1 = np.array([0.1, -0.2, 0.05])
    s = O.Transmission3D_vector(torch.from_numpy(r1[None]), source='plane')
    pts = np.array([[1.0, 0.5, -0.3], [0.0, 2.0, 0.0], [-0.7, 0.1, 0.9]])
    rng = np.random.default_rng(3)
    E0s = rng.normal(size=(3, 2)) + 1j * rng.normal(size=(3, 2))
    E0m = rng.normal(size=(3, 3, 2)) + 1j * rng.normal(size=(3, 3, 2))
    got = s.propagate_ss(torch.from_numpy(pts), k0, alpha, torch.from_numpy(E0m), torch.from_numpy(E0s)).numpy()
    for m in range(3):
        ref = E0m[m] + alpha * k0 ** 2 * _textbook_g3(pts[m] - r1, k0) @ E0s
        assert np.max(np.abs(got[m] - ref)) <= 1e-13 * np.max(np.abs(ref))
    # 2D TM
    t = O.Transmission2D_scalar(torch.tensor([[0.3, -0.1]], dtype=torch.float64), source='plane')
    p2 = np.array([[1.0, 0.2], [-0.5, 0.8]])
    e0s = np.array([[1.0 + 0.5j, -0.2j]])
    e0m = np.array([[0.1, 0.2j], [0.3, -0.4]])
    got = t.propagate_ss(torch.from_numpy(p2), k0, alpha, torch.from_numpy(e0m), torch.from_numpy(e0s)).numpy()
    for m in range(2):
        R = float(np.linalg.norm(p2[m] - np.array([0.3, -0.1])))
        # torch's Bessel functions deviate from the exact ones by ~1e-8 here
        g = complex(0.25j * mp.hankel1(0, k0 * R))
        ref = e0m[m] + alpha * k0 ** 2 * g * e0s[0]
        assert np.max(np.abs(got[m] - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_propagate_ss_coincident_point_takes_e0_meas_row_of_scatterer_index():
    """A measurement point on scatterer j gives a NaN row, replaced by
    E0_meas[j], the measurement field at the row of the scatterer's index
    (Transmission3D.py:226-236), not at the measurement point's own row."""
    k0, alpha = 3.0, 0.02 - 0.001j
    r = np.array([[0.0, 0.0, 0.0], [0.5, 0.1, -0.2]])
    s = O.Transmission3D_vector(torch.from_numpy(r), source='plane')
    pts = np.array([[1.0, 1.0, 1.0], [2.0, -1.0, 0.5], [0.5, 0.1, -0.2]])   # point 2 sits on scatterer 1
    E0s = np.arange(12, dtype=np.float64).reshape(6, 2) * (1 + 0.5j)
    E0m = (np.arange(18, dtype=np.float64).reshape(3, 3, 2) + 100) * (1 - 0.25j)
    got = s.propagate_ss(torch.from_numpy(pts), k0, alpha, torch.from_numpy(E0m), torch.from_numpy(E0s)).numpy()
    assert np.array_equal(got[2], E0m[1])
    for m in range(2):
        ref = E0m[m] + alpha * k0 ** 2 * sum(_textbook_g3(pts[m] - r[j], k0) @ E0s[3 * j:3 * j + 3]
                                             for j in range(2))
        assert np.max(np.abs(got[m] - ref)) <= 1e-13 * np.max(np.abs(ref))


# ---------------------------------------------------------------------------
# periodic wrap
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", list(GREEN))
def test_periodic_shift_invariance(kind):
    """Along a periodic axis, G(r + e) = G(r) for the unit lattice vector e,
    and inside the cell [-1/2, 1/2) the wrap changes nothing.  The 2D kinds
    wrap for periodic == 'x' or 'y' only (an if / elif on the whole string:
    'xy' wraps nothing), the 3D kinds for every axis named."""
    rng = np.random.default_rng(7)
    sp = 2 if kind.startswith("2d") else 3
    axes = ["x", "y"] if sp == 2 else ["x", "y", "z", "xz", "xyz"]
    for per in axes:
        for _ in range(10):
            r = rng.uniform(-0.45, 0.45, size=sp)
            k0 = rng.uniform(2.0, 12.0)
            base = _g(kind, r.copy(), k0, periodic=per)
            for shift in (1.0, -2.0, 3.0):
                rs = r.copy()
                for name in per:
                    rs["xyz".index(name)] += shift
                assert np.allclose(_g(kind, rs, k0, periodic=per), base, rtol=1e-12, atol=0), (kind, per)
            assert np.allclose(base, _g(kind, r, k0), rtol=1e-12, atol=0)
    if sp == 2:
        r = np.array([0.7, 0.8])
        assert np.array_equal(_g(kind, r.copy(), 5.0, periodic="xy"), _g(kind, r.copy(), 5.0))
        assert np.allclose(_g(kind, r.copy(), 5.0, periodic="y"), _g(kind, np.array([0.7, -0.2]), 5.0),
                           rtol=1e-12, atol=0)


@pytest.mark.parametrize("kind", list(GREEN))
def test_kernel_periodic_flags_match_oracle(kind, host_kernels):
    """green.h's MG_F_PERIODIC_X/Y/Z flags (host build) against the oracle's
    wrap of the same difference vectors (2D: one flag at a time, the
    strings 'x' and 'y' the reference honours)."""
    kid = {"2d_tm": 0, "2d_te": 1, "3d_vector": 2, "3d_scalar": 3}[kind]
    flag = {"x": 16, "y": 32, "z": 64}
    d = DIM[kind]
    sp = 2 if kind.startswith("2d") else 3
    rng = np.random.default_rng(11)
    for per in (["x", "y"] if sp == 2 else ["x", "y", "z", "xy", "xyz"]):
        fl = sum(flag[c] for c in per)
        for _ in range(20):
            r = rng.uniform(-2.0, 2.0, size=sp)
            k0 = rng.uniform(1.0, 15.0)
            rr = np.zeros(3)
            rr[:sp] = r
            g = np.zeros(2 * d * d)
            host_kernels.shim_green(kid, rr.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(k0), fl,
                                    ctypes.c_double(0.0), g.ctypes.data_as(ctypes.c_void_p))
            got = (g[0::2] + 1j * g[1::2]).reshape(d, d)
            ref = _g(kind, r.copy(), k0, periodic=per)
            assert np.allclose(got, ref, rtol=1e-13, atol=1e-15 * np.abs(ref).max()), (kind, per, r)


# ---------------------------------------------------------------------------
# regularised NaN fix-up with several candidates
# ---------------------------------------------------------------------------
def _brute_regularized(kind, r, pts, k0, alpha, Ek, E0, radius):
    """Expected propagate(regularize=True) output, by hand: rows of points
    with a scatterer inside the radius (the kind's regularisation test)
    copy the nearest scatterer's field; other rows are alpha k0^2 G Ek + E0."""
    d = DIM[kind]
    N = r.shape[0]
    T = Ek.shape[1]
    out = []
    for m in range(pts.shape[0]):
        dist = np.linalg.norm(r - pts[m], axis=1)
        test = dist * (k0 if kind == "2d_te" else 1.0)   # TE compares k0 R (Transmission2D.py:86-89)
        if np.any(test < radius):
            j = int(np.argmin(dist))
            if kind == "2d_te":
                out.append(np.broadcast_to(Ek[j], (2, T)))      # flat row j of the (2N, T) Ek
            elif d == 3:
                out.append(Ek[3 * j:3 * j + 3])
            else:
                out.append(Ek[j])
            continue
        acc = np.zeros((d, T), dtype=complex)
        for j in range(N):
            acc += _g(kind, pts[m] - r[j], k0) @ Ek[d * j:d * j + d]
        row = alpha * k0 * k0 * acc + E0[m].reshape(d, T)
        out.append(row if d > 1 else row[0])
    return np.stack(out)


@pytest.mark.parametrize("kind", list(GREEN))
def test_regularize_multi_candidate_brute_force(kind):
    """Two scatterers inside the radius of a measurement point: the row takes
    the nearest one's Ek (torch.argmin over all distances); one candidate:
    that one's; none: the Green's-function row."""
    k0, alpha = 2.0, 0.05 - 0.01j
    sp = 2 if kind.startswith("2d") else 3
    r = np.zeros((4, sp))
    r[:, 0] = [0.0, 0.05, 1.0, -1.2]
    r[1, 1] = 0.02
    r[3, 1] = 0.3
    pts = np.zeros((3, sp))
    pts[0, 0] = 0.04                        # near scatterers 0 and 1 (1 is nearer)
    pts[1, 0] = 1.01                        # near scatterer 2 only
    pts[2, :2] = [0.4, 0.9]                 # near none
    radius = 0.1 * (k0 if kind == "2d_te" else 1.0)
    d = DIM[kind]
    rng = np.random.default_rng(2)
    Ek = rng.normal(size=(d * 4, 3)) + 1j * rng.normal(size=(d * 4, 3))
    E0 = rng.normal(size=(3, d, 3)) + 1j * rng.normal(size=(3, d, 3))
    cls = O.CLASSES[kind](torch.from_numpy(r), source='plane')
    E0t = torch.from_numpy(E0 if d > 1 else E0[:, 0])
    got = cls.propagate(torch.from_numpy(pts), torch.from_numpy(Ek), k0, alpha, E0t, regularize=True,
                        radius=radius).numpy()
    ref = _brute_regularized(kind, r, pts, k0, alpha, Ek, E0, radius)
    assert got.shape == ref.shape
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), kind
    # row 2: torch's Bessel functions against the same ones (2D), or exact (3D)
    assert np.allclose(got[2], ref[2], rtol=1e-12, atol=0), kind


# ---------------------------------------------------------------------------
# 3D point sources
# ---------------------------------------------------------------------------
def test_point_sources_3d_dipole_closed_form():
    """3D vector 'point': dipoles sqrt(I) p at r_s = -2u, field G(r - r_s)
    sqrt(I) p with I = 16 pi k0^2 (the reference's intensity rule for point
    sources at distance 2); 3D scalar: sqrt(I) exp(ik R)/(4 pi R)."""
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1.0, 1.0, size=(30, 3))
    k0 = 6.1
    u = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8], [-0.36, 0.48, 0.8]])
    p = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.8, 0.6, 0.0]])
    inten = (k0 * 2.0) ** 2 * 4.0 * np.pi
    vec = O.Transmission3D_vector(torch.from_numpy(pts), source="point").generate_source(
        torch.from_numpy(pts), k0, torch.from_numpy(u), torch.from_numpy(p), 0.2).numpy()
    sca = O.Transmission3D_scalar(torch.from_numpy(pts), source="point").generate_source(
        torch.from_numpy(pts), k0, torch.from_numpy(u), 0.2).numpy()
    assert vec.shape == (30, 3, 3) and sca.shape == (30, 3)
    for t in range(3):
        rs = -2.0 * u[t]
        for m in range(30):
            ref = np.sqrt(inten) * _textbook_g3(pts[m] - rs, k0) @ p[t]
            assert np.max(np.abs(vec[m, :, t] - ref)) <= 1e-13 * np.max(np.abs(ref))
            R = np.linalg.norm(pts[m] - rs)
            refs = np.sqrt(inten) * np.exp(1j * k0 * R) / (4 * np.pi * R)
            assert abs(sca[m, t] - refs) <= 1e-13 * abs(refs)
