"""Host logic of the serving loop and the ensemble (no GPU): request traces,
workload determinism, footprints."""
import math

import numpy as np
import torch

from magreete_amd import serving as S
from magreete_amd import workloads as W
from magreete_amd.ensemble import cost


def test_trace_statistics():
    tr = S.make_trace(4000, rate=50.0, seed=1)
    arr = np.array([r.arrival for r in tr])
    assert np.all(np.diff(arr) >= 0)
    gaps = np.diff(arr)
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0                   # Poisson: mean gap 1/rate
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1                  # exponential: CV 1
    kinds = [r.kind for r in tr]
    frac = {k: kinds.count(k) / len(kinds) for k in S.KINDS}
    assert abs(frac["2d_tm"] - 0.5) < 0.05 and abs(frac["3d_vector"] - 0.25) < 0.05
    keys = [r.key for r in tr]
    reused = sum(1 for i, k in enumerate(keys) if k in set(keys[:i]))
    assert abs(reused / len(tr) - 0.3) < 0.05
    Ns = np.array([r.n_points for r in tr])
    assert Ns.min() >= 256 and Ns.max() <= 16384
    # log-uniform: the median of log N sits mid-range
    assert abs(np.median(np.log(Ns)) - 0.5 * (math.log(256) + math.log(16384))) < 0.3
    Ts = np.array([r.T for r in tr])
    assert Ts.min() >= 1 and Ts.max() <= 256


def test_burst_trace_and_reuse_keeps_geometry():
    tr = S.make_trace(50, rate=math.inf, seed=2, n_range=(64, 200), t_range=(3, 9))
    assert all(r.arrival == 0.0 for r in tr)
    by_key = {}
    for r in tr:
        by_key.setdefault(r.key, []).append(r)
    shared = [rs for rs in by_key.values() if len(rs) > 1]
    assert shared
    a, b = shared[0][0], shared[0][1]
    wa, wb = S.workload(a), S.workload(b)
    assert a.kind == b.kind and a.n_points == b.n_points
    assert torch.equal(wa.points, wb.points) and wa.k0 == wb.k0 and wa.alpha == wb.alpha
    assert wa.measure.shape == (a.n_measure, 2 if a.kind != "3d_vector" else 3)


def test_workload_incidences():
    r = S.Request(0, "3d_vector", 100, 7, key=11, arrival=0.0)
    w = S.workload(r)
    assert w.u.shape == (7, 3) and w.p.shape == (7, 3)
    # polarisations are unit and transverse
    assert np.allclose((w.u * w.p).sum(1).numpy(), 0.0, atol=1e-12)
    assert np.allclose(torch.linalg.norm(w.p, dim=1).numpy(), 1.0)
    r2 = S.Request(1, "2d_te", 100, 5, key=12, arrival=0.0)
    w2 = S.workload(r2)
    assert len(w2.thetas) == 5 and w2.alpha == W.alpha_2d_te(w2.k0)


def test_footprints_and_cost():
    n = 12288
    assert S.resident_bytes(n, "fp64") >= 16 * n * n
    assert S.resident_bytes(n, "bf16x3") >= 24 * n * n
    assert S.transient_bytes(n, 256, 2048 * 3, "bf16x3") > S.transient_bytes(n, 256, 2048 * 3, "fp64")
    ens = W.cfg_ensemble(n_configs=4, n_min=500, n_max=3000, n_dirs=2, seed=0)
    assert all(cost(w) == float(3 * w.points.shape[0]) ** 3 for w in ens)
