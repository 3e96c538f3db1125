"""Propagation with the incident field evaluated in its final pass
(mg_propagate_src) against mg_sources followed by mg_propagate_ws with
beta 1, for every kind, plane waves and beams, with and without the
scatterer split; and the solve pipeline with and without the fused source."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from magreete_amd import engine as E
from magreete_amd import workloads as W
from magreete_amd.pipeline import Solver, problem

DEV = "cuda:0"
SPACE = {"2d_tm": 2, "2d_te": 2, "3d_vector": 3, "3d_scalar": 3}


def say(*a):
    print(*a, flush=True)


def _dirs_pol(kind, T, rng):
    if SPACE[kind] == 2:
        th = torch.from_numpy(rng.uniform(0, 2 * np.pi, T))
        dirs = torch.stack([torch.cos(-th), torch.sin(-th)], dim=1)
        pol = torch.stack([dirs[:, 1], dirs[:, 0]], dim=1) if kind == "2d_te" else None
    else:
        u = torch.from_numpy(rng.normal(size=(T, 3)))
        dirs = u / torch.linalg.norm(u, dim=1, keepdim=True)
        pol = torch.from_numpy(rng.normal(size=(T, 3))) if kind == "3d_vector" else None
    return dirs, pol


@pytest.mark.parametrize("kind", ["2d_tm", "2d_te", "3d_vector", "3d_scalar"])
@pytest.mark.parametrize("beam", [False, True])
@pytest.mark.parametrize("n_meas,n_scat,T", [(40, 1500, 5), (2500, 60, 33), (1, 1, 1)])
def test_fused_source_matches_separate(kind, beam, n_meas, n_scat, T):
    say("fused source", kind, beam, n_meas, n_scat, T)
    rng = np.random.default_rng(n_meas + n_scat + T)
    S, d = SPACE[kind], E.DIM[kind]
    scat = torch.from_numpy(rng.uniform(-1, 1, (n_scat, S))).to(DEV)
    meas = torch.from_numpy(rng.uniform(-1.5, 1.5, (n_meas, S))).to(DEV)
    k0 = 9.0
    alpha = complex(-0.3, 0.7)
    Ek = torch.from_numpy(rng.normal(size=(d * n_scat, T)) + 1j * rng.normal(size=(d * n_scat, T))).to(DEV)
    dirs, pol = _dirs_pol(kind, T, rng)
    ref = E.sources(S, beam, meas, dirs, pol, k0, 0.3).reshape(d * n_meas, T).contiguous()
    E.propagate(kind, meas, scat, k0, alpha, Ek, ref)
    got = torch.full((d * n_meas, T), float("nan"), dtype=torch.complex128, device=DEV)
    E.propagate_src(kind, meas, scat, k0, alpha, Ek, got, beam, dirs, pol, 0.3)
    err = ((got - ref).abs().max() / ref.abs().max()).item()
    assert err <= 1e-14, err


def test_fused_source_pipeline():
    """cfg-shaped 3D vector problem: the solve pipeline's measurement fields
    with the fused source equal the ones with E0 built beforehand."""
    w = W.cfg_3d_vector(n_points=600, n_dirs=8, n_measure=300, seed=2)
    a = Solver(problem(w, DEV), "fp64")
    b = Solver(problem(w, DEV, fused_source=True), "fp64")
    assert b.p.E0m is None
    oa, ob = a.run(), b.run()
    err = ((oa - ob).abs().max() / oa.abs().max()).item()
    say("pipeline fused err", err)
    assert err <= 1e-13
