"""Ensemble executor (concurrent solves on one GPU) against one-at-a-time
solves and the oracle."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from magreete_amd import workloads as W
from magreete_amd.ensemble import Ensemble
from magreete_amd.pipeline import Solver, problem
from oracle import transmission as O

DEV = "cuda:0"


def say(*a):
    print(*a, flush=True)


def colerr(x, ref):
    return (torch.linalg.norm(x - ref, dim=0) / torch.linalg.norm(ref, dim=0)).max().item()


@pytest.fixture(scope="module")
def small_ensemble():
    return W.cfg_ensemble(n_configs=10, n_min=60, n_max=420, n_dirs=4, seed=3)


@pytest.fixture(scope="module")
def sequential(small_ensemble):
    out = []
    for w in small_ensemble:
        s = Solver(problem(w, DEV, measure=False), "fp64")
        s.run()
        out.append(s.X.clone())
    return out


@pytest.mark.parametrize("method,streams", [("fp64", 1), ("fp64", 4), ("bf16x3", 4), ("c64", 3)])
def test_ensemble_matches_sequential(small_ensemble, sequential, method, streams):
    say("ensemble", method, streams)
    res = Ensemble(DEV, streams=streams, method=method).run(small_ensemble)
    tol = 1e-11 if method == "fp64" else 1e-9
    for r, x in zip(res, sequential):
        assert int(r["info"].item()) == 0
        assert colerr(r["X"], x) < tol, (method, streams, colerr(r["X"], x))
        assert r["ms"] > 0
    if method == "fp64" and streams == 1:
        # one worker with the default leaf cap repeats the sequential solve bit for bit
        assert all(torch.equal(r["X"], x) for r, x in zip(res, sequential))


def test_ensemble_member_vs_oracle(small_ensemble):
    say("ensemble oracle")
    w = small_ensemble[1]
    res = Ensemble(DEV, streams=2, method="bf16x3").run([w, small_ensemble[0]])
    o = O.Transmission3D_vector(w.points, source="plane")
    E0 = o.generate_source(w.points, w.k0, w.u, w.p, 0.2)
    Ek = o.solve(w.k0, w.alpha, 0.0, E0, self_interaction=False)
    assert colerr(res[0]["X"].cpu(), Ek) < 1e-9


@pytest.mark.parametrize("method", ["fp64", "bf16x3"])
@pytest.mark.parametrize("group_bytes", [None, 6e6])
def test_grouped_ensemble_matches_sequential(small_ensemble, sequential, method, group_bytes):
    """Grouped lockstep LU (one group, or several when group_bytes caps them)
    against the one-at-a-time FP64 solves; the FP64 grouped path factors with
    the same pivoting rule, the BF16x3 one is refined to 1e-12 in FP64."""
    say("grouped ensemble", method, group_bytes)
    ens = Ensemble(DEV, method=method, grouped=True, group_bytes=group_bytes, tol=1e-12)
    res = ens.run(small_ensemble)
    if group_bytes is not None:
        assert max(r["stats"]["group"] for r in res) < len(small_ensemble)
    for r, x in zip(res, sequential):
        assert int(r["info"].item()) == 0
        err = colerr(r["X"], x)
        assert err < (1e-11 if method == "fp64" else 1e-10), (method, err, r["stats"])
        if method == "bf16x3":
            assert r["stats"]["fallback"] or r["stats"]["residual"] <= 1e-12


def test_grouped_ensemble_measure_vs_oracle():
    """Grouped BF16x3 ensemble of 2D TM configurations with measurement
    points against the oracle's solve and propagate."""
    say("grouped ensemble oracle")
    ws = [W.cfg_2d_tm(n_points=n, n_angles=8, n_measure=40, seed=s) for n, s in ((300, 1), (90, 2), (511, 3))]
    res = Ensemble(DEV, method="bf16x3", grouped=True, measure=True, tol=1e-12).run(ws)
    for w, r in zip(ws, res):
        o = O.Transmission2D_scalar(w.points, source="plane")
        E0 = o.generate_source(w.points, w.k0, w.thetas, 0.2)
        Ek = o.solve(w.k0, w.alpha, 0.0, E0, self_interaction=False).reshape(r["X"].shape)
        assert colerr(r["X"].cpu(), Ek) < 1e-10
        Em = o.propagate(w.measure, Ek, w.k0, w.alpha, o.generate_source(w.measure, w.k0, w.thetas, 0.2))
        assert colerr(r["out"].cpu(), Em.reshape(r["out"].shape)) < 1e-10


def test_grouped_refinement_uneven_convergence():
    """Members of one group that converge after different numbers of
    refinement steps keep their converged solutions while the others go on
    (sixteen cfg4 members with N <= 900, against the one-at-a-time FP64
    solves)."""
    say("grouped uneven convergence")
    full = W.cfg_ensemble(n_configs=256, seed=0)
    ws = [w for w in full if w.points.shape[0] <= 900][:16]
    res = Ensemble(DEV, method="bf16x3", grouped=True, tol=1e-12).run(ws)
    iters = [r["stats"]["iters"] for r in res]
    say("iters", iters)
    assert len(set(iters)) > 1
    for w, r in zip(ws, res):
        s = Solver(problem(w, DEV, measure=False), "fp64")
        s.run()
        assert colerr(r["X"], s.X) < 1e-10, (r["stats"], colerr(r["X"], s.X))
