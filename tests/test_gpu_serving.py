"""Serving loop: a request served in a busy mix returns the same bits as
served alone; results match the FP64 path; admission keeps the reserved
HBM within the budget; resident factors are reused."""
import math

import pytest
import torch

pytestmark = pytest.mark.gpu

from magreete_amd import serving as S
from magreete_amd.pipeline import Solver, problem

DEV = "cuda:0"


def say(*a):
    print(*a, flush=True)


def colerr(x, ref):
    return (torch.linalg.norm(x - ref, dim=0) / torch.linalg.norm(ref, dim=0)).max().item()


@pytest.fixture(scope="module")
def trace():
    return S.make_trace(14, math.inf, seed=5, n_range=(64, 600), t_range=(1, 24), reuse=0.3, n_measure=16)


@pytest.fixture(scope="module")
def mix(trace):
    srv = S.Server(DEV, workers=4, method="bf16x3", keep_outputs=True)
    return srv.run(trace)


def test_alone_vs_in_the_mix(trace, mix):
    say("serving alone vs mix", len(trace))
    assert any(r["hit"] for r in mix), "the trace should reuse a resident factor"
    for req, rm in zip(trace, mix):
        alone = S.Server(DEV, workers=1, method="bf16x3", cache=False, keep_outputs=True).run([req])[0]
        assert torch.equal(alone["X"], rm["X"]), req
        assert torch.equal(alone["E"], rm["E"]), req


def test_mix_matches_fp64(trace, mix):
    say("serving vs fp64")
    for req, rm in list(zip(trace, mix))[:6]:
        w = S.workload(req)
        s = Solver(problem(w, DEV), "fp64")
        s.run()
        assert colerr(rm["X"], s.X) < 1e-9, req
        assert colerr(rm["E"], s.out) < 1e-9, req


def test_admission_within_budget(trace):
    say("serving budget")
    big = max(S.resident_bytes(r.n, "bf16x3") + S.factor_bytes(r.n, "bf16x3")
              + S.transient_bytes(r.n, r.T, r.d * r.n_measure, "bf16x3") for r in trace)
    srv = S.Server(DEV, workers=4, method="bf16x3", hbm_budget=1.5 * big)
    res = srv.run(trace)
    assert all(r is not None for r in res)
    assert srv.peak_reserved <= 1.5 * big


def test_no_cache_no_hits(trace):
    say("serving no cache")
    res = S.Server(DEV, workers=2, method="fp64", cache=False).run(trace[:6])
    assert not any(r["hit"] for r in res)
    assert all(r["latency_ms"] >= r["service_ms"] > 0 for r in res)
