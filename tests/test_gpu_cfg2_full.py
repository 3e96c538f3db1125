"""The 3D vector configuration at full size (N = 4096 points, n = 12288, 128
directions x 2 polarisations, 2048 measurement points at radius 2) through
the solve pipeline, FP64 and BF16x3 + FP64 GMRES-IR, against the oracle's
solve and propagate (torch CPU complex128, ~9 s on the host).  At this size
the pivoted LU's leaf runs on more than 64 workgroups, so its grid barrier
takes the group-counter path."""
import pytest
import torch

pytestmark = pytest.mark.gpu

from magreete_amd import workloads as W
from magreete_amd.pipeline import Solver, problem
from oracle import transmission as O

DEV = "cuda:0"


def say(*a):
    print(*a, flush=True)


def colerr(x, ref):
    x = x.reshape(-1, x.shape[-1])
    ref = ref.reshape(-1, ref.shape[-1])
    return (torch.linalg.norm(x - ref, dim=0) / torch.linalg.norm(ref, dim=0)).max().item()


@pytest.fixture(scope="module")
def cfg2():
    w = W.cfg_3d_vector(n_points=4096, n_dirs=128, n_measure=2048, seed=0)
    torch.set_num_threads(max(1, torch.get_num_threads()))
    o = O.Transmission3D_vector(w.points, source="plane")
    E0s = o.generate_source(w.points, w.k0, w.u, w.p, 0.2)
    E0m = o.generate_source(w.measure, w.k0, w.u, w.p, 0.2)
    say("cfg2 oracle solve")
    Ek = o.solve(w.k0, w.alpha, 0.0, E0s, self_interaction=False)
    E = o.propagate(w.measure, Ek, w.k0, w.alpha, E0m)
    return w, Ek, E


@pytest.mark.parametrize("method,tol", [("fp64", 1e-10), ("bf16x3", 1e-9)])
def test_cfg2_full_size_vs_oracle(cfg2, method, tol):
    w, Ek, E = cfg2
    say("cfg2 full", method)
    s = Solver(problem(w, DEV, fused_source=True), method, tol=1e-10 if method != "fp64" else 1e-12, relax=True)
    out = s.run()
    torch.cuda.synchronize()
    say(method, s.stats)
    assert int(s.info.item()) == 0
    ek, e = colerr(s.X.cpu(), Ek), colerr(out.cpu(), E)
    say("Ek err", ek, "E err", e)
    assert ek <= tol and e <= tol
    if method != "fp64":
        assert not s.stats["fallback"] and s.stats["residual"] <= 1e-10
