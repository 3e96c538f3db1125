"""The 2D TE perturbed-lattice configuration (LDOS at every site and at probe
points, incidence fields in mixed precision) and the diagonal blocks of
M^{-1} it rests on, against numpy's inverse and the oracle."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from magreete_amd import engine as E
from magreete_amd import workloads as W
from magreete_amd.Transmission2D import Transmission2D_vector
from magreete_amd.pipeline import Solver, problem
from oracle import transmission as O

DEV = "cuda:0"


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,d", [(7, 1), (300, 3), (512, 2), (1000, 1), (2900, 2), (3001, 3)])
def test_inverse_diag_blocks_vs_numpy(n, d):
    """Diagonal blocks of M^{-1} from the pivoted LU (triangular inverses in
    place, then the row-column gather with the pivot permutation) against
    numpy's inverse: ragged last blocks, a pivoting-heavy matrix, a subset of
    sites in scrambled order with a repeat."""
    say("inverse diag blocks", n, d)
    rng = np.random.default_rng(n + d)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A[np.arange(n), np.arange(n)] *= 0.01         # small diagonal: rows get swapped
    nsite = n // d
    sites = rng.permutation(nsite)[: max(1, (2 * nsite) // 3)]
    sites = np.concatenate([sites, sites[:1]])
    F = E.Factor(torch.from_numpy(A).to(DEV))
    assert int(F.info.item()) == 0
    blocks = E.inverse_diag_blocks(F, torch.from_numpy(sites), d).cpu().numpy()
    Winv = np.linalg.inv(A)
    ref = np.stack([Winv[d * s:d * s + d, d * s:d * s + d] for s in sites])
    err = np.abs(blocks - ref).max() / np.abs(ref).max()
    # both inverses carry ~eps cond(A) of error; these random matrices with a
    # shrunken diagonal have cond(A) up to ~1e5 (measured 1.8e-11 at n = 3001)
    assert err < 1e-10, err
    with pytest.raises(RuntimeError):
        F.solve(torch.zeros((n, 1), dtype=torch.complex128, device=DEV))


@pytest.mark.parametrize("kind", ["2d_tm", "2d_te", "3d_vector", "3d_scalar"])
def test_site_ldos_paths_agree(kind):
    """LDOS at the scatterers from the diagonal blocks equals the solve
    against G_p^T, with and without self-interaction (whose NaN fix-up puts
    S/V in every entry of the self block)."""
    from magreete_amd import Transmission2D as T2, Transmission3D as T3
    cls = {"2d_tm": T2.Transmission2D_scalar, "2d_te": T2.Transmission2D_vector,
           "3d_vector": T3.Transmission3D_vector, "3d_scalar": T3.Transmission3D_scalar}[kind]
    n = {"2d_tm": 700, "2d_te": 400, "3d_vector": 300, "3d_scalar": 600}[kind]
    say("site ldos paths", kind, n)
    rng = np.random.default_rng(3)
    pts = torch.from_numpy(W.unit_disk(n, rng) if kind.startswith("2d") else W.unit_ball(n, rng))
    k0 = 1.0 / W.mean_nn_spacing(pts.numpy())
    alpha = {"2d_tm": W.alpha_2d_tm, "2d_te": W.alpha_2d_te}.get(kind, W.alpha_3d)(k0)
    radius = 0.2 / k0
    meas = torch.cat([pts[::3], 0.5 * pts[:5]])
    for si in (False, True):
        a = cls(pts, "plane")
        b = cls(pts, "plane")
        b.ldos_site_blocks = False
        la = a.LDOS_measurements(meas, k0, alpha, radius, self_interaction=si)
        lb = b.LDOS_measurements(meas, k0, alpha, radius, self_interaction=si)
        err = (la - lb).abs().max().item() / lb.abs().max().item()
        assert err < 1e-10, (si, err)


@pytest.mark.parametrize("k0a", [0.6, 0.9, 1.2, 1.5])
def test_cfg3_lattice_vs_oracle(k0a):
    """The perturbed triangular lattice at ~1500 sites (n ~ 3000) for each
    k0 a of the configuration: LDOS at every site (diagonal blocks of M^{-1})
    and at 256 probes (solves), batched as magreete does, and 64 incidence
    fields by the BF16x3 factor with FP64 GMRES refinement, against the
    oracle."""
    say("cfg3", k0a)
    w = W.cfg_2d_te_lattice(n_sites=1500, k0a=k0a, n_angles=64, n_probes=256, seed=4)
    pts = torch.cat([w.points, w.probes])
    batches = torch.split(pts, 500)
    dut = Transmission2D_vector(w.points, "plane")
    ora = O.Transmission2D_vector(w.points, "plane")
    a = dut.batch_LDOS_measurements(batches, w.k0, w.alpha, 0.1, self_interaction=False)
    b = ora.batch_LDOS_measurements(batches, w.k0, w.alpha, 0.1, self_interaction=False)
    assert a.shape == b.shape == (pts.shape[0], 1)
    err = (a - b).abs().max().item() / b.abs().max().item()
    say("ldos err", err)
    assert err < 1e-10
    # fields: mixed precision + refinement
    prob = problem(w, DEV, measure=False)
    s = Solver(prob, "bf16x3", tol=1e-10, relax=True)
    s.run()
    say(s.stats)
    cE0 = ora.generate_source(w.points, w.k0, w.thetas, 0.2)
    cEk = ora.solve(w.k0, w.alpha, 0.0, cE0, self_interaction=False).reshape(s.X.shape)
    X = s.X.cpu()
    ferr = (torch.linalg.norm(X - cEk, dim=0) / torch.linalg.norm(cEk, dim=0)).max().item()
    say("field err", ferr)
    assert ferr < 1e-9
