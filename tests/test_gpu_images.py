"""The full-field image flow (magreete.py:568-741): one multi-RHS
propagation of all plotted angles over the grid (magreete_amd.images)
against the oracle's per-angle, per-batch loop, for the four classes, with
grid points on scatterers (NaN fix-up, with and without regularize)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from magreete_amd import Transmission2D as T2
from magreete_amd import Transmission3D as T3
from magreete_amd import images as IM
from magreete_amd import workloads as W
from oracle import transmission as O

DUT = {"2d_tm": T2.Transmission2D_scalar, "2d_te": T2.Transmission2D_vector,
       "3d_vector": T3.Transmission3D_vector, "3d_scalar": T3.Transmission3D_scalar}


@pytest.mark.parametrize("kind", list(DUT))
@pytest.mark.parametrize("regularize", [False, True])
def test_image_flow_vs_reference_loop(kind, regularize):
    print("images", kind, regularize, flush=True)
    ndim = 2 if kind.startswith("2d") else 3
    rng = np.random.default_rng(7)
    n = 150
    pts = W.unit_disk(n, rng) if ndim == 2 else W.unit_ball(n, rng)
    gridsize = (41, 37)
    grid = IM.image_grid(ndim, gridsize, L=1.0, window_width=1.2)
    # put three scatterers exactly on grid points
    pts[:3] = grid.numpy()[[100, 700, 1200]]
    pts = torch.from_numpy(pts)
    k0 = 1.0 / W.mean_nn_spacing(pts.numpy())
    alpha = {"2d_tm": W.alpha_2d_tm, "2d_te": W.alpha_2d_te}.get(kind, W.alpha_3d)(k0)
    thetas = 2 * np.pi * np.arange(16) / 16
    plot = thetas[[0, 5, 11]]
    index = [0, 5, 11]
    dut, ora = DUT[kind](pts, "plane"), O.CLASSES[kind](pts, "plane")
    if ndim == 2:
        E0s = ora.generate_source(pts, k0, thetas, 0.2)
        kw = dict(thetas=plot)
        inc = None
    else:
        u, p = IM.incidence_3d(thetas, polarization_angle_radians=0.3)
        E0s = ora.generate_source(pts, k0, u, 0.2) if kind == "3d_scalar" else ora.generate_source(pts, k0, u, p, 0.2)
        kw = dict(u=u[index], p=p[index])
        inc = lambda a: tuple(t[0] for t in IM.incidence_3d([a], polarization_angle_radians=0.3))
    Ej = ora.solve(k0, alpha, 0.0, E0s, self_interaction=False)
    radius = 0.3 / k0
    E, E0 = IM.image_fields(dut, grid, Ej, index, k0, alpha, 0.2, regularize=regularize, radius=radius, **kw)
    ref = O.image_fields_reference(ora, grid, Ej, thetas, plot, k0, alpha, 0.2, batch_size=500, incidence=inc,
                                   regularize=regularize, radius=radius)
    maps = IM.field_maps(E, grid, gridsize)
    for a in range(len(plot)):
        Ea = E[..., a]
        # the reference keeps scalar fields as (P, 1)
        assert Ea.shape == (ref[a].shape[:1] if dut.d == 1 else ref[a].shape)
        ref[a] = ref[a].reshape(Ea.shape)
        assert not torch.isnan(ref[a]).any()
        err = (Ea - ref[a]).abs().max().item() / ref[a].abs().max().item()
        assert err < 1e-10, (a, err)
        rm = O.field_maps_reference(ref[a], grid, gridsize[0], gridsize[1], dut.d == 1)
        for key, val in rm.items():
            got = maps[key][a]
            m = ~torch.isnan(val)
            assert torch.equal(torch.isnan(got), torch.isnan(val)), key
            assert (got[m] - val[m]).abs().max().item() <= 1e-10 * val[m].abs().max().item(), key
