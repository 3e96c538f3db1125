"""Full-field images of the incident plus scattered field on a grid
(magreete.py:568-741).

The reference loops over every plotted incidence angle and every batch of
grid points, with one single-column propagate per (angle, batch)
(magreete.py:639-673).  Here all plotted angles go through one multi-right-hand
side propagation over the whole grid (the fused Green's-function kernel:
the (d P) x (d N) matrix never exists), and the incident fields of all of
them come from one source evaluation.  Each point's field depends only on
that point, so the batching of the reference changes nothing but memory;
the NaN fix-up of points on scatterers is the drop-in's, as in propagate.
The derived maps (amplitude, longitudinal and transverse parts) follow
magreete.py:683-703; plots and hickle dumps are left to the caller.
"""
import numpy as onp
import torch


def image_grid(ndim, gridsize=(301, 301), L=1.0, window_width=1.2, slice_coordinate=0):
    """The measurement grid of the image flow (magreete.py:573-581): (P, ndim)
    float64, P = ngridx ngridy, x running fastest; in 3D the z = 0 plane,
    its axes rolled by slice_coordinate."""
    ngridx, ngridy = gridsize
    xyratio = ngridx / ngridy
    xs = onp.linspace(0, xyratio, ngridx) - xyratio / 2.0
    ys = onp.linspace(0, 1, ngridy) - 0.5
    if ndim == 2:
        x, y = onp.meshgrid(xs, ys)
        pts = onp.vstack([x.ravel(), y.ravel()]).T
    else:
        x, y, z = onp.roll(onp.meshgrid(xs, ys, [0.0]), slice_coordinate, axis=0)
        pts = onp.vstack([x.ravel(), y.ravel(), z.ravel()]).T
    return torch.from_numpy(pts * L * window_width)


def incidence_3d(thetas, polarization_angle_radians=0.0, switch_angle_scans=False, rotate_u=(0, 0)):
    """Incidence directions u (in the xy plane at angles thetas) and
    polarisations p (z rotated towards the azimuthal direction by the
    polarisation angle) of the 3D image flow (utils.py:138-161); swapped
    when the scan is over polarisations.  Rotated frames are not supported."""
    if tuple(rotate_u) != (0, 0):
        raise NotImplementedError("rotated incidence frames")
    th = onp.asarray(thetas, dtype=onp.float64).reshape(-1)
    u = onp.stack([onp.cos(th), onp.sin(th), onp.zeros(len(th))]).T
    e_alt = onp.zeros_like(u)
    e_alt[:, 2] = 1.0
    e_azim = onp.stack([onp.sin(th), -onp.cos(th), onp.zeros(len(th))]).T
    p = e_alt * onp.cos(polarization_angle_radians) + e_azim * onp.sin(polarization_angle_radians)
    u, p = torch.from_numpy(u), torch.from_numpy(p)
    return (p, u) if switch_angle_scans else (u, p)


def image_fields(solver, grid, Ej, index, k0, alpha, beam_waist, thetas=None, u=None, p=None, regularize=False,
                 radius=0.0):
    """Total fields at the grid for the incidence columns `index` of Ej, as
    (P, A) for scalar classes and (P, d, A) for vector ones, with the
    incident fields E0 of the same shape: one source evaluation and one
    multi-right-hand-side propagation for all A angles (magreete.py:639-673,
    which runs one per angle and grid batch).  2D: thetas of the plotted
    angles; 3D: their directions u (A, 3) and, for the vector class,
    polarisations p (A, 3)."""
    index = torch.as_tensor(index, dtype=torch.int64).reshape(-1)
    cols = Ej.reshape(Ej.shape[0], -1)[:, index.to(Ej.device)]
    if solver.space == 2:
        E0 = solver.generate_source(grid, k0, onp.asarray(thetas, dtype=onp.float64).reshape(-1), beam_waist)
    elif solver.d == 1:
        E0 = solver.generate_source(grid, k0, torch.as_tensor(u).reshape(-1, 3), beam_waist)
    else:
        E0 = solver.generate_source(grid, k0, torch.as_tensor(u).reshape(-1, 3), torch.as_tensor(p).reshape(-1, 3),
                                    beam_waist)
    E0 = E0.reshape(grid.shape[0], -1) if solver.d == 1 else E0.reshape(grid.shape[0], solver.d, -1)
    E = solver.propagate(grid, cols, k0, alpha, E0.to(cols.device), regularize=regularize, radius=radius)
    return E, E0


def field_maps(E, grid, gridsize):
    """Per plotted angle, the maps magreete plots (magreete.py:683-703), each
    (A, ngridy, ngridx): scalar classes: the complex field; vector classes:
    amplitude sqrt(sum |E_a|^2), longitudinal part (2D: along the viewing
    angle atan2(y, x); 3D: along the viewing direction) and transverse part
    (2D: the orthogonal component; 3D: the modulus of the remainder)."""
    ngridx, ngridy = gridsize
    grid = grid.to(E.device)
    if E.dim() == 2:
        return {"field": E.T.reshape(-1, ngridy, ngridx)}
    Ev = E.permute(2, 0, 1)                                    # (A, P, d)
    amp = torch.sqrt(torch.sum(torch.abs(Ev) ** 2, dim=-1))
    if E.shape[1] == 2:
        view = torch.arctan2(grid[:, 1], grid[:, 0])
        lon = Ev[..., 0] * torch.cos(view) - Ev[..., 1] * torch.sin(view)
        tra = Ev[..., 0] * torch.sin(view) + Ev[..., 1] * torch.cos(view)
    else:
        vu = grid / torch.linalg.norm(grid, axis=-1).unsqueeze(-1)
        lon = torch.sum(Ev * vu, dim=-1)
        rest = Ev - lon.unsqueeze(-1) * vu
        tra = torch.sqrt(torch.sum(torch.abs(rest) ** 2, dim=-1))
    shape = (-1, ngridy, ngridx)
    return {"amplitude": amp.reshape(shape), "longitudinal": lon.reshape(shape), "transverse": tra.reshape(shape)}
