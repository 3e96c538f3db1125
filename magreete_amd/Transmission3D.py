"""Drop-in for the reference module Transmission3D (3D coupled dipoles).

Same module-level names (I, self_interaction_integral_vector,
self_interaction_integral_scalar) and classes (Transmission3D_vector,
Transmission3D_scalar) with the reference's signatures
(Transmission3D.py:11-37, 39-439, 565-949).  Compute runs in
libmagreete_hip on the GPU.
"""
import numpy as onp
import torch

from . import _lib as L
from . import engine as E
from ._dropin import _Base, _like

I = torch.tensor(onp.identity(3)).reshape(1, 3, 3)  # Transmission3D.py:11


def self_interaction_integral_vector(k0, radius, self_interaction_type="Rayleigh"):
    """Vector self-interaction integral S (Transmission3D.py:13-24)."""
    volume = 4.0 * onp.pi * radius ** 3 / 3.0
    if self_interaction_type == "full":
        return (1.0 / k0 ** 2) * ((2.0 / 3.0) * onp.exp(1j * k0 * radius) * (1 - 1j * k0 * radius) - 1.0)
    if self_interaction_type == "Rayleigh":
        return -1.0 / (3.0 * k0 ** 2) + radius ** 2 / 3.0 + 1j * k0 * volume / (6.0 * onp.pi)
    raise NotImplementedError


def self_interaction_integral_scalar(k0, radius, self_interaction_type="Rayleigh"):
    """Scalar self-interaction integral S (Transmission3D.py:26-37)."""
    volume = 4.0 * onp.pi * radius ** 3 / 3.0
    if self_interaction_type == "full":
        return (1.0 / k0 ** 2) * (onp.exp(1j * k0 * radius) * (1 - 1j * k0 * radius) - 1.0)
    if self_interaction_type == "Rayleigh":
        return radius ** 2 / 2.0 + 1j * k0 * volume / (4.0 * onp.pi)
    raise NotImplementedError


def _pvec(p):
    """Polarisation rebuilt through atan2/arccos, as (T, 3) (Transmission3D.py:102-104)."""
    p = torch.as_tensor(p, dtype=torch.float64)
    phi = torch.arctan2(p[:, 1], p[:, 0])
    theta = torch.arccos(p[:, 2])
    return torch.stack([torch.sin(theta) * torch.cos(phi), torch.sin(theta) * torch.sin(phi), p[:, 2]], dim=1)


class _Base3D(_Base):
    space = 3

    def _volume(self, radius):
        return 4 * onp.pi * (radius ** 3) / 3

    def _wrap(self, r, periodic):
        # Transmission3D.py:58-69
        for axis, name in enumerate("xyz"):
            if name in periodic:
                r[:, :, axis] += 0.5
                r[:, :, axis] %= 1
                r[:, :, axis] -= 0.5

    def greens(self, r, k0, periodic='', regularize=False, radius=0.0):
        """Green's function of (M, N, 3) difference vectors (Transmission3D.py:47-77, 573-601)."""
        return self._green_raw(r, k0, periodic, regularize, radius)

    def G0(self, points, k0, print_statement='', regularize=False, radius=0.0):
        """(dM, dN) Green's matrix; self blocks zeroed when points is None (Transmission3D.py:240-268, 751-777)."""
        G, ref = self._G0(points, k0, regularize, radius)
        return _like(G, ref)

    def _shift(self, k0, alpha, radius, kind):
        return k0 ** 2 * alpha * self._self_int(k0, radius, kind) / (4 * onp.pi * (radius ** 3) / 3)


class Transmission3D_vector(_Base3D):
    """3D vector (dyadic) coupled dipoles, d = 3 (Transmission3D.py:39-439)."""
    kind = "3d_vector"
    d = 3

    def _self_int(self, k0, radius, kind):
        return self_interaction_integral_vector(k0, radius, kind)

    def generate_source(self, points, k0, u, p, w, print_statement=''):
        """Transmission3D.py:79-145.

        The reference's 'point' branch calls an undefined method
        (Transmission3D.py:136); here it is the field G(r - r_s) p of dipoles
        sqrt(I) p placed at r_s = -2u, with the reference's intensity."""
        if self.source in ("beam", "plane"):
            pol = _pvec(p)
            return _like(E.sources(3, self.source == "beam", points, torch.as_tensor(u), pol, k0, w), points)
        if self.source == "point":
            dev = E.device_of(points)
            intensity = 1.0 * (k0 * 2.0) ** 2 * 4.0 * onp.pi
            loc = (2.0 * (-torch.as_tensor(u, dtype=torch.float64))).to(dev)
            G, _ = E.assemble(self.kind, points.to(dev), loc, k0)           # (3M, 3T)
            moment = (onp.sqrt(intensity) * torch.as_tensor(p, dtype=torch.float64)).to(dev, torch.complex128)
            out = torch.einsum('matb,tb->mat', G.reshape(points.shape[0], 3, -1, 3), moment)
            return _like(out, points)
        if self.source is None:
            return torch.zeros((points.shape[0], u.shape[1], u.shape[0]), dtype=torch.complex128)
        raise NotImplementedError

    def solve(self, k0, alpha, radius, E0j, self_interaction=True, self_interaction_type="Rayleigh"):
        """Transmission3D.py:181-206."""
        return self._solve(k0, alpha, radius, E0j, self_interaction, self_interaction_type, flat=False)

    def propagate(self, points, Ek, k0, alpha, E0j, regularize=False, radius=0.0):
        """Transmission3D.py:147-179."""
        out, _ = self._propagate(points, Ek, k0, alpha, E0j, regularize, radius)
        src = Ek.reshape(self.N, 3, -1)
        out = self._fix_rows(out, points, src, src, regularize, radius)
        return _like(out, Ek)

    def propagate_ss(self, points, k0, alpha, E0_meas, E0_scat, regularize=False, radius=0.0):
        """Transmission3D.py:208-238."""
        points = torch.as_tensor(points)
        out, _ = self._propagate(points, E0_scat, k0, alpha, E0_meas, regularize, radius)
        out = self._fix_rows(out, points, E0_meas, E0_meas, regularize, radius)
        return _like(out, E0_meas)

    def mean_DOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                              self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:270-323."""
        return self._mean_dos(measure_points, k0, alpha, radius, self_interaction, self_interaction_type,
                              regularize, discard_absorption, 2.0 * onp.pi * k0)

    def LDOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                          self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:325-377."""
        return self._ldos(measure_points, k0, alpha, radius, self_interaction, self_interaction_type, regularize,
                          discard_absorption, 2.0 * onp.pi * k0)

    def batch_LDOS_measurements(self, batches, k0, alpha, radius, self_interaction=True,
                                self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:379-439."""
        return self._batch_ldos(batches, k0, alpha, radius, self_interaction, self_interaction_type, regularize,
                                discard_absorption, 2.0 * onp.pi * k0)


class Transmission3D_scalar(_Base3D):
    """3D scalar coupled dipoles, d = 1 (Transmission3D.py:565-949)."""
    kind = "3d_scalar"
    d = 1

    def _self_int(self, k0, radius, kind):
        return self_interaction_integral_scalar(k0, radius, kind)

    def generate_source(self, points, k0, u, w, print_statement=''):
        """Transmission3D.py:603-658.

        The reference's 'point' branch fails on a numpy scalar
        (Transmission3D.py:651); here it is sqrt(I) G(r - r_s), r_s = -2u."""
        if self.source in ("beam", "plane"):
            return _like(E.sources(3, self.source == "beam", points, torch.as_tensor(u), None, k0, w), points)
        if self.source == "point":
            dev = E.device_of(points)
            intensity = 1.0 * (k0 * 2.0) ** 2 * 4.0 * onp.pi
            loc = (2.0 * (-torch.as_tensor(u, dtype=torch.float64))).to(dev)
            G, _ = E.assemble(self.kind, points.to(dev), loc, k0)
            return _like(G * onp.sqrt(intensity), points)
        if self.source is None:
            return torch.zeros((points.shape[0], u.shape[1]), dtype=torch.complex128)
        raise NotImplementedError

    def solve(self, k0, alpha, radius, E0j, self_interaction=True, self_interaction_type="Rayleigh"):
        """Transmission3D.py:693-718."""
        return self._solve(k0, alpha, radius, E0j, self_interaction, self_interaction_type, flat=False)

    def propagate(self, points, Ek, k0, alpha, E0j, regularize=False, radius=0.0):
        """Transmission3D.py:660-691."""
        out, _ = self._propagate(points, Ek, k0, alpha, E0j, regularize, radius)
        src = Ek.reshape(self.N, -1)
        out = self._fix_rows(out, points, src, src, regularize, radius)
        return _like(out, Ek)

    def propagate_ss(self, points, k0, alpha, E0_meas, E0_scat, regularize=False, radius=0.0):
        """Transmission3D.py:720-749."""
        points = torch.as_tensor(points)
        out, _ = self._propagate(points, E0_scat, k0, alpha, E0_meas, regularize, radius)
        out = self._fix_rows(out, points, E0_meas, E0_meas, regularize, radius)
        return _like(out, E0_meas)

    def mean_DOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                              self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:779-832."""
        return self._mean_dos(measure_points, k0, alpha, radius, self_interaction, self_interaction_type,
                              regularize, discard_absorption, 4.0 * onp.pi * k0)

    def LDOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                          self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:834-886."""
        return self._ldos(measure_points, k0, alpha, radius, self_interaction, self_interaction_type, regularize,
                          discard_absorption, 4.0 * onp.pi * k0)

    def batch_LDOS_measurements(self, batches, k0, alpha, radius, self_interaction=True,
                                self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:888-949."""
        return self._batch_ldos(batches, k0, alpha, radius, self_interaction, self_interaction_type, regularize,
                                discard_absorption, 4.0 * onp.pi * k0)
