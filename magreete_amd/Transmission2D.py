"""Drop-in for the reference module Transmission2D (2D coupled dipoles).

Same module-level names (I, self_interaction_integral_scalar,
self_interaction_integral_vector) and classes (Transmission2D_vector for TE,
Transmission2D_scalar for TM) with the reference's signatures
(Transmission2D.py:10-36, 38-453, 570-953).  Compute runs in
libmagreete_hip on the GPU.
"""
import numpy as onp
import scipy.special
import torch

from . import _lib as L
from . import engine as E
from ._dropin import _Base, _like

I = torch.tensor(onp.identity(2)).reshape(1, 1, 2, 2)  # Transmission2D.py:10


def self_interaction_integral_scalar(k0, radius, self_interaction_type="Rayleigh"):
    """TM self-interaction integral S (Transmission2D.py:12-23)."""
    volume = onp.pi * radius ** 2
    if self_interaction_type == "full":
        return -1 / (k0 * k0) + 0.5j * volume * scipy.special.hankel1(1, k0 * radius) / (k0 * radius)
    if self_interaction_type == "Rayleigh":
        return (-1.0 * radius ** 2 / 4.0) * (2.0 * onp.euler_gamma - 1.0 + 2.0 * onp.log(k0 * radius / 2.0)
                                              - 1j * onp.pi)
    raise NotImplementedError


def self_interaction_integral_vector(k0, radius, self_interaction_type="Rayleigh"):
    """TE self-interaction integral S (Transmission2D.py:25-36)."""
    volume = onp.pi * radius ** 2
    if self_interaction_type == "full":
        return -1 / (k0 * k0) + 0.25j * volume * scipy.special.hankel1(1, k0 * radius) / (k0 * radius)
    if self_interaction_type == "Rayleigh":
        return (-1 / (2.0 * k0 * k0) - (radius ** 2 / 8.0)
                * (2.0 * onp.euler_gamma - 1.0 + 2.0 * onp.log(k0 * radius / 2.0) - 1j * onp.pi))
    raise NotImplementedError


def _hankel1(nu, x, exact):
    if nu not in (0, 1):
        raise SystemExit("torch Hankel function only implemented for orders 0 and 1!")
    return _like(E.hankel1(nu, x, exact), x)


class _Base2D(_Base):
    space = 2

    def _volume(self, radius):
        return onp.pi * radius * radius

    def _wrap(self, r, periodic):
        # Transmission2D.py:74-81: exactly 'y' or exactly 'x'
        if periodic == 'y':
            r[:, :, 1] += 0.5
            r[:, :, 1] %= 1
            r[:, :, 1] -= 0.5
        elif periodic == 'x':
            r[:, :, 0] += 0.5
            r[:, :, 0] %= 1
            r[:, :, 0] -= 0.5

    def torch_hankel1(self, nu, x):
        """H_nu^(1)(x), nu in {0, 1}, evaluated by the library's FP64 Bessel kernels."""
        return _hankel1(nu, x, self.bessel == "exact")

    def _dirs(self, thetas):
        th = onp.asarray(thetas, dtype=onp.float64).reshape(-1)
        return torch.from_numpy(onp.stack([onp.cos(-th), onp.sin(-th)], axis=1))

    def _point_source(self, points, k0, thetas, tm):
        """'point' source: dipole at distance 2 in the frame rotated by -theta
        (Transmission2D.py:145-166, 656-670)."""
        dev = E.device_of(points)
        pts = points.to(device=dev, dtype=torch.float64)
        T = len(thetas)
        intensity = 1.0 * k0 * 2.0 * 2.0 * onp.pi
        loc = torch.tensor([[-2.0, 0.0]], dtype=torch.float64, device=dev)
        out = torch.zeros((pts.shape[0], T) if tm else (pts.shape[0], 2, T), dtype=torch.complex128, device=dev)
        for idx in range(T):
            th = thetas[idx]
            c, s = onp.cos(-th), onp.sin(-th)
            rot = torch.tensor([[c, -s], [s, c]], dtype=torch.float64, device=dev)
            rrot = torch.matmul(rot, pts.T).T.contiguous()
            G, _ = E.assemble(self.kind, rrot, loc, k0, flags=self._bf())
            if tm:
                out[:, idx] = onp.sqrt(intensity) * G.reshape(-1)
            else:
                # float32 as in the reference (numpy scalar times a default-dtype tensor)
                moment = (onp.sqrt(intensity) * torch.tensor([0.0, 1.0], device=dev)).type(torch.complex128)
                f = torch.matmul(G.reshape(-1, 2, 2), moment)
                out[:, :, idx] = torch.matmul(rot.type(torch.complex128).T, f.T).T
        return out


class Transmission2D_vector(_Base2D):
    """TE polarisation, d = 2 (Transmission2D.py:38-453)."""
    kind = "2d_te"
    d = 2

    def _self_int(self, k0, radius, kind):
        return self_interaction_integral_vector(k0, radius, kind)

    def _shift(self, k0, alpha, radius, kind):
        return alpha * k0 * k0 * self_interaction_integral_vector(k0, radius, kind) / (onp.pi * radius * radius)

    def torch_greens(self, r, k0, periodic='', regularize=False, radius=0.0):
        """TE dyadic Green's function of (N, M, 2) difference vectors (Transmission2D.py:63-91)."""
        return self._green_raw(r, k0, periodic, regularize, radius)

    def generate_source(self, points, k0, thetas, w, print_statement=''):
        """Transmission2D.py:93-174."""
        if self.source in ("beam", "plane"):
            d = self._dirs(thetas)
            # u = (sin(-theta), cos(-theta)), stored by the reference in a float32
            # tensor (torch.zeros default dtype, Transmission2D.py:112-116)
            pol = torch.stack([d[:, 1], d[:, 0]], dim=1).float().double()
            return _like(E.sources(2, self.source == "beam", points, d, pol, k0, w), points)
        if self.source == "point":
            return _like(self._point_source(points, k0, thetas, tm=False), points)
        if self.source is None:
            return torch.zeros((points.shape[0], 2, len(thetas)), dtype=torch.complex128)
        raise NotImplementedError

    def G0_vector(self, points, k0, print_statement='', regularize=False, radius=0.0):
        """(2M, 2N) TE Green's matrix; self blocks zeroed when points is None (Transmission2D.py:263-280)."""
        G, ref = self._G0(points, k0, regularize, radius)
        return _like(G, ref)

    def solve(self, k0, alpha, radius, E0, self_interaction=True, self_interaction_type="Rayleigh"):
        """Transmission2D.py:207-231."""
        return self._solve(k0, alpha, radius, E0, self_interaction, self_interaction_type, flat=False)

    def propagate(self, points, Ek, k0, alpha, E0, regularize=False, radius=0.0):
        """Transmission2D.py:176-205 (the fix-up copies the flat (2N, T) row idx)."""
        out, _ = self._propagate(points, Ek, k0, alpha, E0, regularize, radius)
        out = self._fix_rows(out, points, Ek, Ek, regularize, radius)
        return _like(out, Ek)

    def propagate_ss(self, points, k0, alpha, E0_meas, E0_scat, regularize=False, radius=0.0):
        """Transmission2D.py:233-261."""
        out, _ = self._propagate(points, E0_scat.reshape(self.N * 2, -1), k0, alpha, E0_meas, regularize, radius)
        out = self._fix_rows(out, points, E0_meas, E0_meas, regularize, radius)
        return _like(out, E0_meas)

    def mean_DOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                              self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:282-336."""
        return self._mean_dos(measure_points, k0, alpha, radius, self_interaction, self_interaction_type,
                              regularize, discard_absorption, 4.0 * k0 * k0)

    def LDOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                          self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:338-390."""
        return self._ldos(measure_points, k0, alpha, radius, self_interaction, self_interaction_type, regularize,
                          discard_absorption, 4.0 * k0 * k0)

    def batch_LDOS_measurements(self, batches, k0, alpha, radius, self_interaction=True,
                                self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:392-453."""
        return self._batch_ldos(batches, k0, alpha, radius, self_interaction, self_interaction_type, regularize,
                                discard_absorption, 4.0 * k0 * k0)


class Transmission2D_scalar(_Base2D):
    """TM polarisation, d = 1 (Transmission2D.py:570-953)."""
    kind = "2d_tm"
    d = 1

    def _self_int(self, k0, radius, kind):
        return self_interaction_integral_scalar(k0, radius, kind)

    def _shift(self, k0, alpha, radius, kind):
        return alpha * k0 * k0 * self_interaction_integral_scalar(k0, radius, kind) / (onp.pi * radius * radius)

    def torch_greens(self, r, k0, periodic='', regularize=False, radius=0.0):
        """0.25i H0(k0 |r|) of (M, N, 2) difference vectors (Transmission2D.py:595-618)."""
        return self._green_raw(r, k0, periodic, regularize, radius)

    def generate_source(self, points, k0, thetas, w, print_statement=''):
        """Transmission2D.py:620-685."""
        if self.source in ("beam", "plane"):
            return _like(E.sources(2, self.source == "beam", points, self._dirs(thetas), None, k0, w), points)
        if self.source == "point":
            return _like(self._point_source(points, k0, thetas, tm=True), points)
        if self.source is None:
            return torch.zeros((points.shape[0], len(thetas)), dtype=torch.complex128)
        raise NotImplementedError

    def G0(self, points, k0, print_statement='', regularize=False, radius=0.0):
        """(M, N) TM Green's matrix; the self term is not zeroed (Transmission2D.py:777-785)."""
        r = self._r()
        flags = (L.MG_F_REGULARIZE if regularize else 0) | self._bf()
        G, _ = E.assemble(self.kind, points.to(r.device), r, k0, flags=flags, radius=radius)
        return _like(G, points)

    def solve(self, k0, alpha, radius, E0j, self_interaction=True, self_interaction_type="Rayleigh"):
        """Transmission2D.py:718-744."""
        return self._solve(k0, alpha, radius, E0j, self_interaction, self_interaction_type, flat=True)

    def propagate(self, points, Ek, k0, alpha, E0j, regularize=False, radius=0.0):
        """Transmission2D.py:687-716."""
        out, _ = self._propagate(points, Ek, k0, alpha, E0j, regularize, radius)
        out = self._fix_rows(out, points, Ek, Ek, regularize, radius)
        return _like(out, Ek)

    def propagate_ss(self, points, k0, alpha, E0_meas, E0_scat, regularize=False, radius=0.0):
        """Transmission2D.py:746-775."""
        out, _ = self._propagate(points, E0_scat, k0, alpha, E0_meas, regularize, radius)
        out = self._fix_rows(out, points, E0_meas, E0_meas, regularize, radius)
        return _like(out, E0_meas)

    def mean_DOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                              self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:787-841."""
        return self._mean_dos(measure_points, k0, alpha, radius, self_interaction, self_interaction_type,
                              regularize, discard_absorption, 4.0 * k0 * k0)

    def LDOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                          self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:843-893."""
        return self._ldos(measure_points, k0, alpha, radius, self_interaction, self_interaction_type, regularize,
                          discard_absorption, 4.0 * k0 * k0)

    def batch_LDOS_measurements(self, batches, k0, alpha, radius, self_interaction=True,
                                self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:895-953."""
        return self._batch_ldos(batches, k0, alpha, radius, self_interaction, self_interaction_type, regularize,
                                discard_absorption, 4.0 * k0 * k0)
