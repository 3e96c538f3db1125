"""Shared implementation of the four drop-in solver classes.

The classes keep the constructor, attributes (r, N, source) and method
signatures of the reference's Transmission2D_{vector,scalar} and
Transmission3D_{vector,scalar} (Transmission2D.py:38-453, 570-953;
Transmission3D.py:39-439, 565-949).  The compute runs on the GPU through
libmagreete_hip: Green's assembly straight into M, blocked LU + multi-RHS
solve, fused propagation, LDOS through solves against G_p^T.  Results come
back on the device of the caller's tensors (CPU tensors in, CPU tensors out).
Progress prints of the reference are not emitted.
"""
import cmath
import math

import numpy as onp
import torch

from . import _lib as L
from . import engine as E
from . import spectral as S


def _like(t, ref):
    """Move a result to the device of the caller's tensor."""
    if ref is not None and not ref.is_cuda:
        return t.cpu()
    return t


class _Base:
    kind = None          # engine kind name
    d = 1                # field components per site
    space = 2            # coordinates per point
    # 2D Bessel functions: "reference" reproduces torch.special's values as
    # the reference computes them (parity); "exact" evaluates them to ~1e-15.
    bessel = "reference"

    def __init__(self, points, source="beam"):
        self.r = points.reshape(-1, self.space)
        self.N = self.r.shape[0]
        self.source = source
        self._rdev = None

    # -- helpers ------------------------------------------------------------
    def _bf(self):
        """Bessel-mode flag added to every Green's evaluation."""
        if self.bessel not in ("reference", "exact"):
            raise ValueError("bessel must be 'reference' or 'exact'")
        return L.MG_F_BESSEL_EXACT if self.bessel == "exact" else 0

    def _dev(self):
        return E.device_of(self.r)

    def _r(self):
        dev = self._dev()
        if self._rdev is None or self._rdev.device != dev:
            self._rdev = self.r.to(device=dev, dtype=torch.float64).contiguous()
        return self._rdev

    def _volume(self, radius):
        raise NotImplementedError

    def _self_int(self, k0, radius, kind):
        raise NotImplementedError

    def _shift(self, k0, alpha, radius, kind):
        """The scalar subtracted on the diagonal of M by the self-interaction."""
        raise NotImplementedError

    def _system(self, k0, alpha, radius, self_interaction, si_type, want64=False, store128=True):
        """M = -alpha k0^2 G0(None), unit diagonal, optional self-interaction shift,
        assembled in one pass (Transmission3D.py:197-203 and siblings)."""
        r = self._r()
        scale = -alpha * k0 * k0
        diag = 1.0
        bad = False
        if self_interaction:
            c = complex(self._shift(k0, alpha, radius, si_type))
            if not (cmath.isfinite(c)):
                bad = True   # c * eye has NaN off the diagonal: the whole M is NaN
            # the reference multiplies c by torch.eye(n), whose default dtype is
            # float32, so the shift it subtracts is c rounded to complex64
            diag = 1.0 - complex(onp.complex64(c))
        flags = L.MG_F_ZERO_SELF | L.MG_F_DIAG | self._bf()
        M, M64 = E.assemble(self.kind, r, r, k0, scale=scale, flags=flags, diag=diag, want64=want64,
                            out=None if store128 else None)
        if bad:
            if M is not None:
                M.fill_(complex(math.nan, math.nan))
            if M64 is not None:
                M64.fill_(complex(math.nan, math.nan))
        return M, M64

    def _raise_singular(self, info):
        i = int(info.item())
        if i < 0:
            raise RuntimeError("the LU factorisation failed (status %d: a grid barrier of its leaf timed out)" % i)
        if i > 0:
            raise torch.linalg.LinAlgError(
                "linalg.solve: The solver failed because the input matrix is singular "
                "(U[%d,%d] is exactly zero)." % (i, i))

    def _factor(self, k0, alpha, radius, self_interaction, si_type):
        M, _ = self._system(k0, alpha, radius, self_interaction, si_type)
        return E.Factor(M)

    # -- G0 -------------------------------------------------------------------
    def _G0(self, points, k0, regularize=False, radius=0.0):
        r = self._r()
        if points is None:
            G, _ = E.assemble(self.kind, r, r, k0, flags=L.MG_F_ZERO_SELF | self._bf())
            return G, self.r
        flags = (L.MG_F_REGULARIZE if regularize else 0) | self._bf()
        G, _ = E.assemble(self.kind, points.to(r.device), r, k0, flags=flags, radius=radius)
        return G, points

    def _green_raw(self, r, k0, periodic, regularize, radius):
        """Green's function of a (M, N, space) tensor of difference vectors."""
        m, n = r.shape[0], r.shape[1]
        self._wrap(r, periodic)
        diffs = r.reshape(-1, self.space)
        dev = E.device_of(diffs)
        origin = torch.zeros((1, self.space), dtype=torch.float64, device=dev)
        flags = (L.MG_F_REGULARIZE if regularize else 0) | self._bf()
        G, _ = E.assemble(self.kind, diffs.to(dev), origin, k0, flags=flags, radius=radius)
        if self.d == 1:
            return _like(G.reshape(m, n), r)
        return _like(G.reshape(m, n, self.d, self.d), r)

    def _wrap(self, r, periodic):
        raise NotImplementedError

    # -- solve / propagate -----------------------------------------------------
    def _solve(self, k0, alpha, radius, E0, self_interaction, si_type, flat):
        F = self._factor(k0, alpha, radius, self_interaction, si_type)
        dev = F.device
        vec = E0.dim() == 1 and flat
        B = E0.reshape(self.d * self.N, -1).to(device=dev, dtype=torch.complex128).contiguous().clone()
        F.solve(B)
        E.check_barriers(dev)
        self._raise_singular(F.info)
        if vec:
            B = B.reshape(-1)
        return _like(B, E0)

    def _propagate(self, points, src, k0, alpha, E0, regularize, radius):
        dev = self._dev()
        npts = points.shape[0]
        T = E0.shape[-1] if E0.dim() > 1 else 1
        out = E0.to(device=dev, dtype=torch.complex128).reshape(self.d * npts, -1).contiguous().clone()
        Ek = src.to(device=dev, dtype=torch.complex128).reshape(self.d * self.N, -1).contiguous()
        flags = (L.MG_F_REGULARIZE if regularize else 0) | self._bf()
        E.propagate(self.kind, points.to(dev), self._r(), k0, alpha * k0 * k0, Ek, out, flags=flags,
                    radius=radius)
        if self.d == 1:
            return out.reshape(npts, -1) if E0.dim() > 1 else out.reshape(npts), T
        return out.reshape(npts, self.d, -1), T

    def _fix_rows(self, Ek_, points, flat_src, reg_src, regularize, radius):
        """Rows whose first entry is NaN take a source row (Transmission3D.py:166-177)."""
        probe = Ek_[:, 0] if Ek_.dim() == 2 else Ek_[:, 0, 0]
        bad = torch.argwhere(torch.isnan(probe))
        if bad.numel() == 0:
            return Ek_
        r = self.r.to(Ek_.device)
        points = points.to(Ek_.device)
        for j in bad:
            if regularize:
                dist = torch.linalg.norm(r - points[j], axis=-1)
                possible = torch.nonzero(dist <= radius)
                idx = torch.argmin(dist) if possible.shape[0] > 1 else possible
                Ek_[j] = reg_src.to(Ek_.device)[idx]
            else:
                Ek_[j] = flat_src.to(Ek_.device)[torch.nonzero(torch.prod(r - points[j] == 0, axis=-1))]
        return Ek_

    # -- LDOS ------------------------------------------------------------------
    def _ldos_dots(self, F, pts, k0, radius, self_interaction, si_type, regularize, chunk=None):
        """Row sums diag(G_p W G_p^T) for the measurement points, by solving
        M X = G_p^T with the factor F and contracting X with G_p^T column by column."""
        dev = F.device
        pts = pts.to(device=dev, dtype=torch.float64).contiguous()
        fill = self._self_int(k0, radius, si_type) / self._volume(radius) if self_interaction else 0.0
        flags = L.MG_F_NANFILL | (L.MG_F_REGULARIZE if regularize else 0) | self._bf()
        n = self.d * self.N
        if chunk is None:
            # keep G_p^T and X within ~4 GiB per chunk
            chunk = max(1, int((4 << 30) // (2 * 16 * n * self.d)))
        out = []
        for s in range(0, pts.shape[0], chunk):
            p = pts[s:s + chunk]
            GT, _ = E.assemble(self.kind, self._r(), p, k0, flags=flags, fill=fill, radius=radius)
            X = GT.clone()
            F.solve(X)
            out.append(E.coldot(GT, X))
            del GT, X
        return torch.cat(out) if out else torch.zeros(0, dtype=torch.complex128, device=dev)

    # LDOS at points that coincide with scatterers (and regularize off): from
    # the diagonal blocks of M^{-1} instead of solves against G_p^T
    ldos_site_blocks = True

    def _site_index(self, pts):
        """For every point, the index of the scatterer at exactly its position
        (the points whose Green's-function entries are NaN), or -1."""
        lut = {}
        for i, row in enumerate(self.r.detach().cpu().tolist()):
            lut.setdefault(tuple(row), i)
        return onp.array([lut.get(tuple(p), -1) for p in pts.detach().cpu().tolist()], dtype=onp.int64)

    def _site_dots(self, W, k0, alpha, radius, self_interaction, si_type):
        """diag(G_p W G_p^T) at scatterer positions from W's d x d diagonal
        blocks.  There G_p = (dval I - M) / c + F with c = k0^2 alpha, dval
        M's diagonal value and F the blocks (S/V) ones(d, d) that the NaN
        fix-up puts on the self blocks (0 without self-interaction); M is
        complex-symmetric and W M = I, so the block of G_p W G_p^T is
        dval (dval W - I) / c^2 + (F (dval W - I) + (dval W - I) F) / c + F W F."""
        d = self.d
        c = complex(alpha * k0 * k0)
        dval, f = 1.0, 0.0
        if self_interaction:
            sh = complex(self._shift(k0, alpha, radius, si_type))
            dval = 1.0 - complex(onp.complex64(sh))        # as _system assembles M's diagonal
            f = complex(self._self_int(k0, radius, si_type) / self._volume(radius))
        eye = torch.eye(d, dtype=torch.complex128, device=W.device)
        A = dval * W - eye
        D = dval * A / (c * c)
        if f != 0.0:
            Fm = torch.full((d, d), f, dtype=torch.complex128, device=W.device)
            D = D + (Fm @ A + A @ Fm) / c + Fm @ W @ Fm
        return torch.diagonal(D, dim1=1, dim2=2).reshape(-1)

    def _dots(self, F, pts, k0, alpha, radius, self_interaction, si_type, regularize):
        """Rows of diag(G_p W G_p^T) for all points (d per point): probes by
        solves against G_p^T, then the scatterer positions from the diagonal
        blocks of M^{-1}, which consumes F."""
        dev = F.device
        d = self.d
        npts = pts.shape[0]
        site = self._site_index(pts) if (self.ldos_site_blocks and not regularize) else onp.full(npts, -1)
        out = torch.empty((npts, d), dtype=torch.complex128, device=dev)
        probes = onp.nonzero(site < 0)[0]
        sites = onp.nonzero(site >= 0)[0]
        if probes.size:
            sel = torch.from_numpy(probes)
            dots = self._ldos_dots(F, pts[sel], k0, radius, self_interaction, si_type, regularize)
            out[sel.to(dev)] = dots.reshape(-1, d)
        if sites.size:
            W = E.inverse_diag_blocks(F, torch.from_numpy(site[sites]), d)
            out[torch.from_numpy(sites).to(dev)] = self._site_dots(W, k0, alpha, radius, self_interaction,
                                                                  si_type).reshape(-1, d)
        return out.reshape(-1)

    def _ldos_finish(self, dots, npts, pref, alpha, discard_absorption):
        alpha_ = onp.real(alpha) if discard_absorption else alpha
        ldos = torch.imag(dots * (pref * alpha_))
        if self.kind == "2d_tm":
            return ldos
        if self.d == 1:
            return torch.sum(ldos.reshape(npts, -1), 1)
        return torch.sum(ldos.reshape(npts, self.d, -1), 1)

    def _ldos(self, measure_points, k0, alpha, radius, self_interaction, si_type, regularize, discard_absorption,
              pref):
        F = self._factor(k0, alpha, radius, self_interaction, si_type)
        dots = self._dots(F, measure_points, k0, alpha, radius, self_interaction, si_type, regularize)
        E.check_barriers(F.device)
        self._raise_singular(F.info)
        return _like(self._ldos_finish(dots, measure_points.shape[0], pref, alpha, discard_absorption),
                     measure_points)

    def _batch_ldos(self, batches, k0, alpha, radius, self_interaction, si_type, regularize, discard_absorption,
                    pref):
        """All batches' points in one pass: the result is the concatenation of
        the per-batch LDOS (Transmission3D.py:428-439), each point's value
        independent of its batch."""
        F = self._factor(k0, alpha, radius, self_interaction, si_type)
        batches = list(batches)
        if not batches:
            return torch.zeros(0, dtype=torch.float64)
        pts = torch.cat([b.to(torch.float64).reshape(-1, self.space).cpu() for b in batches])
        dots = self._dots(F, pts, k0, alpha, radius, self_interaction, si_type, regularize)
        E.check_barriers(F.device)
        self._raise_singular(F.info)
        return _like(self._ldos_finish(dots, pts.shape[0], pref, alpha, discard_absorption), batches[0])

    def _mean_dos(self, measure_points, k0, alpha, radius, self_interaction, si_type, regularize,
                  discard_absorption, pref):
        F = self._factor(k0, alpha, radius, self_interaction, si_type)
        dots = self._dots(F, measure_points, k0, alpha, radius, self_interaction, si_type, regularize)
        E.check_barriers(F.device)
        self._raise_singular(F.info)
        alpha_ = onp.real(alpha) if discard_absorption else alpha
        dos = dots.sum() / measure_points.shape[0]
        dos = dos * (pref * alpha_)
        return _like(torch.imag(dos), measure_points)

    # -- spectral methods ----------------------------------------------------------
    # The matrices come from the assembly kernel; the eigen-decomposition runs
    # on the host (spectral.py).  Results are on the device of self.r.
    def _on_r(self, t):
        return t.to(self.r.device) if torch.is_tensor(t) else t

    def compute_hamiltonian_DOS(self, k0, deltas, file_name, write_eigenvalues=True):
        """Transmission3D.py:440-461, 951-971; Transmission2D.py:455-475, 955-973."""
        r = self._r()
        G, _ = E.assemble(self.kind, r, r, k0, scale=S.hamiltonian_prefactor(self.kind, k0),
                          flags=L.MG_F_ZERO_SELF | self._bf())
        return self._on_r(S.hamiltonian_dos(G, k0, deltas, file_name, write_eigenvalues))

    def compute_eigenvalues_and_scatterer_LDOS(self, k0, alpha, radius, file_name, self_interaction=True,
                                               self_interaction_type="Rayleigh", write_eigenvalues=True):
        """Transmission3D.py:463-500, 973-1009; Transmission2D.py:477-516, 975-1016."""
        M, _ = self._system(k0, alpha, radius, self_interaction, self_interaction_type)
        return self._on_r(S.eigen_ldos(M, self.kind, k0, alpha, self.N, file_name, write_eigenvalues))

    def compute_eigenmodes_IPR(self, k0, alpha, radius, file_name, self_interaction=True,
                               self_interaction_type="Rayleigh", number_eigenmodes=1, write_eigenvalues=True,
                               sorting_type='IPR'):
        """Transmission3D.py:502-564, 1011-1079; Transmission2D.py:518-569, 1018-1066."""
        M, _ = self._system(k0, alpha, radius, self_interaction, self_interaction_type)
        lam, vecs, ipr = S.eigenmodes_ipr(M, self.kind, k0, alpha, radius, file_name, number_eigenmodes,
                                          write_eigenvalues, sorting_type)
        return self._on_r(lam), self._on_r(vecs), self._on_r(ipr)
