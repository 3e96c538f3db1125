"""One coupled-dipole solve on device-resident inputs, with preallocated buffers.

A solve is the unit of work of the benchmark and of the serving loop:
assemble M = I - k0^2 alpha G, factor, solve all incidence columns, and
propagate to the measurement points, for one (configuration, k0).  The same
steps as the drop-in's solve + propagate (Transmission3D.py:181-206, 147-179),
without host round trips.
"""
from dataclasses import dataclass

import torch

from . import _lib as L
from . import engine as E
from . import refine


@dataclass
class Problem:
    kind: str
    points: torch.Tensor     # (N, S) float64, device
    k0: float
    alpha: complex
    E0s: torch.Tensor        # (d N, T) complex128, device: incident field at the scatterers
    measure: torch.Tensor = None   # (M, S) float64, device
    E0m: torch.Tensor = None       # (d M, T) complex128, device: incident field at the measurement points
    bessel_exact: bool = False
    source: tuple = None           # (beam, dirs, pol, waist): E0m evaluated inside the propagation instead

    @property
    def d(self):
        return E.DIM[self.kind]

    @property
    def n(self):
        return self.d * self.points.shape[0]

    @property
    def T(self):
        return self.E0s.shape[1]


class Solver:
    """Buffers for repeated solves of one problem shape on one device.

    method: 'fp64' (complex128 LU), 'c64', 'bf16x3' or 'bf16x6' (complex64
    LU with FP64 GMRES-IR, falling back to the FP64 LU when refinement
    stalls).  relax: the refinement's late products on split-bf16 MFMA
    (refine.MixedSolver)."""

    def __init__(self, prob: Problem, method="fp64", tol=1e-12, max_iters=30, relax=False):
        self.p = prob
        self.method = method
        dev = prob.points.device
        n, T = prob.n, prob.T
        self.M = torch.empty((n, n), dtype=torch.complex128, device=dev)
        self.M32 = None if method == "fp64" else torch.empty((n, n), dtype=torch.complex64, device=dev)
        self.X = torch.empty((n, T), dtype=torch.complex128, device=dev)
        self.out = None
        if prob.measure is not None:
            self.out = torch.empty((prob.d * prob.measure.shape[0], T), dtype=torch.complex128, device=dev)
        self.tol = tol
        self.max_iters = max_iters
        self.relax = relax
        self.stats = {}

    def assemble(self):
        p = self.p
        flags = L.MG_F_ZERO_SELF | L.MG_F_DIAG | (L.MG_F_BESSEL_EXACT if p.bessel_exact else 0)
        E.assemble(p.kind, p.points, p.points, p.k0, scale=-p.alpha * p.k0 * p.k0, flags=flags, diag=1.0,
                   out=self.M, out64=self.M32)

    def factor_solve(self):
        p = self.p
        if self.method == "fp64":
            f = E.Factor(self.M)
            self.X.copy_(p.E0s)
            f.solve(self.X)
            self.info = f.info
            self.factor = f
            self.stats = {"iters": 0, "fallback": False}
            return self.X
        st = refine.solve_mixed(self.M, self.M32, p.E0s, self.X, prec=self.method, tol=self.tol,
                                max_iters=self.max_iters, relax=self.relax)
        self.info = st.pop("info")
        self.stats = st
        return self.X

    def propagate(self):
        p = self.p
        if self.out is None:
            return None
        flags = L.MG_F_BESSEL_EXACT if p.bessel_exact else 0
        if p.source is not None:
            beam, dirs, pol, waist = p.source
            E.propagate_src(p.kind, p.measure, p.points, p.k0, p.alpha * p.k0 * p.k0, self.X, self.out, beam, dirs,
                            pol, waist, flags=flags)
            return self.out
        self.out.copy_(p.E0m)
        E.propagate(p.kind, p.measure, p.points, p.k0, p.alpha * p.k0 * p.k0, self.X, self.out, flags=flags)
        return self.out

    def run(self, phases=None):
        """One solve.  phases: optional dict that receives per-phase milliseconds
        (HIP events on the current stream; synchronises)."""
        if phases is None:
            self.assemble()
            self.factor_solve()
            return self.propagate()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        self.assemble()
        ev[1].record()
        self.factor_solve()
        ev[2].record()
        out = self.propagate()
        ev[3].record()
        ev[3].synchronize()
        for name, a, b in (("assemble", 0, 1), ("factor_solve", 1, 2), ("propagate", 2, 3)):
            phases[name] = phases.get(name, 0.0) + ev[a].elapsed_time(ev[b])
        return out


def source_spec(kind, w, device=None):
    """(space, beam, dirs, pol) of workload w's incident field with the
    drop-in classes' conventions (2D: dirs (cos(-theta), sin(-theta)), TE
    polarisation (-sin theta, cos theta); 3D vector: p rebuilt through
    atan2/arccos); pol None for scalar fields."""
    beam = w.source == "beam"
    if kind in ("2d_tm", "2d_te"):
        th = torch.as_tensor(w.thetas, dtype=torch.float64).reshape(-1)
        dirs = torch.stack([torch.cos(-th), torch.sin(-th)], dim=1)
        pol = None if kind == "2d_tm" else torch.stack([dirs[:, 1], dirs[:, 0]], dim=1).float().double()
        space = 2
    elif kind == "3d_vector":
        from .Transmission3D import _pvec
        dirs, pol, space = torch.as_tensor(w.u, dtype=torch.float64), _pvec(w.p), 3
    elif kind == "3d_scalar":
        dirs, pol, space = torch.as_tensor(w.u, dtype=torch.float64), None, 3
    else:
        raise NotImplementedError(kind)
    if device is not None:
        dirs = dirs.to(device).contiguous()
        pol = None if pol is None else pol.to(device=device, dtype=torch.float64).contiguous()
    return space, beam, dirs, pol


def incident(kind, points, w):
    """Incident field of workload w at points (device tensor), as (d M, T)
    complex128: plane waves or beams along w's directions."""
    space, beam, dirs, pol = source_spec(kind, w)
    out = E.sources(space, beam, points, dirs, pol, w.k0, w.beam_waist)
    return out.reshape(-1, out.shape[-1]).contiguous()


def problem(w, device, measure=True, fused_source=False):
    """The Problem of a workload on a device (positions and incident fields
    moved there; measurement points when the workload has them and measure).
    fused_source: the incident field at the measurement points is not built;
    the propagation evaluates it in its final pass."""
    pts = w.points.to(device=device, dtype=torch.float64).contiguous()
    E0s = incident(w.kind, pts, w)
    meas = E0m = src = None
    if measure and w.measure is not None:
        meas = w.measure.to(device=device, dtype=torch.float64).contiguous()
        if fused_source:
            _, beam, dirs, pol = source_spec(w.kind, w, device)
            src = (beam, dirs, pol, w.beam_waist)
        else:
            E0m = incident(w.kind, meas, w)
    return Problem(w.kind, pts, w.k0, w.alpha, E0s, meas, E0m, source=src)
