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

    method: 'fp64' (complex128 LU), 'c64' or 'bf16x3' (complex64 LU with
    FP64 GMRES-IR, falling back to the FP64 LU when refinement stalls)."""

    def __init__(self, prob: Problem, method="fp64", tol=1e-12, max_iters=30):
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
                                max_iters=self.max_iters)
        self.info = st.pop("info")
        self.stats = st
        return self.X

    def propagate(self):
        p = self.p
        if self.out is None:
            return None
        self.out.copy_(p.E0m)
        flags = L.MG_F_BESSEL_EXACT if p.bessel_exact else 0
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
