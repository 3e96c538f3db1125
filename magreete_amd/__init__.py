"""magreete_amd: MI355X-native coupled-dipole multiple-scattering engine.

Drop-in replacements for the reference modules ``Transmission2D`` and
``Transmission3D`` (``magreete_amd.Transmission2D`` /
``magreete_amd.Transmission3D``) whose hot path runs in the HIP library
``libmagreete_hip.so`` (Green's assembly, blocked complex LU and solves,
fused propagation, LDOS), plus the mixed-precision solver, the batched
solver and the serving loop.
"""
from . import _lib  # noqa: F401

__all__ = ["Transmission2D", "Transmission3D", "engine", "workloads"]
