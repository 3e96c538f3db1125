"""Seeded synthetic inputs for the five coupled-dipole workloads.

The sampling formulas restate the reference's helpers:
uniform points in the unit disk (utils.py:85-97), in the unit ball
(utils.py:100-115), the Fibonacci sphere (utils.py:117-136), the triangular
lattice and the circular crop (lattices.py:49-64, 7-10).  The physical
parameters (k0 from the mean nearest-neighbour spacing, the dressed
polarisabilities with radiative damping) are the ones the workloads define.
Every generator takes an explicit seed; seed 0 draws geometry, seed 1 probes.
"""
from dataclasses import dataclass, field
import math

import numpy as np
import torch
from scipy.spatial import cKDTree


# ---------------------------------------------------------------------------
# sampling formulas
# ---------------------------------------------------------------------------
def unit_disk(n, rng):
    """sqrt(U2)(cos 2piU1, sin 2piU1) (utils.py:90-97)."""
    u1 = rng.uniform(size=n)
    u2 = rng.uniform(size=n)
    return np.stack([np.sqrt(u2) * np.cos(2 * np.pi * u1), np.sqrt(u2) * np.sin(2 * np.pi * u1)], axis=1)


def unit_ball(n, rng, dim=3):
    """Gaussian direction over the norm of (normals, sqrt(Exp(1))) (utils.py:106-113)."""
    normals = rng.normal(0.0, 1.0, size=(n, dim))
    exps = np.sqrt(rng.exponential(1.0, size=(n, 1)))
    return normals / np.linalg.norm(np.concatenate([normals, exps], axis=1), axis=-1).reshape(n, 1)


def fibonacci_sphere(samples):
    """Golden-angle spiral from z=1 to z=-1 (utils.py:117-136)."""
    golden = np.pi * (np.sqrt(5.0) - 1.0)
    i = np.arange(samples)
    z = 1 - (i / (samples - 1)) * 2
    radii = np.sqrt(1 - z ** 2)
    th = i * golden
    return np.stack([np.cos(th) * radii, np.sin(th) * radii, z], axis=1).astype(np.float64)


def triangular_lattice(radius, spacing=1.0):
    """Triangular lattice (rows sqrt(3)/2 apart, every other row shifted by 1/2,
    as lattices.py:49-64) with unit spacing, cropped to a disk (lattices.py:7-10)."""
    nx = int(math.ceil(radius / (math.sqrt(3) / 2))) + 2
    ny = int(math.ceil(radius)) + 2
    x = np.arange(-nx, nx + 1, dtype=np.float64) * math.sqrt(3) / 2
    y = np.arange(-ny, ny + 1, dtype=np.float64)
    grid = np.zeros((x.shape[0], y.shape[0], 2))
    grid[:, :, 0] = x.reshape(-1, 1)
    grid[:, :, 1] = y.reshape(1, -1)
    grid[::2, :, 1] += 0.5
    r = grid.reshape(-1, 2) * spacing
    return r[np.linalg.norm(r, axis=-1) <= radius]


def mean_nn_spacing(points):
    tree = cKDTree(points)
    dist, _ = tree.query(points, k=2)
    return float(dist[:, 1].mean())


def polarisation_pairs(u):
    """Two unit polarisations orthogonal to each direction u: p1 = u x z
    (x-hat where u is parallel to z), p2 = u x p1."""
    zhat = np.array([0.0, 0.0, 1.0])
    p1 = np.cross(u, zhat)
    nrm = np.linalg.norm(p1, axis=1)
    bad = nrm < 1e-12
    p1[bad] = np.array([1.0, 0.0, 0.0])
    nrm[bad] = 1.0
    p1 /= nrm[:, None]
    p2 = np.cross(u, p1)
    p2 /= np.linalg.norm(p2, axis=1)[:, None]
    return p1, p2


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
@dataclass
class Workload:
    kind: str                 # '2d_tm', '2d_te', '3d_vector', '3d_scalar'
    points: torch.Tensor      # (N, dim) float64
    k0: float
    alpha: complex
    source: str
    radius: float = 0.0       # scatterer radius (self-interaction off in these workloads)
    self_interaction: bool = False
    thetas: np.ndarray = None            # 2D incidences
    u: torch.Tensor = None               # 3D directions (T, 3)
    p: torch.Tensor = None               # 3D polarisations (T, 3)
    measure: torch.Tensor = None         # measurement points
    probes: torch.Tensor = None          # LDOS probe points
    beam_waist: float = 0.2
    meta: dict = field(default_factory=dict)

    @property
    def d(self):
        return {'2d_tm': 1, '2d_te': 2, '3d_vector': 3, '3d_scalar': 1}[self.kind]

    @property
    def n(self):
        return self.d * self.points.shape[0]

    @property
    def T(self):
        return len(self.thetas) if self.thetas is not None else self.u.shape[0]


def alpha_2d_tm(k0):
    """1/alpha = 0.5 k0^2 - 0.25i k0^2 (Im(1/alpha) = -k0^2/4)."""
    return 1.0 / (0.5 * k0 ** 2 - 0.25j * k0 ** 2)


def alpha_2d_te(k0):
    """1/alpha = 0.5 k0^2 - i k0^2/8 (TE analogue of the radiative damping)."""
    return 1.0 / (0.5 * k0 ** 2 - 1j * k0 ** 2 / 8.0)


def alpha_3d(k0):
    """1/alpha = 0.3 k0^3 - i k0^3/(6 pi)."""
    return 1.0 / (0.3 * k0 ** 3 - 1j * k0 ** 3 / (6 * np.pi))


def cfg_2d_tm(n_points=1000, n_angles=64, n_measure=1024, seed=0):
    """2D TM, N points uniform in the unit disk, k0 * mean NN spacing = 1,
    plane waves at 2 pi j / T, measurement circle of radius 2."""
    rng = np.random.default_rng(seed)
    pts = unit_disk(n_points, rng)
    k0 = 1.0 / mean_nn_spacing(pts)
    th = 2 * np.pi * np.arange(n_angles) / n_angles
    tm = 2 * np.pi * np.arange(n_measure) / n_measure
    meas = 2.0 * np.stack([np.cos(tm), np.sin(tm)], axis=1)
    return Workload('2d_tm', torch.from_numpy(pts), k0, alpha_2d_tm(k0), 'plane', thetas=th,
                    measure=torch.from_numpy(meas), meta=dict(seed=seed))


def cfg_3d_vector(n_points=4096, n_dirs=128, n_measure=2048, seed=0, scalar=False):
    """3D vector, N points uniform in the unit ball, k0 * mean NN spacing = 1,
    n_dirs Fibonacci directions x 2 polarisations (columns ordered (dir, pol)),
    measurement points on a Fibonacci sphere of radius 2."""
    rng = np.random.default_rng(seed)
    pts = unit_ball(n_points, rng)
    k0 = 1.0 / mean_nn_spacing(pts)
    u = fibonacci_sphere(n_dirs)
    p1, p2 = polarisation_pairs(u)
    if scalar:
        uu, pp = u, None
    else:
        uu = np.repeat(u, 2, axis=0)
        pp = np.empty_like(uu)
        pp[0::2] = p1
        pp[1::2] = p2
    meas = 2.0 * fibonacci_sphere(n_measure)
    return Workload('3d_scalar' if scalar else '3d_vector', torch.from_numpy(pts), k0, alpha_3d(k0), 'plane',
                    u=torch.from_numpy(uu), p=None if pp is None else torch.from_numpy(pp),
                    measure=torch.from_numpy(meas), meta=dict(seed=seed))


def cfg_2d_te_lattice(n_sites=20000, k0a=0.6, n_angles=256, n_probes=4096, sigma=0.1, seed=0):
    """2D TE perturbed triangular lattice with unit spacing, cropped to the disk
    holding ~n_sites sites; Gaussian displacement sigma per coordinate;
    n_probes uniform probe points in the disk (seed + 1)."""
    rng = np.random.default_rng(seed)
    rad = math.sqrt(n_sites * math.sqrt(3) / (2 * math.pi))
    lat = triangular_lattice(rad + 1.0)
    lat = lat + rng.normal(0.0, sigma, size=lat.shape)
    lat = lat[np.linalg.norm(lat, axis=-1) <= rad]
    k0 = k0a / 1.0
    th = 2 * np.pi * np.arange(n_angles) / n_angles
    prng = np.random.default_rng(seed + 1)
    probes = rad * unit_disk(n_probes, prng)
    return Workload('2d_te', torch.from_numpy(lat), k0, alpha_2d_te(k0), 'plane', thetas=th,
                    probes=torch.from_numpy(probes), meta=dict(seed=seed, disk_radius=rad, k0a=k0a))


def cfg_ensemble(n_configs=256, n_min=500, n_max=3000, n_dirs=16, seed=0):
    """Batched ensemble: N log-uniform in [n_min, n_max], uniform ball at the
    density of 4096 points in the unit ball, k0 * spacing uniform in [0.5, 1.5],
    n_dirs Fibonacci directions x 2 polarisations."""
    rng = np.random.default_rng(seed)
    rho = 4096 / (4 * np.pi / 3)
    u = fibonacci_sphere(n_dirs)
    p1, p2 = polarisation_pairs(u)
    uu = np.repeat(u, 2, axis=0)
    pp = np.empty_like(uu)
    pp[0::2] = p1
    pp[1::2] = p2
    out = []
    for c in range(n_configs):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        rad = (n / rho / (4 * np.pi / 3)) ** (1.0 / 3.0)
        pts = rad * unit_ball(n, rng)
        spacing = mean_nn_spacing(pts)
        k0 = rng.uniform(0.5, 1.5) / spacing
        out.append(Workload('3d_vector', torch.from_numpy(pts), k0, alpha_3d(k0), 'plane',
                            u=torch.from_numpy(uu), p=torch.from_numpy(pp), meta=dict(seed=seed, index=c)))
    return out
