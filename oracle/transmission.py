"""Torch-CPU complex128 restatement of the reference coupled-dipole hot path.

Test infrastructure only (see ``oracle/__init__.py``).  Every function cites
the reference lines it restates.  Tensors are CPU torch tensors: positions
float64, fields and matrices complex128.  Progress prints of the reference are
dropped.
"""
import numpy as onp
import scipy.special
import torch

_I2 = torch.eye(2, dtype=torch.float64).reshape(1, 1, 2, 2)  # Transmission2D.py:10
_I3 = torch.eye(3, dtype=torch.float64).reshape(1, 3, 3)     # Transmission3D.py:11


# ---------------------------------------------------------------------------
# self-interaction integrals (closed forms)
# ---------------------------------------------------------------------------
def self_interaction_integral_scalar_2d(k0, radius, kind="Rayleigh"):
    """Transmission2D.py:12-23."""
    area = onp.pi * radius ** 2
    if kind == "full":
        return -1 / (k0 * k0) + 0.5j * area * scipy.special.hankel1(1, k0 * radius) / (k0 * radius)
    if kind == "Rayleigh":
        return (-1.0 * radius ** 2 / 4.0) * (2.0 * onp.euler_gamma - 1.0
                                              + 2.0 * onp.log(k0 * radius / 2.0) - 1j * onp.pi)
    raise NotImplementedError


def self_interaction_integral_vector_2d(k0, radius, kind="Rayleigh"):
    """Transmission2D.py:25-36."""
    area = onp.pi * radius ** 2
    if kind == "full":
        return -1 / (k0 * k0) + 0.25j * area * scipy.special.hankel1(1, k0 * radius) / (k0 * radius)
    if kind == "Rayleigh":
        return (-1 / (2.0 * k0 * k0) - (radius ** 2 / 8.0)
                * (2.0 * onp.euler_gamma - 1.0 + 2.0 * onp.log(k0 * radius / 2.0) - 1j * onp.pi))
    raise NotImplementedError


def self_interaction_integral_vector_3d(k0, radius, kind="Rayleigh"):
    """Transmission3D.py:13-24."""
    vol = 4.0 * onp.pi * radius ** 3 / 3.0
    if kind == "full":
        return (1.0 / k0 ** 2) * ((2.0 / 3.0) * onp.exp(1j * k0 * radius) * (1 - 1j * k0 * radius) - 1.0)
    if kind == "Rayleigh":
        return -1.0 / (3.0 * k0 ** 2) + radius ** 2 / 3.0 + 1j * k0 * vol / (6.0 * onp.pi)
    raise NotImplementedError


def self_interaction_integral_scalar_3d(k0, radius, kind="Rayleigh"):
    """Transmission3D.py:26-37."""
    vol = 4.0 * onp.pi * radius ** 3 / 3.0
    if kind == "full":
        return (1.0 / k0 ** 2) * (onp.exp(1j * k0 * radius) * (1 - 1j * k0 * radius) - 1.0)
    if kind == "Rayleigh":
        return radius ** 2 / 2.0 + 1j * k0 * vol / (4.0 * onp.pi)
    raise NotImplementedError


# ---------------------------------------------------------------------------
# special functions and Green's functions
# ---------------------------------------------------------------------------
def hankel1(nu, x):
    """H_nu^(1)(x) = J_nu + i Y_nu from torch.special, nu in {0, 1} (Transmission2D.py:46-61)."""
    if nu == 0:
        return torch.special.bessel_j0(x) + 1j * torch.special.bessel_y0(x)
    if nu == 1:
        return torch.special.bessel_j1(x) + 1j * torch.special.bessel_y1(x)
    raise SystemExit("torch Hankel function only implemented for orders 0 and 1!")


def _wrap_2d(r, periodic):
    # Transmission2D.py:74-81 / 605-612: only one of 'y' or 'x' is honoured
    if periodic == 'y':
        r[:, :, 1] += 0.5
        r[:, :, 1] %= 1
        r[:, :, 1] -= 0.5
    elif periodic == 'x':
        r[:, :, 0] += 0.5
        r[:, :, 0] %= 1
        r[:, :, 0] -= 0.5


def _wrap_3d(r, periodic):
    # Transmission3D.py:58-69: any of x, y, z
    for axis, name in enumerate("xyz"):
        if name in periodic:
            r[:, :, axis] += 0.5
            r[:, :, axis] %= 1
            r[:, :, axis] -= 0.5


def greens_2d_tm(r, k0, periodic='', regularize=False, radius=0.0):
    """0.25j*H0(k0|r|), (M,N) (Transmission2D.py:595-618)."""
    _wrap_2d(r, periodic)
    R = torch.linalg.norm(r, axis=-1)
    if regularize:
        R = torch.where(R < radius, 0.0, R)
    return 0.25j * hankel1(0, R * k0)


def greens_2d_te(r, k0, periodic='', regularize=False, radius=0.0):
    """TE dyadic, (M,N,2,2) (Transmission2D.py:63-91).

    R̂R̂ is formed before R is scaled by k0, and the regularisation compares
    the already scaled k0*R against radius (Transmission2D.py:83-89)."""
    m, n = r.shape[0], r.shape[1]
    _wrap_2d(r, periodic)
    R = torch.linalg.norm(r, axis=-1)
    RxR = r[:, :, 0:2].reshape(m, n, 1, 2) * r[:, :, 0:2].reshape(m, n, 2, 1)
    RxR /= (R * R).reshape(m, n, 1, 1)
    R *= k0
    if regularize:
        R = torch.where(R < radius, 0.0, R)
    h0 = hankel1(0, R).reshape(m, n, 1, 1)
    h1r = (hankel1(1, R) / R).reshape(m, n, 1, 1)
    return 0.25j * ((_I2 - RxR) * h0 - (_I2 - 2 * RxR) * h1r)


def greens_3d_vector(r, k0, periodic='', regularize=False, radius=0.0):
    """3D dyadic, (M,N,3,3) (Transmission3D.py:47-77)."""
    m, n = r.shape[0], r.shape[1]
    _wrap_3d(r, periodic)
    R = torch.linalg.norm(r, axis=-1).reshape(m, n, 1, 1)
    RxR = r.reshape(m, n, 1, 3) * r.reshape(m, n, 3, 1)
    RxR /= R * R
    if regularize:
        R = torch.where(R < radius, 0.0, R)
    return (_I3 - RxR - (_I3 - 3 * RxR) * (1 / (1j * k0 * R) + (k0 * R) ** -2)) \
        * torch.exp(1j * k0 * R) / (4 * onp.pi * R)


def greens_3d_scalar(r, k0, periodic='', regularize=False, radius=0.0):
    """exp(ik0R)/(4 pi R), (M,N) (Transmission3D.py:573-601)."""
    m, n = r.shape[0], r.shape[1]
    _wrap_3d(r, periodic)
    R = torch.linalg.norm(r, axis=-1).reshape(m, n)
    if regularize:
        R = torch.where(R < radius, 0.0, R)
    return torch.exp(1j * k0 * R) / (4 * onp.pi * R)


# ---------------------------------------------------------------------------
# shared pieces of the solver classes
# ---------------------------------------------------------------------------
def _to_matrix(G):
    # (M,N,d,d) -> (d*M, d*N): row d*m+a, column d*j+b (Transmission3D.py:267)
    d = G.shape[-1]
    return torch.transpose(G, 1, 2).reshape(d * G.shape[0], d * G.shape[1]).to(torch.complex128)


def _system_matrix(G, alpha, k0, self_int, area_or_vol, self_interaction, order2d):
    """M = -alpha k0^2 G, unit diagonal, optional self-interaction shift.

    Transmission2D.py:222-228, 733-739; Transmission3D.py:197-203, 709-715.
    The 2D files write the shift as alpha*k0*k0*S/V, the 3D files as
    k0**2*alpha*S/V; both orders are kept."""
    M = -alpha * k0 * k0 * G
    M.fill_diagonal_(1)
    if self_interaction:
        n = M.shape[0]
        if order2d:
            M -= alpha * k0 * k0 * self_int / area_or_vol * torch.eye(n)
        else:
            M -= k0 ** 2 * alpha * self_int / area_or_vol * torch.eye(n)
    return M


def _fix_nan_entries(G_measure, fill):
    """Every NaN entry of G_measure -> 0 (+fill) (e.g. Transmission3D.py:353-362)."""
    for j in torch.argwhere(torch.isnan(G_measure)):
        G_measure[j[0]][j[1]] = 0
        if fill is not None:
            G_measure[j[0]][j[1]] += fill
    return G_measure


def _ldos_from(G_measure, W, pref, alpha, discard_absorption, rows_per_point):
    """Im(pref*alpha' diag(G W G^T)) summed over the d rows of each point.

    Transmission3D.py:365-374 (einsum 'ij,ji->i', non-conjugate transpose)."""
    ldos = torch.einsum('ij, ji->i', torch.matmul(G_measure, W), G_measure.t())
    alpha_ = onp.real(alpha) if discard_absorption else alpha
    ldos *= pref * alpha_
    ldos = torch.imag(ldos)
    if rows_per_point is None:          # 2D scalar: (M,) as is
        return ldos
    npts = ldos.shape[0] // rows_per_point
    if rows_per_point == 1:             # 3D scalar: reshape(M,-1).sum(1) -> (M,)
        return torch.sum(ldos.reshape(npts, -1), 1)
    return torch.sum(ldos.reshape(npts, rows_per_point, -1), 1)   # vector: (M, 1)


def _mean_dos_from(G_measure, W, npoints, pref, alpha, discard_absorption):
    """Im(pref*alpha' sum((G^T G) o W))/Np (Transmission3D.py:310-321)."""
    dos = (torch.matmul(G_measure.t(), G_measure) * W).sum() / npoints
    alpha_ = onp.real(alpha) if discard_absorption else alpha
    dos *= pref * alpha_
    return torch.imag(dos)


def _fix_nan_rows(E, points, r, flat_source, reg_source, regularize, radius, first):
    """Rows of E whose first entry is NaN take their value from a source row.

    Restates the loops at Transmission3D.py:166-177 (propagate) and 226-236
    (propagate_ss), Transmission2D.py:193-203, 249-259, 704-714, 763-773.
    ``first`` selects E[:, 0] (scalar) or E[:, 0, 0] (vector) for the NaN test;
    ``flat_source`` / ``reg_source`` are the tensors indexed by the scatterer
    index in the exact and regularised branches."""
    probe = E[:, 0] if first == 1 else E[:, 0, 0]
    for j in torch.argwhere(torch.isnan(probe)):
        if regularize:
            dist = torch.linalg.norm(r - points[j], axis=-1)
            possible = torch.nonzero(dist <= radius)
            if possible.shape[0] > 1:
                idx = torch.argmin(dist)
            else:
                idx = possible
            E[j] = reg_source[idx]
        else:
            E[j] = flat_source[torch.nonzero(torch.prod(r - points[j] == 0, axis=-1))]
    return E


# ---------------------------------------------------------------------------
# 2D vector (TE), Transmission2D.py:38-453
# ---------------------------------------------------------------------------
class Transmission2D_vector:
    d = 2

    def __init__(self, points, source="beam"):
        self.r = points.reshape(-1, 2)
        self.N = self.r.shape[0]
        self.source = source

    def torch_greens(self, r, k0, periodic='', regularize=False, radius=0.0):
        return greens_2d_te(r, k0, periodic, regularize, radius)

    def G0_vector(self, points, k0, print_statement='', regularize=False, radius=0.0):
        """Transmission2D.py:263-280."""
        pts = self.r if points is None else points
        G = greens_2d_te(pts.reshape(-1, 1, 2) - self.r.reshape(1, -1, 2), k0,
                         regularize=regularize, radius=radius)
        if points is None:
            for i in range(self.N):
                G[i, i, :, :] = 0
        return _to_matrix(G)

    def generate_source(self, points, k0, thetas, w, print_statement=''):
        """Transmission2D.py:93-174."""
        npts = points.shape[0]
        if self.source in ("beam", "plane"):
            E0 = torch.zeros((npts, len(thetas)), dtype=torch.complex128)
            u = torch.zeros((2, len(thetas)))
            for idx in range(len(thetas)):
                theta = thetas[idx]
                c, s = onp.cos(-theta), onp.sin(-theta)
                u[:, idx] = torch.tensor([s, c])
                rot = torch.tensor([[c, -s], [s, c]])
                rrot = torch.matmul(rot, points.T).T
                if self.source == "beam":
                    a = 2 * rrot[:, 1] / (w * w * k0)
                    E0[:, idx] = torch.exp(1j * rrot[:, 0] * k0 - (rrot[:, 1] ** 2 / (w * w * (1 + 1j * a)))) \
                        / torch.sqrt(1 + 1j * a)
                else:
                    E0[:, idx] = torch.exp(1j * rrot[:, 0] * k0)
            return E0.reshape(npts, 1, len(thetas)) * u
        if self.source == 'point':
            dist = 2.0
            intensity = 1.0 * k0 * dist * 2.0 * onp.pi
            E0 = torch.zeros((npts, 2, len(thetas)), dtype=torch.complex128)
            for idx in range(len(thetas)):
                theta = thetas[idx]
                c, s = onp.cos(-theta), onp.sin(-theta)
                rot = torch.tensor([[c, -s], [s, c]])
                rrot = torch.matmul(rot, points.T).T
                loc = dist * torch.tensor([-1.0, 0.0])
                moment = onp.sqrt(intensity) * torch.tensor([0.0, 1.0])
                E0[:, :, idx] = torch.matmul(greens_2d_te(rrot.reshape(-1, 1, 2) - loc.reshape(1, -1, 2), k0),
                                             moment.type(torch.complex128)).squeeze()
                E0[:, :, idx] = torch.matmul(rot.type(torch.complex128).T, E0[:, :, idx].T).T
            return E0
        if self.source is None:
            return torch.zeros((npts, 2, len(thetas)), dtype=torch.complex128)
        raise NotImplementedError

    def solve(self, k0, alpha, radius, E0, self_interaction=True, self_interaction_type="Rayleigh"):
        """Transmission2D.py:207-231."""
        S = self_interaction_integral_vector_2d(k0, radius, self_interaction_type) if self_interaction else 0
        M = _system_matrix(self.G0_vector(None, k0), alpha, k0, S, onp.pi * radius * radius,
                           self_interaction, order2d=True)
        return torch.linalg.solve(M, E0.reshape(2 * self.N, -1))

    def propagate(self, points, Ek, k0, alpha, E0, regularize=False, radius=0.0):
        """Transmission2D.py:176-205.  The fix-up copies the flat (2N,T) row idx."""
        E = torch.matmul(alpha * k0 * k0 * self.G0_vector(points, k0, regularize=regularize, radius=radius),
                         Ek).reshape(points.shape[0], 2, -1) + E0
        return _fix_nan_rows(E, points, self.r, Ek, Ek, regularize, radius, first=2)

    def propagate_ss(self, points, k0, alpha, E0_meas, E0_scat, regularize=False, radius=0.0):
        """Transmission2D.py:233-261."""
        E = torch.matmul(alpha * k0 * k0 * self.G0_vector(points, k0, regularize=regularize, radius=radius),
                         E0_scat.reshape(self.r.shape[0] * 2, -1)).reshape(points.shape[0], 2, -1) + E0_meas
        return _fix_nan_rows(E, points, self.r, E0_meas, E0_meas, regularize, radius, first=2)

    def _inverse(self, k0, alpha, radius, self_interaction, kind):
        S = self_interaction_integral_vector_2d(k0, radius, kind) if self_interaction else 0
        M = _system_matrix(self.G0_vector(None, k0), alpha, k0, S, onp.pi * radius * radius,
                           self_interaction, order2d=True)
        return torch.linalg.solve(M, torch.eye(len(M), dtype=torch.complex128))

    def _measure(self, pts, k0, radius, self_interaction, kind, regularize):
        G = self.G0_vector(pts, k0, regularize=regularize, radius=radius)
        fill = self_interaction_integral_vector_2d(k0, radius, kind) / (onp.pi * radius * radius) \
            if self_interaction else None
        return _fix_nan_entries(G, fill)

    def mean_DOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                              self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:282-336."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        G = self._measure(measure_points, k0, radius, self_interaction, self_interaction_type, regularize)
        return _mean_dos_from(G, W, measure_points.shape[0], 4.0 * k0 * k0, alpha, discard_absorption)

    def LDOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                          self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:338-390."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        G = self._measure(measure_points, k0, radius, self_interaction, self_interaction_type, regularize)
        return _ldos_from(G, W, 4.0 * k0 * k0, alpha, discard_absorption, 2)

    def batch_LDOS_measurements(self, batches, k0, alpha, radius, self_interaction=True,
                                self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:392-453."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        out = []
        for b in range(len(batches)):
            G = self._measure(batches[b], k0, radius, self_interaction, self_interaction_type, regularize)
            out.append(_ldos_from(G, W, 4.0 * k0 * k0, alpha, discard_absorption, 2))
        return torch.cat(out)


# ---------------------------------------------------------------------------
# 2D scalar (TM), Transmission2D.py:570-953
# ---------------------------------------------------------------------------
class Transmission2D_scalar:
    d = 1

    def __init__(self, points, source="beam"):
        self.r = points.reshape(-1, 2)
        self.N = self.r.shape[0]
        self.source = source

    def torch_greens(self, r, k0, periodic='', regularize=False, radius=0.0):
        return greens_2d_tm(r, k0, periodic, regularize, radius)

    def G0(self, points, k0, print_statement='', regularize=False, radius=0.0):
        """Transmission2D.py:777-785: no zeroing of the self term."""
        return greens_2d_tm(points.reshape(-1, 1, 2) - self.r.reshape(1, -1, 2), k0,
                            regularize=regularize, radius=radius)

    def generate_source(self, points, k0, thetas, w, print_statement=''):
        """Transmission2D.py:620-685."""
        npts = points.shape[0]
        E0 = torch.zeros((npts, len(thetas)), dtype=torch.complex128)
        if self.source is None:
            return E0
        if self.source not in ("beam", "plane", "point"):
            raise NotImplementedError
        for idx in range(len(thetas)):
            theta = thetas[idx]
            c, s = onp.cos(-theta), onp.sin(-theta)
            rot = torch.tensor([[c, -s], [s, c]])
            rrot = torch.matmul(rot, points.T).T
            if self.source == "beam":
                a = 2 * rrot[:, 1] / (w * w * k0)
                E0[:, idx] = torch.exp(1j * rrot[:, 0] * k0 - (rrot[:, 1] ** 2 / (w * w * (1 + 1j * a)))) \
                    / torch.sqrt(1 + 1j * a)
            elif self.source == "plane":
                E0[:, idx] = torch.exp(1j * rrot[:, 0] * k0)
            else:
                intensity = 1.0 * k0 * 2.0 * 2.0 * onp.pi
                loc = 2.0 * torch.tensor([-1.0, 0.0])
                E0[:, idx] = onp.sqrt(intensity) * greens_2d_tm(rrot.reshape(-1, 1, 2) - loc.reshape(1, -1, 2),
                                                                k0).reshape(npts)
        return E0

    def solve(self, k0, alpha, radius, E0j, self_interaction=True, self_interaction_type="Rayleigh"):
        """Transmission2D.py:718-744."""
        S = self_interaction_integral_scalar_2d(k0, radius, self_interaction_type) if self_interaction else 0
        M = _system_matrix(self.G0(self.r, k0), alpha, k0, S, onp.pi * radius * radius,
                           self_interaction, order2d=True)
        return torch.linalg.solve(M, E0j)

    def propagate(self, points, Ek, k0, alpha, E0j, regularize=False, radius=0.0):
        """Transmission2D.py:687-716."""
        E = torch.matmul(alpha * k0 * k0 * self.G0(points, k0, regularize=regularize, radius=radius), Ek) + E0j
        return _fix_nan_rows(E, points, self.r, Ek, Ek, regularize, radius, first=1)

    def propagate_ss(self, points, k0, alpha, E0_meas, E0_scat, regularize=False, radius=0.0):
        """Transmission2D.py:746-775."""
        points = torch.tensor(points)
        E = torch.matmul(alpha * k0 * k0 * self.G0(points, k0, regularize=regularize, radius=radius),
                         E0_scat) + E0_meas
        return _fix_nan_rows(E, points, self.r, E0_meas, E0_meas, regularize, radius, first=1)

    def _inverse(self, k0, alpha, radius, self_interaction, kind):
        S = self_interaction_integral_scalar_2d(k0, radius, kind) if self_interaction else 0
        M = _system_matrix(self.G0(self.r, k0), alpha, k0, S, onp.pi * radius * radius,
                           self_interaction, order2d=True)
        return torch.linalg.solve(M, torch.eye(len(M), dtype=torch.complex128))

    def _measure(self, pts, k0, radius, self_interaction, kind, regularize):
        G = self.G0(pts, k0, regularize=regularize, radius=radius)
        fill = self_interaction_integral_scalar_2d(k0, radius, kind) / (onp.pi * radius * radius) \
            if self_interaction else None
        return _fix_nan_entries(G, fill)

    def mean_DOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                              self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:787-841."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        G = self._measure(measure_points, k0, radius, self_interaction, self_interaction_type, regularize)
        return _mean_dos_from(G, W, measure_points.shape[0], 4.0 * k0 * k0, alpha, discard_absorption)

    def LDOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                          self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:843-893."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        G = self._measure(measure_points, k0, radius, self_interaction, self_interaction_type, regularize)
        return _ldos_from(G, W, 4.0 * k0 * k0, alpha, discard_absorption, None)

    def batch_LDOS_measurements(self, batches, k0, alpha, radius, self_interaction=True,
                                self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission2D.py:895-953."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        out = []
        for b in range(len(batches)):
            G = self._measure(batches[b], k0, radius, self_interaction, self_interaction_type, regularize)
            out.append(_ldos_from(G, W, 4.0 * k0 * k0, alpha, discard_absorption, None))
        return torch.cat(out)


# ---------------------------------------------------------------------------
# 3D vector, Transmission3D.py:39-439
# ---------------------------------------------------------------------------
def _pvec(p):
    # Transmission3D.py:102-104: polarisation rebuilt through atan2/arccos
    phi = torch.arctan2(p[:, 1], p[:, 0])
    theta = torch.arccos(p[:, 2])
    return torch.stack([torch.sin(theta) * torch.cos(phi), torch.sin(theta) * torch.sin(phi), p[:, 2]])


def _beam_3d(points, k0, u, w):
    # Transmission3D.py:96-100
    rpara = torch.matmul(points, u.T)
    rperp = torch.linalg.norm(points.reshape(-1, 3, 1)
                              - rpara.reshape(points.shape[0], 1, u.shape[0]) * u.T.reshape(1, 3, -1), axis=1)
    a = 2 * rperp / (w * w * k0)
    return torch.exp(1j * rpara * k0 - (rperp ** 2 / (w * w * (1 + 1j * a)))) / torch.sqrt(1 + 1j * a)


class Transmission3D_vector:
    d = 3

    def __init__(self, points, source="beam"):
        self.r = points.reshape(-1, 3)
        self.N = self.r.shape[0]
        self.source = source

    def greens(self, r, k0, periodic='', regularize=False, radius=0.0):
        return greens_3d_vector(r, k0, periodic, regularize, radius)

    def G0(self, points, k0, print_statement='', regularize=False, radius=0.0):
        """Transmission3D.py:240-268."""
        pts = self.r if points is None else points
        G = greens_3d_vector(pts.reshape(-1, 1, 3) - self.r.reshape(1, -1, 3), k0,
                             regularize=regularize, radius=radius)
        if points is None:
            for i in range(self.N):
                G[i, i, :, :] = 0
        return _to_matrix(G)

    def generate_source(self, points, k0, u, p, w, print_statement=''):
        """Transmission3D.py:79-145.

        'point' calls an undefined method in the reference (Transmission3D.py:136);
        here it is the field of dipoles p placed at -2u, G(r - r_s) p, scaled
        as the reference's other point sources."""
        if self.source in ('beam', 'plane'):
            if self.source == 'beam':
                E = _beam_3d(points, k0, u, w)
            else:
                E = torch.exp(1j * torch.matmul(points, u.T) * k0)
            return E.reshape(points.shape[0], 1, -1) * _pvec(p).reshape(1, 3, -1)
        if self.source == 'point':
            intensity = 1.0 * (k0 * 2.0) ** 2 * 4.0 * onp.pi
            loc = 2.0 * (-u)
            moment = onp.sqrt(intensity) * p
            G = greens_3d_vector(points.reshape(-1, 1, 3) - loc.reshape(1, -1, 3), k0)  # (M, Ndirs, 3, 3)
            return torch.einsum('mtab,tb->mat', G, moment.type(torch.complex128))
        if self.source is None:
            return torch.zeros((points.shape[0], u.shape[1], u.shape[0]), dtype=torch.complex128)
        raise NotImplementedError

    def solve(self, k0, alpha, radius, E0j, self_interaction=True, self_interaction_type="Rayleigh"):
        """Transmission3D.py:181-206."""
        S = self_interaction_integral_vector_3d(k0, radius, self_interaction_type) if self_interaction else 0
        M = _system_matrix(self.G0(None, k0), alpha, k0, S, 4 * onp.pi * (radius ** 3) / 3,
                           self_interaction, order2d=False)
        return torch.linalg.solve(M, E0j.reshape(3 * self.N, -1))

    def propagate(self, points, Ek, k0, alpha, E0j, regularize=False, radius=0.0):
        """Transmission3D.py:147-179."""
        E = torch.matmul(alpha * k0 * k0 * self.G0(points, k0, regularize=regularize, radius=radius),
                         Ek).reshape(points.shape[0], 3, -1) + E0j
        src = Ek.reshape(self.r.shape[0], 3, -1)
        return _fix_nan_rows(E, points, self.r, src, src, regularize, radius, first=2)

    def propagate_ss(self, points, k0, alpha, E0_meas, E0_scat, regularize=False, radius=0.0):
        """Transmission3D.py:208-238."""
        points = torch.tensor(points)
        E = torch.matmul(alpha * k0 * k0 * self.G0(points, k0, regularize=regularize, radius=radius),
                         E0_scat).reshape(points.shape[0], 3, -1) + E0_meas
        return _fix_nan_rows(E, points, self.r, E0_meas, E0_meas, regularize, radius, first=2)

    def _inverse(self, k0, alpha, radius, self_interaction, kind):
        S = self_interaction_integral_vector_3d(k0, radius, kind) if self_interaction else 0
        M = _system_matrix(self.G0(None, k0), alpha, k0, S, 4 * onp.pi * (radius ** 3) / 3,
                           self_interaction, order2d=False)
        return torch.linalg.solve(M, torch.eye(len(M), dtype=torch.complex128))

    def _measure(self, pts, k0, radius, self_interaction, kind, regularize):
        G = self.G0(pts, k0, regularize=regularize, radius=radius)
        fill = self_interaction_integral_vector_3d(k0, radius, kind) / (4 * onp.pi * (radius ** 3) / 3) \
            if self_interaction else None
        return _fix_nan_entries(G, fill)

    def mean_DOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                              self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:270-323."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        G = self._measure(measure_points, k0, radius, self_interaction, self_interaction_type, regularize)
        return _mean_dos_from(G, W, measure_points.shape[0], 2.0 * onp.pi * k0, alpha, discard_absorption)

    def LDOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                          self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:325-377."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        G = self._measure(measure_points, k0, radius, self_interaction, self_interaction_type, regularize)
        return _ldos_from(G, W, 2.0 * onp.pi * k0, alpha, discard_absorption, 3)

    def batch_LDOS_measurements(self, batches, k0, alpha, radius, self_interaction=True,
                                self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:379-439."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        out = []
        for b in range(len(batches)):
            G = self._measure(batches[b], k0, radius, self_interaction, self_interaction_type, regularize)
            out.append(_ldos_from(G, W, 2.0 * onp.pi * k0, alpha, discard_absorption, 3))
        return torch.cat(out)


# ---------------------------------------------------------------------------
# 3D scalar, Transmission3D.py:565-949
# ---------------------------------------------------------------------------
class Transmission3D_scalar:
    d = 1

    def __init__(self, points, source="beam"):
        self.r = points.reshape(-1, 3)
        self.N = self.r.shape[0]
        self.source = source

    def greens(self, r, k0, periodic='', regularize=False, radius=0.0):
        return greens_3d_scalar(r, k0, periodic, regularize, radius)

    def G0(self, points, k0, print_statement='', regularize=False, radius=0.0):
        """Transmission3D.py:751-777."""
        pts = self.r if points is None else points
        G = greens_3d_scalar(pts.reshape(-1, 1, 3) - self.r.reshape(1, -1, 3), k0,
                             regularize=regularize, radius=radius)
        if points is None:
            for i in range(self.N):
                G[i, i] = 0
        return G

    def generate_source(self, points, k0, u, w, print_statement=''):
        """Transmission3D.py:603-658.

        'point' fails in the reference (a .type() call on a numpy scalar,
        Transmission3D.py:651); here it is sqrt(I) G(r - r_s)."""
        if self.source == 'beam':
            return _beam_3d(points, k0, u, w)
        if self.source == 'plane':
            return torch.exp(1j * torch.matmul(points, u.T) * k0)
        if self.source == 'point':
            intensity = 1.0 * (k0 * 2.0) ** 2 * 4.0 * onp.pi
            loc = 2.0 * (-u)
            return greens_3d_scalar(points.reshape(-1, 1, 3) - loc.reshape(1, -1, 3), k0) * onp.sqrt(intensity)
        if self.source is None:
            return torch.zeros((points.shape[0], u.shape[1]), dtype=torch.complex128)
        raise NotImplementedError

    def solve(self, k0, alpha, radius, E0j, self_interaction=True, self_interaction_type="Rayleigh"):
        """Transmission3D.py:693-718."""
        S = self_interaction_integral_scalar_3d(k0, radius, self_interaction_type) if self_interaction else 0
        M = _system_matrix(self.G0(None, k0), alpha, k0, S, 4 * onp.pi * (radius ** 3) / 3,
                           self_interaction, order2d=False)
        return torch.linalg.solve(M, E0j.reshape(self.N, -1))

    def propagate(self, points, Ek, k0, alpha, E0j, regularize=False, radius=0.0):
        """Transmission3D.py:660-691."""
        E = torch.matmul(alpha * k0 * k0 * self.G0(points, k0, regularize=regularize, radius=radius),
                         Ek).reshape(points.shape[0], E0j.shape[1]) + E0j
        src = Ek.reshape(self.r.shape[0], -1)
        return _fix_nan_rows(E, points, self.r, src, src, regularize, radius, first=1)

    def propagate_ss(self, points, k0, alpha, E0_meas, E0_scat, regularize=False, radius=0.0):
        """Transmission3D.py:720-749."""
        points = torch.tensor(points)
        E = torch.matmul(alpha * k0 * k0 * self.G0(points, k0, regularize=regularize, radius=radius),
                         E0_scat).reshape(points.shape[0], -1) + E0_meas
        return _fix_nan_rows(E, points, self.r, E0_meas, E0_meas, regularize, radius, first=1)

    def _inverse(self, k0, alpha, radius, self_interaction, kind):
        S = self_interaction_integral_scalar_3d(k0, radius, kind) if self_interaction else 0
        M = _system_matrix(self.G0(None, k0), alpha, k0, S, 4 * onp.pi * (radius ** 3) / 3,
                           self_interaction, order2d=False)
        return torch.linalg.solve(M, torch.eye(len(M), dtype=torch.complex128))

    def _measure(self, pts, k0, radius, self_interaction, kind, regularize):
        G = self.G0(pts, k0, regularize=regularize, radius=radius)
        fill = self_interaction_integral_scalar_3d(k0, radius, kind) / (4 * onp.pi * (radius ** 3) / 3) \
            if self_interaction else None
        return _fix_nan_entries(G, fill)

    def mean_DOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                              self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:779-832."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        G = self._measure(measure_points, k0, radius, self_interaction, self_interaction_type, regularize)
        return _mean_dos_from(G, W, measure_points.shape[0], 4.0 * onp.pi * k0, alpha, discard_absorption)

    def LDOS_measurements(self, measure_points, k0, alpha, radius, self_interaction=True,
                          self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:834-886."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        G = self._measure(measure_points, k0, radius, self_interaction, self_interaction_type, regularize)
        return _ldos_from(G, W, 4.0 * onp.pi * k0, alpha, discard_absorption, 1)

    def batch_LDOS_measurements(self, batches, k0, alpha, radius, self_interaction=True,
                                self_interaction_type="Rayleigh", regularize=False, discard_absorption=False):
        """Transmission3D.py:888-949."""
        W = self._inverse(k0, alpha, radius, self_interaction, self_interaction_type)
        out = []
        for b in range(len(batches)):
            G = self._measure(batches[b], k0, radius, self_interaction, self_interaction_type, regularize)
            out.append(_ldos_from(G, W, 4.0 * onp.pi * k0, alpha, discard_absorption, 1))
        return torch.cat(out)


# ---------------------------------------------------------------------------
# spectral methods (eigenvalues of G or M, computed with torch's LAPACK)
# ---------------------------------------------------------------------------
# per class: Hamiltonian prefactor of G (as a function of k0), the LDOS
# prefactor's numerator and power of k0, the CSV suffix of the deltas
SPECTRAL = {
    "3d_vector": dict(ham=lambda k0: 6 * onp.pi / k0, ldos=(2.0 * onp.pi, 3), deltas_suffix=""),   # Transmission3D.py:451, 498
    "3d_scalar": dict(ham=lambda k0: 4 * onp.pi / k0, ldos=(4.0 * onp.pi, 3), deltas_suffix=""),   # Transmission3D.py:960, 1004
    "2d_te": dict(ham=lambda k0: 8.0, ldos=(4.0, 2), deltas_suffix="_vector"),                   # Transmission2D.py:465, 513
    "2d_tm": dict(ham=lambda k0: 4.0, ldos=(4.0, 2), deltas_suffix="_scalar"),                   # Transmission2D.py:964, 1008
}


def hamiltonian_dos(G, k0, deltas, file_name, write_eigenvalues):
    """Transmission3D.py:440-461 (and siblings): eigenvalues of the scaled
    Green's matrix with a zero diagonal, the resolvent mean at deltas - i."""
    G.fill_diagonal_(0)
    lambdas = torch.linalg.eigvals(G)
    k0_ = onp.round(k0 / (2.0 * onp.pi), 1)
    if write_eigenvalues:
        onp.savetxt(file_name + '_hamiltonian_lambdas_' + str(k0_) + '_vector.csv',
                    onp.stack([torch.real(lambdas).numpy(), torch.imag(lambdas).numpy()]).T)
    resonance = deltas - 1j
    resolvant = torch.mean(1 / (resonance.reshape(-1, 1) - lambdas.reshape(1, -1)), axis=-1)
    return torch.imag(resolvant) / onp.pi


def eigen_ldos(M, kind, k0, alpha, npoints, file_name, write_eigenvalues):
    """Transmission3D.py:463-500 (and siblings): eigenvalues of M and the
    scatterer-averaged LDOS Im(pref sum (1 - delta)^2 / delta / N)."""
    deltas = torch.linalg.eigvals(M)
    k0_ = onp.round(k0 / (2.0 * onp.pi), 1)
    if write_eigenvalues:
        onp.savetxt(file_name + '_deltas_' + str(k0_) + SPECTRAL[kind]["deltas_suffix"] + '.csv',
                    onp.stack([torch.real(deltas).numpy(), torch.imag(deltas).numpy()]).T)
    num, power = SPECTRAL[kind]["ldos"]
    dos = ((1 - deltas) ** 2 / deltas).sum() / npoints
    dos *= num / (k0 ** power * alpha)
    return torch.imag(dos)


def eigenmodes_ipr(M, kind, k0, alpha, radius, file_name, number_eigenmodes, write_eigenvalues, sorting_type):
    """Transmission3D.py:502-564, 1011-1078; Transmission2D.py:518-569,
    1018-1066 without the plots and prints: eigenpairs of M, inverse
    participation ratios of the eigenvectors, and the rescaled eigenvalues
    (3D: Monsarrat's dimensionless lambdas; 2D: 1 - k0^2 alpha delta),
    sorted by IPR or by damping.  The particle volume is defined whether or
    not self-interaction is on (the reference's 3D forms use it unbound
    without self-interaction)."""
    deltas, vecs = torch.linalg.eig(M)
    iprs = torch.sum(torch.abs(vecs ** 4), axis=0) / (torch.sum(torch.abs(vecs ** 2), axis=0)) ** 2
    k0_ = onp.round(k0 / (2.0 * onp.pi), 1)
    if kind.startswith("3d"):
        volume = 4 * onp.pi * (radius ** 3) / 3
        lam = (1.0 - deltas) / (k0 ** 2 * alpha)
        if kind == "3d_vector":
            kares = onp.sqrt(1.0 + 3.0 / onp.real(alpha / volume))
            gamma0 = (2.0 / 9.0) * (1 + 3.0 / onp.real(alpha / volume))
            lam = 6.0 * volume * k0 ** 2 * lam / (kares * gamma0)
        else:
            kares = onp.sqrt(2.0 / onp.real(alpha / volume))
            gamma0 = 4.0 / (3.0 * onp.real(alpha / volume))
            lam = 4.0 * volume * k0 ** 2 * lam / (kares * gamma0)
        name = '_lambdas_'
    else:
        lam = 1.0 - k0 ** 2 * alpha * deltas
        name = '_deltas_'
    if write_eigenvalues:
        onp.savetxt(file_name + name + str(k0_) + '.csv',
                    onp.stack([torch.real(lam).numpy(), torch.imag(lam).numpy(), iprs.numpy()]).T)
    if sorting_type == 'IPR':
        iprs, idx = torch.sort(iprs, descending=True)
        lam = lam[idx]
        vecs = vecs[:, idx]
    elif sorting_type == 'damping':
        idx = torch.argsort(torch.imag(lam), descending=False)
        lam = lam[idx]
        iprs = iprs[idx]
        vecs = vecs[:, idx]
    else:
        raise NotImplementedError
    return lam, vecs[:, 0:number_eigenmodes], iprs


def _spectral_methods(kind, d, system_matrix, green_matrix):
    """The three spectral methods of a class, over its M and zero-self G."""
    def compute_hamiltonian_DOS(self, k0, deltas, file_name, write_eigenvalues=True):
        G = SPECTRAL[kind]["ham"](k0) * green_matrix(self, k0)
        return hamiltonian_dos(G, k0, deltas, file_name, write_eigenvalues)

    def compute_eigenvalues_and_scatterer_LDOS(self, k0, alpha, radius, file_name, self_interaction=True,
                                               self_interaction_type="Rayleigh", write_eigenvalues=True):
        M = system_matrix(self, k0, alpha, radius, self_interaction, self_interaction_type)
        return eigen_ldos(M, kind, k0, alpha, self.r.shape[0], file_name, write_eigenvalues)

    def compute_eigenmodes_IPR(self, k0, alpha, radius, file_name, self_interaction=True,
                               self_interaction_type="Rayleigh", number_eigenmodes=1, write_eigenvalues=True,
                               sorting_type='IPR'):
        M = system_matrix(self, k0, alpha, radius, self_interaction, self_interaction_type)
        return eigenmodes_ipr(M, kind, k0, alpha, radius, file_name, number_eigenmodes, write_eigenvalues,
                              sorting_type)
    return compute_hamiltonian_DOS, compute_eigenvalues_and_scatterer_LDOS, compute_eigenmodes_IPR


def _m_2d_te(self, k0, alpha, radius, si, kind):
    S = self_interaction_integral_vector_2d(k0, radius, kind) if si else 0
    return _system_matrix(self.G0_vector(None, k0), alpha, k0, S, onp.pi * radius * radius, si, order2d=True)


def _m_2d_tm(self, k0, alpha, radius, si, kind):
    S = self_interaction_integral_scalar_2d(k0, radius, kind) if si else 0
    return _system_matrix(self.G0(self.r, k0), alpha, k0, S, onp.pi * radius * radius, si, order2d=True)


def _m_3d_vector(self, k0, alpha, radius, si, kind):
    S = self_interaction_integral_vector_3d(k0, radius, kind) if si else 0
    return _system_matrix(self.G0(None, k0), alpha, k0, S, 4 * onp.pi * (radius ** 3) / 3, si, order2d=False)


def _m_3d_scalar(self, k0, alpha, radius, si, kind):
    S = self_interaction_integral_scalar_3d(k0, radius, kind) if si else 0
    return _system_matrix(self.G0(None, k0), alpha, k0, S, 4 * onp.pi * (radius ** 3) / 3, si, order2d=False)


(Transmission2D_vector.compute_hamiltonian_DOS, Transmission2D_vector.compute_eigenvalues_and_scatterer_LDOS,
 Transmission2D_vector.compute_eigenmodes_IPR) = _spectral_methods(
    "2d_te", 2, _m_2d_te, lambda self, k0: self.G0_vector(None, k0))
(Transmission2D_scalar.compute_hamiltonian_DOS, Transmission2D_scalar.compute_eigenvalues_and_scatterer_LDOS,
 Transmission2D_scalar.compute_eigenmodes_IPR) = _spectral_methods(
    "2d_tm", 1, _m_2d_tm, lambda self, k0: self.G0(self.r, k0))
(Transmission3D_vector.compute_hamiltonian_DOS, Transmission3D_vector.compute_eigenvalues_and_scatterer_LDOS,
 Transmission3D_vector.compute_eigenmodes_IPR) = _spectral_methods(
    "3d_vector", 3, _m_3d_vector, lambda self, k0: self.G0(None, k0))
(Transmission3D_scalar.compute_hamiltonian_DOS, Transmission3D_scalar.compute_eigenvalues_and_scatterer_LDOS,
 Transmission3D_scalar.compute_eigenmodes_IPR) = _spectral_methods(
    "3d_scalar", 1, _m_3d_scalar, lambda self, k0: self.G0(None, k0))


# ---------------------------------------------------------------------------
# full-field images (magreete.py:568-741), per angle and per grid batch
# ---------------------------------------------------------------------------
def image_fields_reference(solver, measurement_points, Ej, thetas, thetas_plot, k0, alpha, beam_waist, batch_size,
                           incidence=None, regularize=False, radius=0.0):
    """magreete.py:639-677: for every plotted angle, the grid in chunks of
    batch_size (torch.split), the incident field of that one angle and a
    one-column propagate of Ej[:, index]; returns the list over angles of
    Eall, (P,) for scalar and (P, d) for vector classes.  incidence(angle)
    gives the 3D (u, p) rows of one angle (utils.vector_3d_u_and_p)."""
    scalar = solver.d == 1
    ndim = measurement_points.shape[1]
    out = []
    for angle in thetas_plot:
        index = onp.where((thetas - angle) ** 2 < 1e-8)[0][0]
        Eall = []
        for batch_points in torch.split(measurement_points, batch_size):
            if ndim == 2:
                E0_meas = solver.generate_source(batch_points, k0, [angle], beam_waist)
            elif scalar:
                u_angle, _ = incidence(angle)
                E0_meas = solver.generate_source(batch_points, k0, u_angle.reshape(1, 3), beam_waist)
            else:
                u_angle, p_angle = incidence(angle)
                E0_meas = solver.generate_source(batch_points, k0, u_angle.reshape(1, 3), p_angle.reshape(1, 3),
                                                 beam_waist)
            if scalar:
                E0_meas = E0_meas.reshape(batch_points.shape[0], 1)
            else:
                E0_meas = E0_meas.reshape(batch_points.shape[0], ndim if ndim == 2 else 3, 1)
            E_meas = solver.propagate(batch_points, Ej[:, index].unsqueeze(-1), k0, alpha, E0_meas,
                                      regularize=regularize, radius=radius)
            Eall.append(E_meas)
        Eall = torch.cat(Eall, dim=0)
        if not scalar:
            Eall = Eall.squeeze(-1)
        out.append(Eall)
    return out


def field_maps_reference(Eall, measurement_points, ngridx, ngridy, scalar):
    """magreete.py:683-703 for one angle."""
    if scalar:
        return {"field": Eall.reshape(ngridy, ngridx)}
    ndim = measurement_points.shape[1]
    if ndim == 2:
        viewing_angle = torch.arctan2(measurement_points[:, 1], measurement_points[:, 0])
    else:
        viewing_unit_vector = measurement_points / torch.linalg.norm(measurement_points, axis=-1).unsqueeze(-1)
    amp = torch.sqrt(torch.sum(torch.absolute(Eall) ** 2, axis=-1))
    if ndim == 2:
        lon = Eall[:, 0] * torch.cos(viewing_angle) - Eall[:, 1] * torch.sin(viewing_angle)
        tra = Eall[:, 0] * torch.sin(viewing_angle) + Eall[:, 1] * torch.cos(viewing_angle)
    else:
        lon = torch.sum(Eall * viewing_unit_vector, axis=1)
        tra = Eall - lon.reshape(-1, 1) * viewing_unit_vector
        tra = torch.sqrt(torch.sum(torch.absolute(tra) ** 2, axis=1))
    return {"amplitude": amp.reshape(ngridy, ngridx), "longitudinal": lon.reshape(ngridy, ngridx),
            "transverse": tra.reshape(ngridy, ngridx)}


CLASSES = {
    "2d_tm": Transmission2D_scalar,
    "2d_te": Transmission2D_vector,
    "3d_vector": Transmission3D_vector,
    "3d_scalar": Transmission3D_scalar,
}
