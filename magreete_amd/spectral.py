"""Spectral methods of the solver classes: eigenvalues of the Green's matrix
or of M, the scatterer LDOS from M's spectrum and the eigenmodes' inverse
participation ratios (Transmission3D.py:440-564, 951-1079;
Transmission2D.py:455-569, 955-1066).

The matrices are assembled on the GPU by the same kernel as the solves; the
non-Hermitian eigen-decomposition runs on the host with torch's LAPACK
(zgeev), as in the reference, so eigenvalues come in LAPACK's order.  The
CSV files keep the reference's names and columns; its plots and progress
prints are not produced.
"""
import numpy as onp
import torch

# kind -> (prefactor of the Hamiltonian's Green's matrix as a function of k0,
#          numerator and power of k0 of the LDOS prefactor, CSV suffix of the deltas)
_CONST = {
    "3d_vector": (lambda k0: 6.0 * onp.pi / k0, 2.0 * onp.pi, 3, ""),    # Transmission3D.py:451, 498
    "3d_scalar": (lambda k0: 4.0 * onp.pi / k0, 4.0 * onp.pi, 3, ""),    # Transmission3D.py:960, 1004
    "2d_te": (lambda k0: 8.0, 4.0, 2, "_vector"),                        # Transmission2D.py:465, 513
    "2d_tm": (lambda k0: 4.0, 4.0, 2, "_scalar"),                        # Transmission2D.py:964, 1008
}


def hamiltonian_prefactor(kind, k0):
    return _CONST[kind][0](k0)


def _k0_label(k0):
    return str(onp.round(k0 / (2.0 * onp.pi), 1))


def _save(path, cols):
    onp.savetxt(path, onp.stack([c.numpy() for c in cols]).T)


def hamiltonian_dos(G, k0, deltas, file_name, write_eigenvalues):
    """Im of the mean resolvent 1 / (delta - i - lambda) / pi over the
    eigenvalues lambda of the scaled zero-diagonal Green's matrix G."""
    G = G.cpu()
    G.fill_diagonal_(0)
    lam = torch.linalg.eigvals(G)
    if write_eigenvalues:
        _save(file_name + '_hamiltonian_lambdas_' + _k0_label(k0) + '_vector.csv', [lam.real, lam.imag])
    z = torch.as_tensor(deltas).cpu() - 1j
    res = (1.0 / (z.reshape(-1, 1) - lam.reshape(1, -1))).mean(dim=-1)
    return res.imag / onp.pi


def eigen_ldos(M, kind, k0, alpha, npoints, file_name, write_eigenvalues):
    """Im(pref sum (1 - delta)^2 / delta / N) over the eigenvalues of M."""
    delta = torch.linalg.eigvals(M.cpu())
    if write_eigenvalues:
        _save(file_name + '_deltas_' + _k0_label(k0) + _CONST[kind][3] + '.csv', [delta.real, delta.imag])
    num, power = _CONST[kind][1], _CONST[kind][2]
    val = ((1 - delta) ** 2 / delta).sum() / npoints
    val = val * (num / (k0 ** power * alpha))
    return val.imag


def eigenmodes_ipr(M, kind, k0, alpha, radius, file_name, number_eigenmodes, write_eigenvalues, sorting_type):
    """Eigenpairs of M, the eigenvectors' inverse participation ratios
    sum |v|^4 / (sum |v|^2)^2, and the eigenvalues rescaled as the
    reference's (3D: (1 - delta) / (k0^2 alpha) in Monsarrat's units, with
    the particle volume defined also without self-interaction, where the
    reference leaves it unbound; 2D: 1 - k0^2 alpha delta), sorted by
    decreasing IPR or increasing damping."""
    delta, vecs = torch.linalg.eig(M.cpu())
    a2 = vecs.abs() ** 2
    ipr = (a2 * a2).sum(dim=0) / a2.sum(dim=0) ** 2
    if kind.startswith("3d"):
        vol = 4.0 * onp.pi * radius ** 3 / 3.0
        x = onp.real(alpha / vol)
        if kind == "3d_vector":
            kares, gamma0, scale = onp.sqrt(1.0 + 3.0 / x), (2.0 / 9.0) * (1.0 + 3.0 / x), 6.0
        else:
            kares, gamma0, scale = onp.sqrt(2.0 / x), 4.0 / (3.0 * x), 4.0
        lam = scale * vol * k0 ** 2 * ((1.0 - delta) / (k0 ** 2 * alpha)) / (kares * gamma0)
        stem = '_lambdas_'
    else:
        lam = 1.0 - k0 ** 2 * alpha * delta
        stem = '_deltas_'
    if write_eigenvalues:
        _save(file_name + stem + _k0_label(k0) + '.csv', [lam.real, lam.imag, ipr])
    if sorting_type == 'IPR':
        ipr, order = torch.sort(ipr, descending=True)
        lam, vecs = lam[order], vecs[:, order]
    elif sorting_type == 'damping':
        order = torch.argsort(lam.imag, descending=False)
        lam, ipr, vecs = lam[order], ipr[order], vecs[:, order]
    else:
        raise NotImplementedError
    return lam, vecs[:, :number_eigenmodes], ipr
