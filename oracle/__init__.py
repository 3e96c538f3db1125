"""CPU oracle for the coupled-dipole hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``magreete_amd``) imports
this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker or
as the timed CPU baseline, never as the thing measured or shipped.

``oracle.transmission`` restates, on torch CPU in complex128, the ★ methods of
the four reference classes (Transmission2D.py:10-453, 570-953;
Transmission3D.py:11-439, 565-949): Green's functions, ``G0``, sources,
``solve``, ``propagate``, ``propagate_ss``, ``LDOS_measurements``,
``batch_LDOS_measurements`` and ``mean_DOS_measurements``.  The reference is
itself a composition of torch ops, so the same ops on the same torch give the
reference's numbers to rounding.

How each function is pinned (tests/test_oracle_*.py):

* ``hankel1``: mpmath golden values (tests/golden/bessel_golden.json),
  Wronskian, small-argument log form.
* Green's functions: a textbook form written independently
  (e^{ikR}/(4πR)[(1 + i/kR − 1/(kR)²) I + (−1 − 3i/kR + 3/(kR)²) R̂R̂] for 3D,
  the 2D dyadic written with H0/H2), mpmath golden values, symmetry
  G(r) = G(−r), the R→0 limits of Im G.
* ``solve``: N=1 closed form with and without self-interaction; numpy LAPACK
  (an independent library from torch's MKL) and mpmath LU on small systems.
* ``propagate`` / ``propagate_ss``: N=1 closed form, linearity over columns.
* LDOS / mean DOS: the site identity (diag of G W G^T from diag(W)) and
  mean_DOS == mean(LDOS).
* sources: plane-wave closed form and |E0| = 1, Gaussian-beam formula by mpmath;
  the 2D ``'point'`` sources as a dipole at R^T (-2, 0) through scipy's Hankel
  functions (rotation covariance of the Green's functions).

* the ``regularize`` NaN fix-up, multi-candidate branch included (several
  scatterers within ``radius`` of one measurement point: the nearest one's
  Ek), against a brute-force restatement on a hand-placed configuration
  (tests/test_oracle_pins.py).

Every function has a pin of this kind; none is parity unpinned.
"""
