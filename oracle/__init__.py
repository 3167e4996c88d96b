"""CPU oracle for the Kennett reflectivity path — TEST INFRASTRUCTURE ONLY.

This package is a plain NumPy (float64 / complex128) restatement of the
reference algorithm (tmnestor/SeismicInversion, Kennett_Reflectivity/).
Each function cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU
baseline. The product package ``seismicinversion_amd`` never imports it.

Pinning status of each function (what checks it besides its own formula):

* ``complex_slowness``, ``vertical_slowness`` — closed form and the
  eta^2 = (s+p)(s-p), Im(eta) > 0 invariants (tests/test_oracle.py:81-97), and
  the golden vectors.
* ``interface_coefficients`` — Tu = Td^T, Rd/Ru symmetry, the ocean-bottom
  zero pattern, the normal-incidence impedance formula at p = 0, identical
  media (Rd = Ru = 0, Td = Tu = I), and energy-flux unitarity for elastic
  media; end-to-end through the golden reflectivities.
* ``reflectivity_batch`` / ``reflectivity`` — golden vectors written by the
  reference itself (tests/golden/make_golden.py), layer-splitting
  invariance, the two-layer closed form, the free-surface identity.
* ``bessel_sum``, ``hermitian_ifft``, ``gather`` — golden gather written by the
  reference's float64 ``compute_gather``; J0 against ``scipy.special.j0``.
* ``kennett_torch_ref`` — AD through the torch restatement against the golden
  Jacobian/Hessian of the reference and central finite differences.
* ``reflectivity_batch(..., ctype=np.clongdouble)`` — the same operations in
  x87 extended precision, the exact value of the parity tests: against
  40-digit mpmath (``kennett_mp``), which satisfies the free-surface identity
  to its own precision.
* ``taup_lm_ref`` — the LM loop against the reference's own 3-iteration
  invert_taup trajectory (tests/golden/invert_taup_2pct.npz) to 1e-12.
"""
