"""CPU oracle for the block-tridiagonal-arrowhead (BTA) solver path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: a numpy/scipy
restatement of the reference algorithms on the hot path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to check or to time the CPU side.  The product path
(``pyinla_amd``) never imports it and fails loudly when its HIP library is
missing.

Modules
-------
bta    : pobtaf / pobtas / pobtasi / logdet block recurrences (the arithmetic
         the reference hands to the external Serinv package) plus dense
         conversions and the dense oracle.
model  : f(theta) of the Gaussian spatio-temporal model and its central
         finite-difference gradient, restated from the reference driver.
scatter: the reference's sparse -> BTA block scatter and the selected-inverse
         gather back to a sparse matrix.

Pinning status (what each function is checked against, in tests/test_oracle.py):
  pobtaf/pobtf  - known-factor construction (Cholesky uniqueness), dense LAPACK,
                  Kronecker logdet identity, b=1 tridiagonal closed form,
                  Schur-complement identity.
  logdet        - the same, plus the closed form ln(n+1) of tridiag(-1,2,-1).
  pobtas/pobts  - residual, dense cho_solve, Kronecker solve identity.
  pobtasi/pobtsi- dense inverse on the BTA pattern, Kronecker inverse identity.
  scatter/gather- dense slicing of the same matrix.
  model.f       - dense slogdet restatement, x-invariance identity of the
                  Gaussian marginal, closed-form Gaussian log-density.
  model FD grad - dense float64 torch autograd of f.
"""
