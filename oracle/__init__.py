"""CPU oracle for the TVSA implicit step engine -- TEST INFRASTRUCTURE ONLY.

This package is a plain NumPy/SciPy restatement of the upstream simulator's step engine
(residual, coloured finite-difference Jacobian, SDIRK2/TR-BDF2 Newton-Krylov with the SuperLU
theta-block preconditioner, step runner, balances).  Every function cites the upstream
file:line it restates.  It exists for two reasons only:

* as the checker that the HIP path (``dac_vtsa_amd``) is compared with in ``tests/`` and in
  ``__graft_entry__.smoke()``;
* as the timed CPU baseline (``bench.py``'s ``cpu_baseline`` leg).

Nothing in the product package imports it, and it is never a fallback: the product fails
loudly when its HIP library is missing.

Pinning (what makes this restatement trustworthy):
* known answers held by the upstream tests (isotherm hand values, p_sat(373.15));
* invariants (species/energy conservation of rhs + accumulators, T<->e round trip,
  axisymmetry, Ergun inversion round trip, softplus split identity);
* golden vectors written by running the upstream code where it is installed
  (``tests/golden/make_golden.py``): rhs/acc/diag on seeded states for every step type,
  FD Jacobian values, and a short SDIRK2 and TR-BDF2 integration;
* brute force on tiny inputs (FD Jacobian vs one-column FD, integrator vs RK4).
Functions with none of these checks are marked "parity unpinned" in their docstrings.
"""
