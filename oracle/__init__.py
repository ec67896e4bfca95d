"""CPU FP64 oracle for the implicit cpGIMP MPM load-step Newton loop.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy restatement of the
reference algorithm (GeoWarp: mpm/gimp.py, mpm/mpm_utils.py,
materials/plasticity.py, solvers/sparse_differentiation.py and the three
solvers/quasi_static_*_solver_*.py classes).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import
it, and only as the checker.  The product path (``geowarp_amd``) never
imports it and fails loudly when its HIP library is missing.

Every function cites the reference file:line it restates.  Quirks of the
reference are reproduced on purpose (transposed increment gradient, the 1e-8
floor bias, the perturbed initial F, the post-update convergence test, ...).

Parity pins (what checks each part of the oracle, see tests/test_oracle_*.py):

* GIMP weights: closed form of the domain-averaged linear hat (brute-force
  quadrature), partition of unity, linear reproduction, FD of the weights.
* Residuals: rigid-translation patch test, uniform-stretch patch test, global
  force balance, closed-form stresses, gravity-column known answer
  (bar_compaction_cpGIMP_J2.py:126), FD of the per-particle energy-free terms.
* Jacobian: coloured central FD (the reference's 5^d colouring semantics).
* Return mappings: yield-surface consistency of the converged state, J2 closed
  form, Drucker-Prager cone closed form (shape factor 0), NorSand local
  residual equations to round-off.
* Coupled u-p terms (tests/test_oracle_coupled.py): consistent-mass rows and
  linear-pressure recovery, the pressure block's symmetry and null space, the
  Darcy term of a linear pressure equal to the boundary flux, the undrained
  column carrying the load in pore pressure, Terzaghi's series, the penalty
  indenter force in closed form.
* Linear solves: scipy.sparse.linalg.spsolve.

cpu_port.cpp / cpu_port.py restate the 3D u-p Neo-Hookean assembly in C++17 with
OpenMP (the benchmark's CPU baseline).  Its Jacobian is exact, by forward-mode
dual numbers, and is checked against this package's residual and coloured
central differences (tests/test_cpu_port.py).
"""
