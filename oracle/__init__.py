"""CPU oracle for famrdp_jax's right-hand-side and TVD-RK3 path.

THIS PACKAGE IS TEST INFRASTRUCTURE.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker.  The product (``famrdp-jax_amd/``) never
imports, links or executes anything here, and fails loudly when its HIP
library is missing instead of falling back to this code.

Contents (each function cites the reference file:line it restates):

* ``numerics``: eos, MUSCL-2 (minmod / van Leer), Roe + Harten, inviscid RHS,
  viscous RHS, RK3 (numpy forward; also runs on CPU float64 torch tensors
  as the autograd oracle with JAX's derivative conventions, see ``_ops``).
* ``blocks``: halo exchange, BCs, the full ``step``.
* ``metric``: 4th-order metrics, Fortran probe readers, the build-defined CFL
  time step.
* ``c_step`` (with ``c/step_omp.c``): the full ``step`` of one periodic block
  restated in C + OpenMP in the numpy oracle's evaluation order, for the CPU
  baseline on every host thread; pinned against ``blocks.step``
  (``tests/test_oracle_c.py``).  Not imported here (it needs its built
  library).

Pinning (what checks each function independently of its own formula) is
listed in ``tests/test_oracle_*.py``; every function above is pinned by at
least one known-answer test, closed form, invariant or brute-force check.
The only result no check can pin without running the reference is the
bit-level XLA:CPU rounding of the composed step; the step tolerances
(1e-12 per step, 1e-9 after 100 steps) are what absorb it.
"""
from . import numerics, blocks, metric  # noqa: F401
