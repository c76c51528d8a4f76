"""CPU float64 oracle for the veros mixing hot path.

TEST INFRASTRUCTURE ONLY. This package restates, in NumPy, the reference
algorithms that the HIP library replaces, so that tests can check the GPU
results on identical inputs. Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
(or the timed CPU baseline) — never as a product code path.

Modules:

* ``oracle.tdma`` — ``create_water_masks``, ``solve_implicit`` and the two
  tridiagonal solvers (LAPACK ``dgtsv`` over concatenated wet cells, as
  veros/core/operators.py:60-77 does, and a column-wise Thomas sweep as
  veros/core/operators.py:133-156 does), plus a dense brute-force solver.
* ``oracle.isoneutral`` — ``isoneutral_diffusion_pre``
  (veros/core/isoneutral/isoneutral.py:10-229) and the isoneutral tracer
  diffusion kernel with its energy branch (veros/core/isoneutral/diffusion.py:9-283,
  veros/core/diffusion.py:9-62).

How each function is pinned (golden vectors from the reference's own NumPy
path, closed forms, invariants, brute force) is listed in the module headers.
"""
