"""FP64 CPU oracle for the batched parameter-sweep integrator.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this directory; only the
test-suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it, and
only as the checker (or as the timed CPU comparison), never as a fallback for the HIP path.

Contents:

* ``lindblad``  -- numeric restatement of the reduced real Lindblad generator
  (reference ``gqis/solver.py:202-253`` layout, ``:399-493`` reduction, ``:431-433`` observable).
* ``integrate`` -- FP64 fixed-step RK4 / tableau RK / adaptive Tsit5 (``tsit5a``) / adaptive DOP853
  with the reference's stage-time, mean, trace and controller semantics
  (``gqis/cuda_solvers.py:470-517, 705-908``; ``gqis/solver.py:2126-2193``).
* ``models``    -- the numeric benchmark models (two-level LZS, four-level chain, random N-level,
  forced Duffing).
* ``rk4_affine.c`` -- the same fixed-step RK4 on an affine model in plain C (OpenMP over points),
  used as the fast oracle for large samples and as the timed CPU baseline.
* ``python_cpu`` -- the reference's own CPU path restated (per-point Python RK4 on a lambdified
  RHS in a spawn pool), timed beside the C baseline.

Every function here is pinned by closed forms, the reference's known-answer tests, SciPy's
``solve_ivp`` or matrix exponentials (see ``tests/test_oracle.py``).
"""
