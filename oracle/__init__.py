"""CPU restatement of lyspeq's per-chunk QML step: TEST INFRASTRUCTURE ONLY.

This package is the correctness oracle and the timed CPU baseline. It is
imported only by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``, always as the checker and never as the
thing measured or shipped.  The product path (``lyspeq_amd``) never imports
it and fails loudly when its HIP library is missing.

Modules
-------
``bins``      k/z binning, redshift-bin lookup, z-bin weight functions
              (core/global_numbers.cpp:209-471).
``interp``    uniform-grid linear and bilinear interpolation with the
              reference's clamp/extrapolate arithmetic
              (mathtools/discrete_interpolation.cpp:61-72,221-242).
``chunk``     the chunk estimator: construction, S/Q assembly, LU inverse,
              continuum marginalisation, d/b/t, Fisher and accumulation
              (core/chunk_estimate.cpp:141-914).  Also a Cholesky form of the
              same mathematics, used only to measure the LU-vs-Cholesky
              rounding floor.
``driver``    the Newton-Raphson consumer of F, d, b, t: 1/2 scaling,
              stableInvertSym, theta and sigma
              (core/quadratic_estimate.cpp:261-318,456-457,
              mathtools/matrix_helper.cpp:184-204,219-230,294-358).
``tables``    the S/Q table generator: QAWO quadrature for the derivative
              table and the c2r-FFT signal table
              (core/sq_table.cpp:201-323, core/fiducial_cosmology.cpp:273-331).

How each function is pinned (golden vectors, invariants, closed forms) is
listed in each module's docstring and exercised by ``tests/test_oracle_golden.py`` and ``tests/test_bootstrap_oracle.py``.
Parity status: every function above is pinned, except where a module header
says "parity unpinned".
"""
