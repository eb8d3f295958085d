"""CPU FP64 restatement of the parallel-in-time IEKS ODE solver (the oracle).

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP product
path: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` may import it, and only to check or to time the CPU baseline.
The product (``pof`` over ``libpof_hip.so``) never imports, links or calls
anything in here, and has no CPU fallback.

It restates, in NumPy float64, the algorithm of the reference package ``pof``
(nathanaelbosch/parallel-in-time-ode-filters).  Every function cites the
reference file:line it follows.  The reference depends on jax 0.3.23 /
jaxlib 0.3.22 (``poetry.lock:467-468,494-495``), tornadox 0.0.1
(``poetry.lock:1312-1313``) and scipy 1.9.2 (``poetry.lock:1165-1166``); none
of jax/jaxlib/tornadox is importable here, so their published algorithms are
restated:

* ``jax.lax.associative_scan``: the odd/even recursion (pair-combine,
  recurse, fill the evens; ``reverse=True`` = reverse, scan, reverse).
* ``jnp.linalg.qr(mode="r")``: LAPACK ``dgeqrf`` (NumPy calls the same
  LAPACK routine, so the Householder sign convention is the same).
* ``jax.jacfwd``: forward-mode AD, restated with dual numbers.
* ``tornadox.init.TaylorMode``: Taylor-mode AD, restated with truncated
  power series (jets); returns the derivatives y^(k)(t0) and a zero P0.

How each part is pinned (the tests that check the oracle itself):

* transitions / preconditioner / projections: closed forms
  (``A_ij = C(q-i, j-i)``, ``Q_ij = 1/(2q+1-i-j)``, ``P A P^-1 = dt^(j-i)/(j-i)!``)
  and the reference's ``tests/test_transtions.py`` layout checks.
* Taylor-mode init: closed-form derivatives of polynomial fields.
* EK1 linearisation: ``H m + b == r(m)`` (reference ``tests/test_observations.py:33``),
  finite differences of the residual.
* parallel filter / smoother: brute-force conditioning of the dense joint
  Gaussian on tiny models; parallel == sequential on random models; the
  degenerate all-zero model of reference ``tests/test_filtsmooth.py:41-62``.
* full solve: iteration counts and rmse values of the reference's own
  work-precision CSVs (``experiments/3_work_precision_diagram/data/*.csv``),
  the README figure's decoded means (``readmefig.svg``), and the zero
  ODE residual at convergence.

The filter's ``nll`` (one-step predictive terms) and the smoother's ``obj``
(with the reference's call order, ``||QL^-1 (m_k - F m_{k+1})||^2``) are pinned
on tiny linear models against the dense joint Gaussian, and so is
``sigma_squared`` for d = 1.

Parity unpinned: the raw entries of Cholesky factors (only ``L L^T`` is
unique), and ``sigma_squared`` for d > 1, whose terms the reference whitens
with the factor's transpose (``yhat^T (L^T L)^-1 yhat``), a value that depends
on the signs of the LAPACK factors along the chain.
"""
