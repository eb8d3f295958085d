"""pof: parallel-in-time probabilistic ODE filters on AMD MI355X (gfx950).

Drop-in for the reference package ``pof`` (nathanaelbosch/parallel-in-time-ode-filters):
``pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init=...)`` returns
``(MVNSqrt(means[len(ts), d], chol_covs[len(ts), d, D]), info)``, computed by
hand-written HIP kernels in ``libpof_hip.so`` behind the C ABI of
``include/pof.h``.  Arrays come back as float64 torch tensors on the GPU.
"""
