"""Sequential square-root filter + RTS smoother on the GPU (reference ``pof/sequential_filtsmooth``).

``linear_filtsmooth(x0, dtm, dom)`` runs the noiseless square-root Kalman
filter (``filter.py:33-56``) and the RTS smoother (``smoother.py:8-48``) in
``libpof_hip.so``, one GPU worker stepping through time, and returns
``(MVNSqrt(means[N+1, D], chols[N+1, D, D]), nll, obj, ssq)``
(``__init__.py:13-16``); ``nll`` is the summed log-likelihood, as in the
reference.  The EKS of ``filtsmooth`` (``__init__.py:5-10``) is
``pof.solver.sequential_eks_solve``.
"""
from ..parallel_filtsmooth import _run


def linear_filtsmooth(x0, linear_transitions, linear_observations):
    """(``sequential_filtsmooth/__init__.py:13-16``): ``(MVNSqrt, nll, obj, ssq)``."""
    out, _, nll, obj, ssq = _run(x0, linear_transitions, linear_observations, False, sequential=True)
    return out, nll, obj, ssq


def linear_noiseless_filter(x0, discrete_transition_models, discrete_observation_models):
    """Filtered states (``filter.py:33-56``): ``(MVNSqrt, nll, ssq)``."""
    _, filt, nll, _, ssq = _run(x0, discrete_transition_models, discrete_observation_models, True,
                                sequential=True)
    return filt, nll, ssq
