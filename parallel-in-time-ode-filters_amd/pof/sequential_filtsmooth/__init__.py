"""Sequential square-root filter + RTS smoother on the GPU (reference ``pof/sequential_filtsmooth``).

``linear_filtsmooth(x0, dtm, dom)`` runs the square-root Kalman filter
(``filter.py:33-56``) and the RTS smoother (``smoother.py:8-48``) in
``libpof_hip.so``, one GPU worker stepping through time, and returns
``(MVNSqrt(means[N+1, D], chols[N+1, D, D]), nll, obj, ssq)``
(``__init__.py:13-16``); ``nll`` is the summed log-likelihood, as in the
reference.  ``extended_kalman_filter`` / ``filtsmooth`` take the residual
model of ``set_up_solver`` (a device vector field) and linearise it at each
predicted mean (``filter.py:9-30``); ``smoothing`` runs the RTS recursion over
given filtered states.
"""
import ctypes

import numpy as np

from .. import _lib
from ..parallel_filtsmooth import _constant_transition, _run, _smooth
from ..utils import MVNSqrt, as_f64_host


def linear_filtsmooth(x0, linear_transitions, linear_observations):
    """(``sequential_filtsmooth/__init__.py:13-16``): ``(MVNSqrt, nll, obj, ssq)``."""
    out, _, nll, obj, ssq = _run(x0, linear_transitions, linear_observations, False, sequential=True)
    return out, nll, obj, ssq


def linear_noiseless_filter(x0, discrete_transition_models, discrete_observation_models):
    """Filtered states (``filter.py:33-56``): ``(MVNSqrt, nll, ssq)``."""
    _, filt, nll, _, ssq = _run(x0, discrete_transition_models, discrete_observation_models, True,
                                sequential=True)
    return filt, nll, ssq


def extended_kalman_filter(x0, discrete_transition_models, continuous_observation_model):
    """EKF (``filter.py:9-30``): ``(MVNSqrt(means[N+1, D], chols[N+1, D, D]), nll, ssq)``.

    ``discrete_transition_models`` holds one (constant) ``(F, QL)`` per step
    (``[N, D, D]``, as ``set_up_solver`` builds it); ``continuous_observation_model``
    is the residual model of ``set_up_solver``.
    """
    import torch

    from ..observations import NonlinearModel, _problem_for

    om = continuous_observation_model
    if not isinstance(om, NonlinearModel) or not hasattr(om.f, "kind"):
        raise TypeError("extended_kalman_filter needs the residual model of set_up_solver (a device vector field)")
    Fs = discrete_transition_models.F if hasattr(discrete_transition_models, "F") else discrete_transition_models[0]
    if np.ndim(Fs) != 3:
        raise ValueError("pass one (F, QL) per step ([N, D, D]), as set_up_solver builds them")
    N = int(np.shape(Fs)[0])
    F, QL = _constant_transition(discrete_transition_models)
    spec = _problem_for(om)
    D = spec.D
    spec.ts = np.linspace(0.0, N * float(spec.ts[1] - spec.ts[0]), N + 1)
    m0 = np.ascontiguousarray(as_f64_host(x0.mean).reshape(D))
    L0 = np.ascontiguousarray(as_f64_host(x0.chol).reshape(D, D))
    dev = torch.device("cuda", torch.cuda.current_device())
    fm = torch.empty((N + 1, D), dtype=torch.float64, device=dev)
    fL = torch.empty((N + 1, D, D), dtype=torch.float64, device=dev)
    nll = ctypes.c_double()
    ssq = ctypes.c_double()
    prob = spec.cstruct()
    ctx = _lib.context(dev.index)
    torch.cuda.synchronize(dev)
    _lib.check(
        _lib.load().pof_ekf(ctx, ctypes.byref(prob), _lib.hptr(m0), _lib.hptr(L0), _lib.hptr(F), _lib.hptr(QL),
                            _lib.dptr(fm), _lib.dptr(fL), ctypes.byref(nll), ctypes.byref(ssq), _lib.stream_ptr(dev)),
        ctx, "pof_ekf",
    )
    return MVNSqrt(fm, fL), nll.value, ssq.value


def smoothing(discrete_transition_models, filter_trajectory):
    """Sequential square-root RTS smoother (``smoother.py:8-28``): ``(MVNSqrt, obj)``."""
    return _smooth(discrete_transition_models, filter_trajectory, True)


def filtsmooth(x0, linear_transitions, continuous_observation_model):
    """EKF then RTS smoother (``__init__.py:5-10``): ``(MVNSqrt, nll, obj, ssq)``."""
    out, nll, ssq = extended_kalman_filter(x0, linear_transitions, continuous_observation_model)
    out, obj = smoothing(linear_transitions, out)
    return out, nll, obj, ssq
