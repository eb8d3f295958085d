"""Square-root parallel filter + smoother on the GPU (reference ``pof/parallel_filtsmooth``).

``linear_filtsmooth(x0, dtm, dom)`` runs the noiseless parallel Kalman filter
(associative scan, ``filter.py:18-47``) and the parallel RTS smoother
(``smoother.py:8-22``) in ``libpof_hip.so`` and returns
``(MVNSqrt(means[N+1, D], chols[N+1, D, D]), nll, obj, ssq)``
(``__init__.py:5-10``).  The transition model must be constant over time
(the solver always builds it so, ``convenience.py:19``).  The observations
may carry a noise factor ``cholR`` (``[N, d, d]``, from
``uncertain_linearize`` or ``linearize_unscented``); EK1 produces ``0``.
"""
import ctypes

import numpy as np

from .. import _lib
from ..utils import MVNSqrt, as_f64_device, as_f64_host


def _constant_transition(dtm):
    F = as_f64_host(dtm.F if hasattr(dtm, "F") else dtm[0])
    QL = as_f64_host(dtm.QL if hasattr(dtm, "QL") else dtm[1])
    if F.ndim == 3:
        if not (np.all(F == F[:1]) and np.all(QL == QL[:1])):
            raise NotImplementedError("time-varying transition models are not supported; pass one (F, QL) pair")
        F, QL = F[0], QL[0]
    return np.ascontiguousarray(F), np.ascontiguousarray(QL)


def _run(x0, dtm, dom, want_filter, sequential=False):
    import torch

    F, QL = _constant_transition(dtm)
    H = as_f64_device(dom.H if hasattr(dom, "H") else dom[0])
    c = as_f64_device(dom.b if hasattr(dom, "b") else dom[1])
    cholR = dom.cholR if hasattr(dom, "cholR") else dom[2]
    Ns, d, D = H.shape
    R = None
    if cholR is not None:
        R = as_f64_device(cholR)
        if tuple(R.shape) != (Ns, d, d):
            raise ValueError(f"cholR has shape {tuple(R.shape)}, expected {(Ns, d, d)}")
        if not bool((R != 0).any()):
            R = None  # noiseless: the kernels' R = 0 path
    m0 = as_f64_host(x0.mean).reshape(D)
    L0 = as_f64_host(x0.chol).reshape(D, D)
    dev = H.device
    sm = torch.empty((Ns + 1, D), dtype=torch.float64, device=dev)
    sL = torch.empty((Ns + 1, D, D), dtype=torch.float64, device=dev)
    fm = torch.empty_like(sm) if want_filter else None
    fL = torch.empty_like(sL) if want_filter else None
    nll = ctypes.c_double()
    obj = ctypes.c_double()
    ssq = ctypes.c_double()
    ctx = _lib.context(dev.index)
    _lib.check(
        _lib.load().pof_linear_filtsmooth(
            ctx, D, d, Ns, _lib.hptr(m0), _lib.hptr(L0), _lib.hptr(F), _lib.hptr(QL), _lib.dptr(H), _lib.dptr(c),
            _lib.dptr(R), 1 if sequential else 0, _lib.dptr(sm), _lib.dptr(sL), _lib.dptr(fm), _lib.dptr(fL), ctypes.byref(nll), ctypes.byref(obj),
            ctypes.byref(ssq), _lib.stream_ptr(dev)),
        ctx, "pof_linear_filtsmooth",
    )
    return MVNSqrt(sm, sL), (MVNSqrt(fm, fL) if want_filter else None), nll.value, obj.value, ssq.value


def linear_filtsmooth(x0, linear_transitions, linear_observations):
    """(``parallel_filtsmooth/__init__.py:5-10``): ``(MVNSqrt, nll, obj, ssq)``."""
    out, _, nll, obj, ssq = _run(x0, linear_transitions, linear_observations, False)
    return out, nll, obj, ssq


def linear_noiseless_filtering(x0, transition_models, observation_models):
    """Filtered states (``filter.py:18-47``): ``(MVNSqrt, nll, obj, ssq)``.

    ``obj`` is the smoother's objective (the filter's own value is discarded
    by the reference, ``__init__.py:6-9``).
    """
    _, filt, nll, obj, ssq = _run(x0, transition_models, observation_models, True)
    return filt, nll, obj, ssq


def _smooth(transition_models, filter_trajectory, sequential):
    import torch

    F, QL = _constant_transition(transition_models)
    fm = as_f64_device(filter_trajectory.mean)
    fL = as_f64_device(filter_trajectory.chol)
    n1, D = fm.shape
    if fL.shape != (n1, D, D):
        raise ValueError(f"filter_trajectory.chol has shape {tuple(fL.shape)}, expected {(n1, D, D)}")
    sm = torch.empty_like(fm)
    sL = torch.empty_like(fL)
    obj = ctypes.c_double()
    ctx = _lib.context(fm.device.index)
    torch.cuda.synchronize(fm.device)
    _lib.check(
        _lib.load().pof_smoothing(ctx, D, n1 - 1, _lib.hptr(F), _lib.hptr(QL), _lib.dptr(fm), _lib.dptr(fL),
                                  1 if sequential else 0, _lib.dptr(sm), _lib.dptr(sL), ctypes.byref(obj),
                                  _lib.stream_ptr(fm.device)),
        ctx, "pof_smoothing",
    )
    return MVNSqrt(sm, sL), obj.value


def smoothing(transition_models, filter_trajectory):
    """Parallel square-root RTS smoother over given filtered states (``smoother.py:8-22``):
    ``(MVNSqrt(means[N+1, D], chols[N+1, D, D]), obj)``."""
    return _smooth(transition_models, filter_trajectory, False)
