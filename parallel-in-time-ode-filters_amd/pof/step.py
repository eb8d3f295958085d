"""One IEKS iteration (reference ``pof/step.py``)."""
import ctypes

from . import _lib
from .observations import linearize
from .utils import MVNSqrt, as_f64_device


def linearize_at_previous_states(om, prev_states):
    """EK1 at ``states.mean[1:]`` (``step.py:12-22``)."""
    m = prev_states.mean if isinstance(prev_states, MVNSqrt) else prev_states
    return linearize(om, MVNSqrt(as_f64_device(m)[1:], None))


def ieks_step(*, om=None, dtm=None, x0=None, states, calibrate=True, sequential=False, setup=None):
    """(``step.py:33-45``): linearise at ``states.mean[1:]``, filter + smoother, chol *= sqrt(ssq).

    Parallel (associative scans) or, with ``sequential=True``, the sequential
    square-root filter and RTS smoother (``step.py:39-40``; there ``nll`` is
    the summed log-likelihood, as in the reference).  The native step needs the
    problem description; pass ``setup=`` (from ``set_up_solver``).  Returns
    ``(MVNSqrt(mean[N+1, D], chol[N+1, D, D]), nll, obj, ssq)``.
    """
    import torch

    if setup is None:
        raise TypeError("ieks_step needs setup= (the dictionary returned by pof.convenience.set_up_solver)")
    spec = setup["spec"]
    m = as_f64_device(states.mean if isinstance(states, MVNSqrt) else states)
    N, D = m.shape
    if N != spec.Ns + 1 or D != spec.D:
        raise ValueError(f"states have shape {tuple(m.shape)}, expected ({spec.Ns + 1}, {spec.D})")
    om_ = torch.empty_like(m)
    oL = torch.empty((N, D, D), dtype=torch.float64, device=m.device)
    nll = ctypes.c_double()
    obj = ctypes.c_double()
    ssq = ctypes.c_double()
    prob = spec.cstruct()
    ctx = _lib.context(m.device.index)
    _lib.check(
        _lib.load().pof_ieks_step(ctx, ctypes.byref(prob), _lib.dptr(m), 1 if sequential else 0,
                                  _lib.dptr(om_), _lib.dptr(oL),
                                  ctypes.byref(nll), ctypes.byref(obj), ctypes.byref(ssq), _lib.stream_ptr(m.device)),
        ctx, "pof_ieks_step",
    )
    if not calibrate:
        oL = oL / (ssq.value ** 0.5)
    return MVNSqrt(om_, oL), nll.value, obj.value, ssq.value
