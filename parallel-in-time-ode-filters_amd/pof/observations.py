"""Observation models and their linearisations (reference ``pof/observations.py``).

``linearize`` (EK1), ``linearize_ek0``, ``uncertain_linearize`` and
``linearize_regularized`` run on the GPU (``pof_linearize`` /
``pof_linearize_model``), vectorised over a leading batch axis of states.
"""
import ctypes
from typing import Any, NamedTuple

import numpy as np

from . import _lib
from .utils import MVNSqrt, as_f64_device


class NonlinearModel(NamedTuple):
    """The ODE residual ``r(x) = E1 x - f(E0 x)`` (``observations.py:14-20``, built at ``convenience.py:28``).

    ``f`` must be a device vector field; ``E0``/``E1`` are the preconditioned
    projections ``E_k P``.
    """

    f: Any
    E0: Any
    E1: Any

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        return x @ self.E1.T - self.f(None, x @ self.E0.T)


class AffineModel(NamedTuple):
    """``y = H x + b`` with noise factor ``cholR`` (``observations.py:23-32``)."""

    H: Any
    b: Any
    cholR: Any


def _problem_for(om):
    from .convenience import ProblemSpec

    f = om.f
    D = om.E0.shape[1]
    q = D // f.d - 1
    # dt enters only through P; recover |dt| from P_00 = |dt|^(q + 1/2) / q!
    import math

    p00 = float(om.E0[0, 0])
    dt = (p00 * math.factorial(q)) ** (1.0 / (q + 0.5))
    spec = ProblemSpec(f, np.zeros(f.d), np.array([0.0, dt]), q)
    return spec


def linearize(f: NonlinearModel, x):
    """EK1 (``observations.py:35-40``) at the means ``x.mean[n, D]`` on the GPU.

    Returns ``AffineModel(H[n, d, D], b[n, d], cholR = 0[n, d, d])`` as device tensors.
    """
    import torch

    if not isinstance(f, NonlinearModel) or not hasattr(f.f, "kind"):
        raise TypeError("linearize needs the residual model built by set_up_solver (a device vector field)")
    m = as_f64_device(x.mean if isinstance(x, MVNSqrt) else x)
    single = m.dim() == 1
    if single:
        m = m[None]
    n, D = m.shape
    spec = _problem_for(f)
    d = spec.d
    H = torch.empty((n, d, D), dtype=torch.float64, device=m.device)
    c = torch.empty((n, d), dtype=torch.float64, device=m.device)
    prob = spec.cstruct()
    ctx = _lib.context(m.device.index)
    _lib.check(
        _lib.load().pof_linearize(ctx, ctypes.byref(prob), _lib.dptr(m), n, _lib.dptr(H), _lib.dptr(c),
                                  _lib.stream_ptr(m.device)),
        ctx, "pof_linearize",
    )
    R = torch.zeros((n, d, d), dtype=torch.float64, device=m.device)
    if single:
        return AffineModel(H[0], c[0], R[0])
    return AffineModel(H, c, R)


def _linearize_model(f, x, method, params=None, need_chol=False):
    """Shared driver of ``pof_linearize_model``: returns ``AffineModel`` device tensors."""
    import torch

    if not isinstance(f, NonlinearModel) or not hasattr(f.f, "kind"):
        raise TypeError("the residual model built by set_up_solver (a device vector field) is required")
    if isinstance(x, MVNSqrt):
        m, L = x.mean, x.chol
    else:
        m, L = x, None
    m = as_f64_device(m)
    single = m.dim() == 1
    if single:
        m = m[None]
    n, D = m.shape
    Lt = None
    if need_chol:
        if L is None:
            raise ValueError("this linearisation needs the state's factor (x.chol)")
        Lt = as_f64_device(L).reshape(n, D, D).contiguous()
    spec = _problem_for(f)
    d = spec.d
    dO = d + D if method == _lib.LIN_METHODS["regularized"] else d
    H = torch.empty((n, dO, D), dtype=torch.float64, device=m.device)
    c = torch.empty((n, dO), dtype=torch.float64, device=m.device)
    R = torch.empty((n, dO, dO), dtype=torch.float64, device=m.device)
    prob = spec.cstruct()
    pa = np.array(params if params is not None else [1.0, 1.0, 0.0, np.nan], dtype=np.float64)
    ctx = _lib.context(m.device.index)
    _lib.check(
        _lib.load().pof_linearize_model(ctx, ctypes.byref(prob), method, _lib.hptr(pa), _lib.dptr(m),
                                        _lib.dptr(Lt), n, _lib.dptr(H), _lib.dptr(c), _lib.dptr(R),
                                        _lib.stream_ptr(m.device)),
        ctx, "pof_linearize_model",
    )
    torch.cuda.synchronize(m.device)
    if single:
        return AffineModel(H[0], c[0], R[0])
    return AffineModel(H, c, R)


def uncertain_linearize(f: NonlinearModel, x: MVNSqrt):
    """EK1 with ``cholR = tria(H L)`` (``observations.py:43-49``)."""
    return _linearize_model(f, x, _lib.LIN_METHODS["uncertain"], need_chol=True)


def linearize_ek0(f: NonlinearModel, x):
    """EK0 (``observations.py:52-61``): ``H = E1`` of the IWP without the preconditioner, ``cholR = 0``."""
    return _linearize_model(f, x, _lib.LIN_METHODS["ek0"])


def linearize_regularized(f: NonlinearModel, x, l):
    """(``observations.py:64-89``): ``H = [J; I]``, ``b = [r(m); -m]``, ``cholR = blockdiag(0, I / sqrt(l))``."""
    return _linearize_model(f, x, _lib.LIN_METHODS["regularized"], params=[float(l), 1.0, 0.0, np.nan])
