"""Noiseless observation models and EK1 linearisation (reference ``pof/observations.py``)."""
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
