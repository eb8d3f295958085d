"""Initial state and initial trajectories (reference ``pof/initialization.py``)."""
import ctypes

import numpy as np

from . import _lib
from .convenience import ProblemSpec
from .utils import MVNSqrt, as_f64_device


def taylor_mode_init(f, y0, num_derivatives):
    """``x0`` from Taylor-mode derivatives, not preconditioned (``initialization.py:15-22``).

    Returns host arrays ``(mean[D], chol[D, D])``; the factor is zero.
    """
    spec = ProblemSpec(f, y0, np.array([0.0, 1.0]), num_derivatives)
    D = spec.D
    F = np.empty((D, D))
    QL = np.empty((D, D))
    P = np.empty(D)
    PI = np.empty(D)
    m0 = np.empty(D)
    L0 = np.empty((D, D))
    prob = spec.cstruct()
    _lib.check(
        _lib.load().pof_setup(ctypes.byref(prob), _lib.hptr(F), _lib.hptr(QL), _lib.hptr(P), _lib.hptr(PI),
                              _lib.hptr(m0), _lib.hptr(L0)),
        what="pof_setup",
    )
    # pof_setup returns PI * x0 with dt = 1; at dt = 1 the scaling is (q-i)!
    return MVNSqrt(m0 / PI, L0)


def _initial(method, f, y0, order, ts):
    spec = ProblemSpec(f, y0, ts, order, init=method)
    m = np.empty((spec.ts.shape[0], spec.D))
    prob = spec.cstruct()
    _lib.check(_lib.load().pof_initial_means(ctypes.byref(prob), _lib.hptr(m)), what="pof_initial_means")
    return m


def constant_init(*, y0, order, ts, f):
    """``[y0, f(y0), 0, ...]`` repeated, zero factors, not preconditioned (``initialization.py:42-56``)."""
    import torch

    m = _initial("constant", f, y0, order, ts)
    spec = ProblemSpec(f, y0, ts, order)
    t = np.asarray(spec.ts)
    # pof_initial_means returns PI * x; undo PI for the reference's un-preconditioned constant_init
    from .transitions import IWP, nordsieck_preconditioner

    _, PI = nordsieck_preconditioner(IWP(spec.d, spec.order), t[1] - t[0])
    m = m / np.diag(PI)[None, :]
    mean = as_f64_device(m)
    return MVNSqrt(mean, torch.zeros(mean.shape + (spec.D,), dtype=torch.float64, device=mean.device))


def coarse_ekf_init(*, f, y0, order, ts, N=100):
    """Sequential EKS on ``N = 100`` equispaced points, looked up at ``floor(ts / coarse_dt)``
    (``initialization.py:112-129``); returns ``P``-scaled means (not ``PI``-multiplied), zero factors.

    The reference's default is ``N=10`` but its only caller passes ``N=100``
    (``convenience.py:81``); that is the size the device path implements.
    """
    import torch

    if N != 100:
        raise NotImplementedError("the coarse EKS runs on N = 100 points (convenience.py:81)")
    spec = ProblemSpec(f, y0, ts, order, init="coarse")
    m = np.empty((spec.ts.shape[0], spec.D))
    prob = spec.cstruct()
    ctx = _lib.context(torch.cuda.current_device())
    _lib.check(_lib.load().pof_coarse_means(ctx, ctypes.byref(prob), _lib.hptr(m)), ctx, "pof_coarse_means")
    from .transitions import IWP, nordsieck_preconditioner

    _, PI = nordsieck_preconditioner(IWP(spec.d, spec.order), spec.ts[1] - spec.ts[0])
    m = m / np.diag(PI)[None, :]
    mean = as_f64_device(m)
    return MVNSqrt(mean, torch.zeros(mean.shape + (spec.D,), dtype=torch.float64, device=mean.device))


def prior_init(*, f, y0, order, ts):
    """Prior predictions of x0 to every ``t_k`` with steps ``ts[1:]`` (``initialization.py:83-89``); means only."""
    import torch

    m = _initial("prior", f, y0, order, ts)
    mean = as_f64_device(m)
    D = mean.shape[1]
    return MVNSqrt(mean, torch.zeros(mean.shape + (D,), dtype=torch.float64, device=mean.device))
