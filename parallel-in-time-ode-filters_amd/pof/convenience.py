"""Solver set-up (reference ``pof/convenience.py``): problem descriptor, transitions, x0, initial trajectory."""
import ctypes

import numpy as np

from . import _lib
from .ivp import VectorField
from .observations import AffineModel, NonlinearModel, linearize
from .transitions import IWP, TransitionModel, discretize_transitions, projection_matrix
from .utils import MVNSqrt, as_f64_device, as_f64_host

INIT_METHODS = {"prior": _lib.INIT_PRIOR, "constant": _lib.INIT_CONSTANT, "coarse": _lib.INIT_COARSE}


class ProblemSpec:
    """Host-side description of one IVP solve, convertible to the C ``pof_problem``."""

    def __init__(self, f, y0, ts, order, init="prior"):
        if not isinstance(f, VectorField):
            raise TypeError(
                "f must be a device vector field from pof.ivp (e.g. pof.ivp.fitzhughnagumo().f); "
                f"got {type(f).__name__}"
            )
        self.f = f
        self.y0 = as_f64_host(y0).reshape(-1)
        self.ts = as_f64_host(ts).reshape(-1)
        self.order = int(order)
        if self.y0.shape[0] != f.d:
            raise ValueError(f"y0 has dimension {self.y0.shape[0]}, {f.name} needs {f.d}")
        if self.ts.shape[0] < 2:
            raise ValueError("ts needs at least two time points")
        if init not in INIT_METHODS:
            raise Exception(f"method={init} not found")
        self.init = init
        self.d = f.d
        self.D = self.d * (self.order + 1)

    @property
    def Ns(self):
        return self.ts.shape[0] - 1

    def cstruct(self):
        p = _lib.Problem()
        p.kind = self.f.kind
        p.d = self.d
        p.q = self.order
        p.init = INIT_METHODS[self.init]
        p.n_times = self.ts.shape[0]
        p.ts_host = _lib.hptr(self.ts)
        p.y0_host = _lib.hptr(self.y0)
        for i, v in enumerate(self.f.params_array()):
            p.params[i] = v
        return p


def set_up_solver(*, f, y0, ts, order):
    """(``convenience.py:13-45``): transitions, preconditioner, projections, residual model, x0.

    Returns the reference's dictionary keys; matrices are host float64 arrays,
    ``dtm`` is the preconditioned transition pair repeated for every step, ``om`` the
    residual model ``E1 x - f(E0 x)`` and ``x0`` the preconditioned Taylor-mode
    initial state.
    """
    spec = ProblemSpec(f, y0, ts, order)
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
    iwp = IWP(wiener_process_dimension=spec.d, num_derivatives=spec.order)
    E0 = projection_matrix(iwp, 0) * P[None, :]
    E1 = projection_matrix(iwp, 1) * P[None, :]
    return {
        "f": f,
        "y0": spec.y0,
        "ts": spec.ts,
        # one (F, QL) per step, the same pair repeated (convenience.py:19), as read-only views
        "dtm": TransitionModel(np.broadcast_to(F, (spec.Ns, D, D)), np.broadcast_to(QL, (spec.Ns, D, D))),
        "om": NonlinearModel(f, E0, E1),
        "x0": MVNSqrt(m0, L0),
        "E0": E0,
        "P": np.diag(P),
        "PI": np.diag(PI),
        "order": spec.order,
        "iwp": iwp,
        "spec": spec,
    }


def get_initial_trajectory(setup, method="prior"):
    """(``convenience.py:76-92``): device means ``[len(ts), D]`` and zero factors.

    ``constant`` and ``coarse`` are multiplied by ``PI``; ``prior`` is returned
    without preconditioning, as in the reference.  ``coarse`` runs the
    sequential EKS on 100 points on the GPU.  Only the means enter the IEKS.
    """
    import torch

    spec = setup["spec"]
    if method not in INIT_METHODS:
        raise Exception(f"method={method} not found")
    spec = ProblemSpec(spec.f, spec.y0, spec.ts, spec.order, init=method)
    m = np.empty((spec.ts.shape[0], spec.D))
    prob = spec.cstruct()
    if method == "coarse":
        ctx = _lib.context(torch.cuda.current_device())
        _lib.check(_lib.load().pof_coarse_means(ctx, ctypes.byref(prob), _lib.hptr(m)), ctx, "pof_coarse_means")
    else:
        _lib.check(_lib.load().pof_initial_means(ctypes.byref(prob), _lib.hptr(m)), what="pof_initial_means")

    mean = as_f64_device(m)
    return MVNSqrt(mean, torch.zeros(mean.shape + (spec.D,), dtype=torch.float64, device=mean.device))


def linearize_observation_model(observation_model, trajectory):
    """vmap of ``linearize`` over a trajectory (``convenience.py:9-10``)."""
    return linearize(observation_model, trajectory)
