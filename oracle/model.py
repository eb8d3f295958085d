"""IWP prior, Nordsieck preconditioner, solver setup, initial trajectories, EK1 (oracle).

Restates ``pof/transitions.py``, ``pof/convenience.py``,
``pof/initialization.py`` and ``pof/observations.py:35-40`` in float64 NumPy.
"""
import math
from typing import NamedTuple

import numpy as np
import scipy.linalg

from .linalg import mv, tria
from .problems import taylor_mode


class MVNSqrt(NamedTuple):
    """Gaussian in square-root form (``pof/utils.py:9-11``)."""

    mean: np.ndarray
    chol: np.ndarray


class TransitionModel(NamedTuple):
    """``x' | x ~ N(F x, QL QL^T)`` (``pof/transitions.py:15-19``)."""

    F: np.ndarray
    QL: np.ndarray


class AffineModel(NamedTuple):
    """Noiseless affine observation ``y = H x + b`` (``pof/observations.py:23-32``)."""

    H: np.ndarray
    b: np.ndarray
    cholR: np.ndarray


class IWP(NamedTuple):
    """(``pof/transitions.py:22-24``)"""

    wiener_process_dimension: int
    num_derivatives: int


def preconditioned_discretize_1d(q):
    """``A = flip(pascal_lower(q+1))``, ``L_Q = chol(flip(hilbert(q+1)))`` (``transitions.py:37-41``)."""
    A = np.flip(scipy.linalg.pascal(q + 1, kind="lower", exact=False))
    Q = np.flip(scipy.linalg.hilbert(q + 1))
    return np.ascontiguousarray(A), np.linalg.cholesky(np.ascontiguousarray(Q))


def preconditioned_discretize(iwp):
    """``kron(I_d, .)`` of the 1-d pair (``transitions.py:44-50``)."""
    A1, L1 = preconditioned_discretize_1d(iwp.num_derivatives)
    I = np.eye(iwp.wiener_process_dimension)
    return np.kron(I, A1), np.kron(I, L1)


def nordsieck_preconditioner_1d(q, dt):
    """``P = diag(|dt|^(q-i+1/2) / (q-i)!)``, ``PI = P^-1`` (``transitions.py:53-61``)."""
    powers = np.arange(q, -1, -1)
    scales = np.array([float(math.factorial(int(p))) for p in powers])
    powers = powers + 0.5
    return np.diag(np.abs(dt) ** powers / scales), np.diag(np.abs(dt) ** (-powers) * scales)


def nordsieck_preconditioner(iwp, dt):
    """(``transitions.py:64-68``)"""
    P, PI = nordsieck_preconditioner_1d(iwp.num_derivatives, dt)
    I = np.eye(iwp.wiener_process_dimension)
    return np.kron(I, P), np.kron(I, PI)


def non_preconditioned_discretize(iwp, dt):
    """``(P F PI, P L_Q)`` for one step ``dt`` (``transitions.py:71-77``)."""
    P, PI = nordsieck_preconditioner(iwp, dt)
    F, QL = preconditioned_discretize(iwp)
    return P @ F @ PI, P @ QL


def projection_matrix(iwp, k):
    """``E_k = kron(I_d, e_k^T)`` (``transitions.py:80-88``)."""
    e = np.eye(1, iwp.num_derivatives + 1, k)
    return np.kron(np.eye(iwp.wiener_process_dimension), e)


def discretize_transitions(iwp, times=None, steps=None):
    """Stacked non-preconditioned transitions (``transitions.py:96-100``)."""
    if steps is None:
        times = np.asarray(times, dtype=np.float64)
        steps = times[1:] - times[:-1]
    pairs = [non_preconditioned_discretize(iwp, float(h)) for h in np.asarray(steps)]
    return TransitionModel(np.stack([p[0] for p in pairs]), np.stack([p[1] for p in pairs]))


# --------------------------------------------------------------------------
# initialisation (pof/initialization.py)
# --------------------------------------------------------------------------
def taylor_mode_init(problem, q):
    """``x0 = (concat(m0^T), kron(I_d, P0))`` with P0 = 0 (``initialization.py:15-22``)."""
    m0 = taylor_mode(problem, q)  # [q+1, d]
    d = problem.d
    P0 = np.zeros((q + 1, q + 1))
    return MVNSqrt(np.concatenate(m0.T), np.kron(np.eye(d), P0))


def constant_init(problem, q, ts):
    """``[y0, f(y0), 0, ...]`` per dimension, repeated, zero chol (``initialization.py:42-56``)."""
    y0 = problem.y0
    d = problem.d
    dy0 = problem.f(None, y0)
    x0 = np.concatenate([y0[:, None], dy0[:, None], np.zeros((d, q - 1))], axis=1).reshape(1, -1)
    N = len(ts)
    traj = np.repeat(x0, N, axis=0)
    D = traj.shape[1]
    return MVNSqrt(traj, np.zeros((N, D, D)))


def sqrt_predict(F, QL, x):
    """(``pof/sequential_filtsmooth/filter.py:59-66``)"""
    m, L = x
    return MVNSqrt(mv(F, m), tria(np.concatenate([F @ L, QL], axis=-1)))


def prior_init(problem, q, ts):
    """Predict the (non-preconditioned) x0 to every ``t_k`` with ``steps = ts[1:]``.

    Follows ``initialization.py:75-89``: the result is not multiplied by ``PI``
    (``convenience.py:88-90``).
    """
    x0 = taylor_mode_init(problem, q)
    iwp = IWP(problem.d, q)
    dtm = discretize_transitions(iwp, steps=np.asarray(ts, dtype=np.float64)[1:])
    pred = sqrt_predict(dtm.F, dtm.QL, MVNSqrt(x0.mean[None], x0.chol[None]))
    return MVNSqrt(
        np.concatenate([x0.mean[None], pred.mean]),
        np.concatenate([x0.chol[None], pred.chol]),
    )


# --------------------------------------------------------------------------
# setup (pof/convenience.py:13-45) and initial trajectory (76-92)
# --------------------------------------------------------------------------
class Setup(NamedTuple):
    problem: object
    ts: np.ndarray
    q: int
    iwp: IWP
    F: np.ndarray
    QL: np.ndarray
    P: np.ndarray
    PI: np.ndarray
    E0: np.ndarray  # E0_raw @ P
    E1: np.ndarray  # E1_raw @ P
    x0: MVNSqrt  # preconditioned

    @property
    def Ns(self):
        return len(self.ts) - 1

    @property
    def dtm(self):
        Ns = self.Ns
        return TransitionModel(
            np.broadcast_to(self.F, (Ns,) + self.F.shape), np.broadcast_to(self.QL, (Ns,) + self.QL.shape)
        )


def set_up_solver(problem, ts, q):
    """(``convenience.py:13-45``); ``dt = ts[1] - ts[0]`` for the whole grid (uniform grid assumed)."""
    ts = np.asarray(ts, dtype=np.float64)
    dt = ts[1] - ts[0]
    iwp = IWP(problem.d, q)
    F, QL = preconditioned_discretize(iwp)
    P, PI = nordsieck_preconditioner(iwp, dt)
    E0 = projection_matrix(iwp, 0) @ P
    E1 = projection_matrix(iwp, 1) @ P
    x0 = taylor_mode_init(problem, q)
    x0 = MVNSqrt(PI @ x0.mean, PI @ x0.chol)
    return Setup(problem, ts, q, iwp, F, QL, P, PI, E0, E1, x0)


def get_initial_trajectory(setup, method="prior"):
    """(``convenience.py:76-92``); ``coarse`` needs the sequential EKS (see solver.coarse_ekf_init)."""
    if method == "constant":
        s = constant_init(setup.problem, setup.q, setup.ts)
        return MVNSqrt(mv(setup.PI, s.mean), setup.PI @ s.chol)
    if method == "prior":
        return prior_init(setup.problem, setup.q, setup.ts)
    if method == "coarse":
        from .solver import coarse_ekf_init

        s = coarse_ekf_init(setup.problem, setup.q, setup.ts, N=100)
        return MVNSqrt(mv(setup.PI, s.mean), setup.PI @ s.chol)
    raise Exception(f"method={method} not found")


# --------------------------------------------------------------------------
# EK1 (pof/observations.py:35-40; pof/step.py:12-22)
# --------------------------------------------------------------------------
def residual(setup, m):
    """``om(x) = E1 x - f(E0 x)`` (``convenience.py:28``)."""
    return mv(setup.E1, m) - setup.problem.f(None, mv(setup.E0, m))


def linearize(setup, m):
    """EK1 at means ``m[..., D]``: ``H = E1 - J_f(E0 m) E0``, ``b = r(m) - H m``, ``cholR = 0``."""
    y = mv(setup.E0, m)
    J = setup.problem.jac(y)
    H = setup.E1 - J @ setup.E0
    res = residual(setup, m)
    b = res - mv(H, m)
    d = setup.problem.d
    return AffineModel(H, b, np.zeros(m.shape[:-1] + (d, d)))


def linearize_at_previous_states(setup, states):
    """Linearise at ``states.mean[1:]`` (``step.py:12-22``)."""
    return linearize(setup, states.mean[1:])
