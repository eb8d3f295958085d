"""IWP prior transitions (reference ``pof/transitions.py``), computed by the native library."""
import ctypes
from typing import NamedTuple

import numpy as np

from . import _lib


class TransitionModel(NamedTuple):
    """Linear transition model in square-root form: x' | x ~ N(F x, QL QL^T) (``transitions.py:15-19``)."""

    F: np.ndarray
    QL: np.ndarray


class IWP(NamedTuple):
    """(``transitions.py:22-24``)"""

    wiener_process_dimension: int
    num_derivatives: int


def _native(iwp, dt):
    d, q = int(iwp.wiener_process_dimension), int(iwp.num_derivatives)
    D = d * (q + 1)
    F = np.empty((D, D))
    QL = np.empty((D, D))
    P = np.empty(D)
    PI = np.empty(D)
    _lib.check(
        _lib.load().pof_transitions(d, q, float(dt), _lib.hptr(F), _lib.hptr(QL), _lib.hptr(P), _lib.hptr(PI)),
        what="pof_transitions",
    )
    return F, QL, P, PI


def preconditioned_discretize(iwp: IWP):
    """``(kron(I_d, flip(pascal)), kron(I_d, chol(flip(hilbert))))`` (``transitions.py:44-50``)."""
    F, QL, _, _ = _native(iwp, 1.0)
    return F, QL


def nordsieck_preconditioner(iwp: IWP, dt: float):
    """``(P, PI)`` as dense diagonal matrices (``transitions.py:64-68``)."""
    _, _, P, PI = _native(iwp, dt)
    return np.diag(P), np.diag(PI)


def non_preconditioned_discretize(iwp: IWP, dt: float):
    """``(P F PI, P QL)`` (``transitions.py:71-77``)."""
    F, QL, P, PI = _native(iwp, dt)
    return (P[:, None] * F) * PI[None, :], P[:, None] * QL


def projection_matrix(iwp: IWP, derivative_to_project_onto: int):
    """``E_k = kron(I_d, e_k^T)`` (``transitions.py:80-88``)."""
    q1 = iwp.num_derivatives + 1
    return np.kron(np.eye(iwp.wiener_process_dimension), np.eye(1, q1, derivative_to_project_onto))


def get_transition_model(iwp: IWP, dt: float):
    """(``transitions.py:91-93``)"""
    return TransitionModel(*non_preconditioned_discretize(iwp, dt))


def discretize_transitions(iwp: IWP, times=None, steps=None):
    """Stacked non-preconditioned transitions (``transitions.py:96-100``)."""
    if steps is None:
        times = np.asarray(times, dtype=np.float64)
        steps = times[1:] - times[:-1]
    pairs = [non_preconditioned_discretize(iwp, float(h)) for h in np.asarray(steps, dtype=np.float64)]
    return TransitionModel(np.stack([p[0] for p in pairs]), np.stack([p[1] for p in pairs]))
