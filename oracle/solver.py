"""IEKS driver: ``ieks_step``, ``crit``, ``solve``, ``sequential_eks_solve`` (oracle).

Restates ``pof/step.py:33-45``, ``pof/convergence_criteria.py:4-13`` and
``pof/solver.py:11-96`` (including the double calibration of the returned
factors, ``step.py:42-44`` + ``solver.py:66-68``).
"""
import numpy as np

from . import parallel, sequential
from .linalg import mv
from .model import MVNSqrt, get_initial_trajectory, linearize_at_previous_states, set_up_solver


def ieks_step(setup, states, calibrate=True, is_sequential=False):
    """(``pof/step.py:33-45``): ``(states, nll, obj, ssq)``."""
    om = linearize_at_previous_states(setup, states)
    fs = sequential.linear_filtsmooth if is_sequential else parallel.linear_filtsmooth
    out, nll, obj, ssq = fs(setup.x0, setup.F, setup.QL, om)
    if calibrate:
        out = MVNSqrt(out.mean, ssq**0.5 * out.chol)
    return out, nll, obj, ssq


def isclose(a, b, rtol, atol):
    """``jnp.isclose`` for finite inputs: ``|a - b| <= atol + rtol * |b|`` (NaN -> False)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    with np.errstate(invalid="ignore"):
        out = np.abs(a - b) <= atol + rtol * np.abs(b)
    both_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    return out | both_inf


def crit(obj, obj_old, nll, nll_old, states, states_old, rtol=1e-6, atol=1e-9):
    """(``pof/convergence_criteria.py:4-13``); ``means`` test uses rtol=1e-13, atol=1e-8."""
    isnan = np.isnan(obj) or np.isnan(nll)
    obj_conv = bool(isclose(obj_old, obj, rtol, atol))
    means_conv = bool(np.all(isclose(states_old.mean, states.mean, 1e-13, 1e-8)))
    return isnan or obj_conv or means_conv


def solve(problem, ts, order, init="prior", calibrate=True, maxiters=10_000, is_sequential=False,
          fixed_iters=None, return_history=False):
    """(``pof/solver.py:11-73``): ``(MVNSqrt(mean[N, d], chol[N, d, D]), info)``.

    ``fixed_iters`` (not in the reference) runs exactly that many iterations
    without the convergence test, for timing-mode comparisons.
    """
    setup = set_up_solver(problem, ts, order)
    states = get_initial_trajectory(setup, method=init)
    nll = obj = ssq = 0.0
    k = 0
    history = []
    while True:
        if fixed_iters is not None:
            if k >= fixed_iters:
                break
        elif k >= 1:
            if not (k <= maxiters) or crit(obj, obj_old, nll, nll_old, states, states_old):
                break
        states_old, nll_old, obj_old = states, nll, obj
        states, nll, obj, ssq = ieks_step(setup, states_old, calibrate=True, is_sequential=is_sequential)
        k += 1
        if return_history:
            history.append((states, nll, obj, ssq))
    info = {"iterations": float(k), "nll": nll, "obj": obj, "sigma_squared": ssq, "calibrated": False}
    if calibrate:
        states = MVNSqrt(states.mean, ssq**0.5 * states.chol)
        info["calibrated"] = True
    ys = MVNSqrt(mv(setup.E0, states.mean), setup.E0 @ states.chol)
    if return_history:
        info["history"] = history
        info["setup"] = setup
        info["states"] = states
    return ys, info


def sequential_eks_solve(problem, ts, order, return_full_states=False, calibrate=True):
    """(``pof/solver.py:76-96``): single-pass sequential EKS, calibrated once."""
    setup = set_up_solver(problem, ts, order)
    states, nll, obj, ssq = sequential.filtsmooth(setup.x0, setup.F, setup.QL, setup, setup.Ns)
    info = {"nll": nll, "obj": obj, "sigma_squared": ssq, "calibrated": False}
    if calibrate:
        states = MVNSqrt(states.mean, ssq**0.5 * states.chol)
        info["calibrated"] = True
    A = setup.P if return_full_states else setup.E0
    return MVNSqrt(mv(A, states.mean), A @ states.chol), info


def jax_linspace(start, stop, num):
    """``jnp.linspace(start, stop, num)`` (jax 0.3.23, endpoint=True): ``start (1 - k/div) + stop k/div``
    for ``k < div = num - 1``, then ``stop`` (the rounding differs from numpy's ``start + k step``)."""
    div = num - 1
    step = np.arange(div, dtype=np.float64) / div
    return np.concatenate([start * (1.0 - step) + stop * step, [stop]])


def coarse_ekf_init(problem, q, ts, N=10):
    """(``pof/initialization.py:112-129``): EKS on a coarse grid, nearest-left lookup.

    ``coarse_ts = jnp.linspace(ts[0], ts[-1], N)``; ``idx = floor(ts / coarse_dt)``
    (not ``ts - ts[0]``, as in the reference); the gather ``l[idx]`` counts a
    negative index from the end and clamps out-of-range indices (JAX indexing).
    """
    ts = np.asarray(ts, dtype=np.float64)
    coarse_ts = jax_linspace(ts[0], ts[-1], N)
    coarse_dt = coarse_ts[1] - coarse_ts[0]
    out, _ = sequential_eks_solve(problem, coarse_ts, q, return_full_states=True)
    idxs = np.floor(ts / coarse_dt).astype(np.int64)
    idxs = np.where(idxs < 0, idxs + N, idxs)
    idxs = np.clip(idxs, 0, N - 1)
    return MVNSqrt(out.mean[idxs], out.chol[idxs])
