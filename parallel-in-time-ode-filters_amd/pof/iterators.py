"""IEKS as a generator of iterates (reference ``pof/iterators.py:20-41``).

Each ``next()`` runs one IEKS iteration on the GPU (``pof.step.ieks_step``:
linearise at the previous means, parallel filter and smoother, factors
scaled by ``sqrt(sigma_squared)``) and yields
``(states, nll, obj, sigma_squared)``; the generator stops after the
iteration at which the convergence rule of ``pof.solver.solve`` holds
(``pof/convergence_criteria.py:4-13``), so it yields exactly as many iterates
as ``solve`` runs iterations.
"""
from .convenience import get_initial_trajectory, set_up_solver
from .convergence_criteria import crit
from .step import ieks_step


def ieks_iterator(*, f, y0, ts, order, init="prior"):
    """Returns ``(iterator, setup)`` like the reference; ``states`` are preconditioned device arrays."""
    setup = set_up_solver(f=f, y0=y0, ts=ts, order=order)
    states = get_initial_trajectory(setup, method=init)
    return _ieks_iterator(setup, states), setup


def _ieks_iterator(setup, init_traj):
    states, nll, obj, ssq = ieks_step(states=init_traj, setup=setup)
    yield states, nll, obj, ssq
    while True:
        nll_old, obj_old, states_old = nll, obj, states
        states, nll, obj, ssq = ieks_step(states=states_old, setup=setup)
        yield states, nll, obj, ssq
        if crit(obj, obj_old, nll, nll_old, states, states_old):
            break
