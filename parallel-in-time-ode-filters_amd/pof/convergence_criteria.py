"""IEKS stopping rule (reference ``pof/convergence_criteria.py:4-13``).

The solver evaluates this rule inside ``libpof_hip.so`` after every
iteration; this function is the same rule on tensors for callers that drive
``ieks_step`` themselves.
"""
import math


def _isclose_scalar(a, b, rtol, atol):
    if math.isnan(a) or math.isnan(b):
        return False
    if math.isinf(a) or math.isinf(b):
        return a == b
    return abs(a - b) <= atol + rtol * abs(b)


def crit(obj, obj_old, nll, nll_old, states, states_old, *, rtol=1e-6, atol=1e-9):
    """NaN, or ``isclose(obj_old, obj, rtol, atol)``, or all means close at ``rtol=1e-13, atol=1e-8``."""
    import torch

    obj, obj_old, nll = float(obj), float(obj_old), float(nll)
    isnan = math.isnan(obj) or math.isnan(nll)
    obj_converged = _isclose_scalar(obj_old, obj, rtol, atol)
    m, mo = states.mean, states_old.mean
    means_converged = bool(torch.isclose(torch.as_tensor(mo), torch.as_tensor(m), rtol=1e-13, atol=1e-8).all())
    return isnan or obj_converged or means_converged
