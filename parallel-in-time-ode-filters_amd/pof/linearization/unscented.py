"""Unscented (sigma-point) linearisation on the GPU (reference ``pof/linearization/unscented.py``)."""
import numpy as np

from .. import _lib
from ..observations import NonlinearModel, _linearize_model
from ..utils import MVNSqrt


def linearize_unscented(f: NonlinearModel, x: MVNSqrt, alpha: float = 1.0, beta: float = 0.0, kappa: float = None):
    """Statistical linear regression through 2D + 1 sigma points (``unscented.py:19-39``).

    ``H = F`` with ``F^T = (L L^T)^-1 Psi``, ``b = m_f - F m``, ``cholR`` = the
    Cholesky downdate of ``tria(sqrt(wc) (f_i - m_f))`` by the columns of ``F L``
    (``utils.py:115-185``).  ``kappa=None`` means ``3 + D``.  Needs a
    nonsingular ``x.chol``.
    """
    params = [1.0, float(alpha), float(beta), np.nan if kappa is None else float(kappa)]
    return _linearize_model(f, x, _lib.LIN_METHODS["unscented"], params=params, need_chol=True)
