"""Alternative linearisations of the residual model (oracle).

Test infrastructure only: restates ``pof/observations.py:43-89``
(``uncertain_linearize``, ``linearize_ek0``, ``linearize_regularized``),
``pof/linearization/unscented.py:19-70`` (``linearize_unscented`` and its
sigma points) and ``pof/utils.py:115-185`` (``cholesky_update_many``).  Each
works on one state ``(m[D], L[D, D])`` of the preconditioned residual model
``r(x) = E1 x - f(E0 x)`` of ``setup`` (``convenience.py:28``).
"""
import numpy as np

from .linalg import solve_triangular, tria
from .model import AffineModel, linearize, projection_matrix, residual


def linearize_ek0(setup, m):
    """(``observations.py:52-61``): ``H = E1`` of the IWP without the preconditioner
    (``projection_matrix(IWP(d, q), 1)``, not ``E1 P``), ``b = r(m) - H m``, ``cholR = 0``."""
    res = residual(setup, m)
    H = projection_matrix(setup.iwp, 1)
    d = res.shape[-1]
    return AffineModel(H, res - H @ m, np.zeros((d, d)))


def uncertain_linearize(setup, m, L):
    """(``observations.py:43-49``): EK1 ``H, b`` and ``cholR = tria(H L)``."""
    om = linearize(setup, m)
    return AffineModel(om.H, om.b, tria(om.H @ L))


def linearize_regularized(setup, m, lam):
    """(``observations.py:64-89``): ``H = [J; I_D]``, ``b = [r(m); -m]`` (``r(m)``, not
    ``r(m) - J m``, as written there), ``cholR = blockdiag(0_d, I_D / sqrt(lam))``."""
    om = linearize(setup, m)
    J = om.H
    res = residual(setup, m)
    d, D = J.shape
    H = np.concatenate([J, np.eye(D)], axis=0)
    b = np.concatenate([res, -m])
    R = np.zeros((d + D, d + D))
    R[d:, d:] = np.eye(D) / np.sqrt(lam)
    return AffineModel(H, b, R)


def _cholesky_update(chol, v, multiplier):
    """(``utils.py:124-175``): Algorithm 3.1 of the rank-one Cholesky update as one loop over columns."""
    n = chol.shape[0]
    diag = np.diag(chol).copy()
    omega = np.array(v, dtype=np.float64)
    b = 1.0
    new_diag = np.zeros(n)
    new_cols = np.zeros((n, n))
    for i in range(n):
        col = chol[:, i]
        dm = diag[i]
        w = omega[i]
        nd = np.sqrt(dm * dm + multiplier / b * (w * w))
        gamma = dm * dm * b + multiplier * (w * w)
        omega = omega - (w / dm) * col
        new_cols[i] = nd * (col / dm + (multiplier * w / gamma) * omega)
        b = b + multiplier * (w / dm) ** 2
        new_diag[i] = nd
    out = new_cols.T.copy()
    out[np.diag_indices(n)] = new_diag
    out[np.triu_indices(n, 1)] = 0.0
    return np.where(np.isfinite(out), out, 0.0)


def cholesky_update_many(chol, vectors, multiplier):
    """(``utils.py:115-121``): successive rank-one updates ``L L^T + multiplier v v^T``."""
    with np.errstate(all="ignore"):
        for v in vectors:
            chol = _cholesky_update(chol, v, multiplier)
    return chol


def unscented_weights(n, alpha, beta, kappa):
    """(``unscented.py:56-63``)"""
    lam = alpha**2 * (n + kappa) - n
    wm = np.full(2 * n + 1, 1.0 / (2.0 * (n + lam)))
    wm[0] = lam / (n + lam)
    wc = wm.copy()
    wc[0] = lam / (n + lam) + (1.0 - alpha**2 + beta)
    return wm, wc, lam


def sigma_points(m, L, alpha, beta, kappa):
    """(``unscented.py:42-53``): ``[m, m + c_j, m - c_j]`` with ``c_j`` the columns of ``sqrt(n + lam) L``."""
    n = m.shape[0]
    if kappa is None:
        kappa = 3.0 + n
    wm, wc, lam = unscented_weights(n, alpha, beta, kappa)
    sc = np.sqrt(n + lam) * L
    pts = m[None, :] + np.concatenate([np.zeros((1, n)), sc.T, -sc.T], axis=0)
    return pts, wm, wc


def linearize_unscented(setup, m, L, alpha=1.0, beta=0.0, kappa=None):
    """(``unscented.py:19-39``): statistical linear regression through the sigma points.

    ``F = (L L^T)^-1 Psi`` transposed, ``b = m_f - F m``, ``cholR`` = the
    Cholesky downdate of ``tria(sqrt(wc) (f_i - m_f))`` by the columns of ``F L``.
    """
    pts, wm, wc = sigma_points(m, L, alpha, beta, kappa)
    f_pts = residual(setup, pts)
    m_f = wm @ f_pts
    Psi = ((pts - m[None, :]).T * wc[None, :]) @ (f_pts - m_f[None, :])
    # cho_solve((L, True), Psi): (L L^T) X = Psi
    X = solve_triangular(L, solve_triangular(L, Psi, lower=True), lower=True, trans=True)
    F = X.T
    sqrt_phi = np.sqrt(wc)[:, None] * (f_pts - m_f[None, :])
    chol = tria(sqrt_phi.T)
    chol = cholesky_update_many(chol, (F @ L).T, -1.0)
    return AffineModel(F, m_f - F @ m, chol)
