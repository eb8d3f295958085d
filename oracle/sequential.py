"""Sequential square-root Kalman filter / EKF and RTS smoother (oracle).

Restates ``pof/sequential_filtsmooth/{__init__,filter,smoother}.py``.  The
step functions are vectorised over a leading batch axis so one call can run
many independent models; the time recursion is a Python loop.
"""
import numpy as np

from .linalg import mT, mv, solve_triangular, tria
from .model import MVNSqrt, linearize
from .parallel import mvn_loglikelihood, objective, whiten


def sqrt_predict(F, QL, m, L):
    """(``filter.py:59-66``)"""
    return mv(F, m), tria(np.concatenate([F @ L, np.broadcast_to(QL, L.shape)], axis=-1))


def sqrt_update(H, cholR, c, m, L):
    """(``filter.py:69-92``): returns ``(m, L, ell, ssq)``."""
    nx = m.shape[-1]
    ny = c.shape[-1]
    y_hat = mv(H, m) + c
    y_diff = 0 - y_hat
    Mblk = np.concatenate(
        [
            np.concatenate([H @ L, cholR], axis=-1),
            np.concatenate([L, np.zeros(L.shape[:-2] + (nx, ny))], axis=-1),
        ],
        axis=-2,
    )
    S = tria(Mblk)
    Lnew = S[..., ny:, ny:]
    G = S[..., ny:, :ny]
    I = S[..., :ny, :ny]
    wres = whiten(y_diff, I)
    ssq = np.sum(wres * wres, -1) / ny
    mnew = m + mv(G, solve_triangular(I, y_diff, lower=True))
    ell = mvn_loglikelihood(y_diff, I)
    return mnew, Lnew, ell, ssq


def linear_noiseless_filter(x0, F, QL, om):
    """(``filter.py:33-56``): ``(MVNSqrt, ell, ssq)``; ``ell`` is +sum loglik."""
    H, c, cholR = om
    N = c.shape[0]
    m, L = x0.mean, x0.chol
    ms, Ls = [m], [L]
    ssq = 0.0
    ell = 0.0
    for k in range(N):
        m, L = sqrt_predict(F, QL, m, L)
        m, L, e, s = sqrt_update(H[k], cholR[k], c[k], m, L)
        ssq += s
        ell += e
        ms.append(m)
        Ls.append(L)
    return MVNSqrt(np.stack(ms), np.stack(Ls)), float(ell), float(ssq / N)


def extended_kalman_filter(x0, F, QL, setup, N):
    """EKF linearising at the predicted state (``filter.py:9-30``)."""
    m, L = x0.mean, x0.chol
    ms, Ls = [m], [L]
    ssq = 0.0
    ell = 0.0
    for _ in range(N):
        m, L = sqrt_predict(F, QL, m, L)
        H, c, cholR = linearize(setup, m)
        m, L, e, s = sqrt_update(H, cholR, c, m, L)
        ssq += s
        ell += e
        ms.append(m)
        Ls.append(L)
    return MVNSqrt(np.stack(ms), np.stack(Ls)), float(ell), float(ssq / N)


def sqrt_smooth(F, QL, mf, Lf, ms, Ls):
    """(``smoother.py:31-48``)"""
    nx = F.shape[-1]
    Phi = np.concatenate(
        [
            np.concatenate([F @ Lf, np.broadcast_to(QL, Lf.shape)], axis=-1),
            np.concatenate([Lf, np.zeros_like(Lf)], axis=-1),
        ],
        axis=-2,
    )
    T = tria(Phi)
    Phi11 = T[..., :nx, :nx]
    Phi21 = T[..., nx:, :nx]
    Phi22 = T[..., nx:, nx:]
    gain = mT(solve_triangular(Phi11, mT(Phi21), lower=True, trans=True))
    mean = mf + mv(gain, ms - mv(F, mf))
    chol = tria(np.concatenate([Phi22, gain @ Ls], axis=-1))
    return mean, chol


def smoothing(F, QL, filt):
    """Reverse-time RTS (``smoother.py:8-28``): ``(MVNSqrt, obj)``."""
    N = filt.mean.shape[0]
    m, L = filt.mean[-1], filt.chol[-1]
    ms, Ls = [m], [L]
    for k in range(N - 2, -1, -1):
        m, L = sqrt_smooth(F, QL, filt.mean[k], filt.chol[k], m, L)
        ms.append(m)
        Ls.append(L)
    means = np.stack(ms[::-1])
    chols = np.stack(Ls[::-1])
    return MVNSqrt(means, chols), objective(F, QL, means)


def linear_filtsmooth(x0, F, QL, om):
    """(``__init__.py:13-16``)"""
    out, nll, ssq = linear_noiseless_filter(x0, F, QL, om)
    out, obj = smoothing(F, QL, out)
    return out, nll, obj, ssq


def filtsmooth(x0, F, QL, setup, N):
    """(``__init__.py:5-10``)"""
    out, nll, ssq = extended_kalman_filter(x0, F, QL, setup, N)
    out, obj = smoothing(F, QL, out)
    return out, nll, obj, ssq
