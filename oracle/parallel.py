"""Square-root parallel Kalman filter and RTS smoother (oracle).

Restates ``pof/parallel_filtsmooth/{__init__,filter,smoother}.py`` with the
element construction, the associative operators and ``jax.lax.associative_scan``'s
odd/even recursion order, vectorised over elements in float64.
"""
import numpy as np

from .linalg import mT, mv, solve_triangular, tria
from .model import MVNSqrt


# --------------------------------------------------------------------------
# jax.lax.associative_scan (odd/even recursion) over tuples of stacked arrays
# --------------------------------------------------------------------------
def associative_scan(fn, elems, reverse=False):
    """Inclusive scan with ``fn(earlier, later)`` in JAX's recursion order.

    ``reverse=True`` reverses the inputs, scans, and reverses the outputs, so
    ``fn`` then receives ``(later-in-time, earlier-in-time)`` aggregates.
    """
    elems = tuple(np.asarray(e) for e in elems)
    if reverse:
        elems = tuple(e[::-1] for e in elems)

    def _scan(el):
        n = el[0].shape[0]
        if n < 2:
            return el
        reduced = fn(tuple(e[0:-1:2] for e in el), tuple(e[1::2] for e in el))
        odd = _scan(reduced)
        if n % 2 == 0:
            even = fn(tuple(e[:-1] for e in odd), tuple(e[2::2] for e in el))
        else:
            even = fn(odd, tuple(e[2::2] for e in el))
        even = tuple(np.concatenate([e[:1], r]) for e, r in zip(el, even))
        out = []
        for ev, od in zip(even, odd):
            res = np.empty((ev.shape[0] + od.shape[0],) + ev.shape[1:], dtype=np.float64)
            res[0::2] = ev
            res[1::2] = od
            out.append(res)
        return tuple(out)

    out = _scan(elems)
    if reverse:
        out = tuple(e[::-1] for e in out)
    return out


# --------------------------------------------------------------------------
# filter (pof/parallel_filtsmooth/filter.py)
# --------------------------------------------------------------------------
def filter_elements(F, QL, H, c, cholR, ms, Ls):
    """``_get_element`` vectorised over k (``filter.py:50-81``).

    ``ms``/``Ls`` are the per-element priors: ``x0`` for k = 0 and zeros after.
    """
    nx = F.shape[-1]
    ny = c.shape[-1]
    Nk = H.shape[0]
    F = np.broadcast_to(F, (Nk, nx, nx))
    QL = np.broadcast_to(QL, (Nk, nx, nx))
    m1 = mv(F, ms)
    N1 = tria(np.concatenate([F @ Ls, QL], axis=-1))
    Psi_ = np.concatenate(
        [
            np.concatenate([H @ N1, cholR], axis=-1),
            np.concatenate([N1, np.zeros((Nk, nx, ny))], axis=-1),
        ],
        axis=-2,
    )
    Psi = tria(Psi_)
    Psi11 = Psi[:, :ny, :ny]
    Psi21 = Psi[:, ny:, :ny]
    U = Psi[:, ny:, ny:]
    K = mT(solve_triangular(Psi11, mT(Psi21), lower=True, trans=True))
    HF = H @ F
    A = F - K @ HF
    b = m1 + mv(K, -mv(H, m1) - c)
    Z = mT(solve_triangular(Psi11, HF, lower=True))
    eta = mv(mT(solve_triangular(Psi11, mT(Z), lower=True, trans=True)), -c)
    if nx > ny:
        Z = np.concatenate([Z, np.zeros((Nk, nx, nx - ny))], axis=-1)
    else:
        Z = tria(Z)
    return A, b, U, eta, Z


def sqrt_filtering_operator(e1, e2):
    """Combine (earlier ``e1``, later ``e2``) (``filter.py:117-142``), vectorised."""
    A1, b1, U1, eta1, Z1 = e1
    A2, b2, U2, eta2, Z2 = e2
    n = A1.shape[0]
    nx = Z2.shape[-1]
    I = np.broadcast_to(np.eye(nx), (n, nx, nx))
    Xi = np.concatenate(
        [
            np.concatenate([mT(U1) @ Z2, I], axis=-1),
            np.concatenate([Z2, np.zeros_like(A1)], axis=-1),
        ],
        axis=-2,
    )
    T = tria(Xi)
    Xi11 = T[:, :nx, :nx]
    Xi21 = T[:, nx : 2 * nx, :nx]
    Xi22 = T[:, nx : 2 * nx, nx:]
    M = solve_triangular(Xi11, mT(U1) @ mT(A2), lower=True)
    A = A2 @ A1 - mT(M) @ mT(Xi21) @ A1
    m = solve_triangular(Xi11, mT(U1), lower=True)
    b = mv(A2 @ (I - mT(m) @ mT(Xi21)), b1 + mv(U1 @ mT(U1), eta2)) + b2
    _U = mT(solve_triangular(Xi11, mT(U1) @ mT(A2), lower=True))
    U = tria(np.concatenate([_U, U2], axis=-1))
    _e = solve_triangular(Xi11, mT(Xi21), lower=True, trans=True)
    eta = mv(mT(A1) @ (I - mT(_e) @ mT(U1)), eta2 - mv(Z2 @ mT(Z2), b1)) + eta1
    Z = tria(np.concatenate([mT(A1) @ Xi22, Z1], axis=-1))
    return A, b, U, eta, Z


def get_obs(F, QL, H, c, cholR, m, L):
    """Predicted observation ``(H F m + c, tria([H tria([F L | QL]) | R]))`` (``filter.py:84-93``)."""
    pm = mv(F, m)
    pL = tria(np.concatenate([F @ L, np.broadcast_to(QL, L.shape)], axis=-1))
    om = mv(H, pm) + c
    oL = tria(np.concatenate([H @ pL, cholR], axis=-1))
    return om, oL


def mvn_loglikelihood(x, L):
    """(``pof/utils.py:22-30``)"""
    dim = L.shape[-1]
    y = solve_triangular(L, x, lower=True)
    nc = np.sum(np.log(np.abs(np.diagonal(L, axis1=-2, axis2=-1))), axis=-1) + dim * np.log(2 * np.pi) / 2.0
    return -0.5 * np.sum(y * y, -1) - nc


def whiten(m, L):
    """``solve_triangular(L^T, m)`` with the default ``lower=False`` (``pof/utils.py:110-112``)."""
    return solve_triangular(mT(L), m, lower=False)


def objective(F, QL, means):
    """``sum_k ||QL^-1 (m_k - F m_{k+1})||^2`` (``pof/utils.py:97-101`` with the call order of ``filter.py:43-45``)."""
    r = solve_triangular(np.broadcast_to(QL, (means.shape[0] - 1,) + QL.shape), means[:-1] - mv(F, means[1:]), lower=True)
    return float(np.sum(r * r))


def linear_noiseless_filtering(x0, F, QL, om):
    """Parallel filter (``filter.py:18-47``): returns ``(MVNSqrt, nll, obj, ssq)``."""
    H, c, cholR = om
    N, d = c.shape
    D = F.shape[-1]
    ms = np.zeros((N, D))
    Ls = np.zeros((N, D, D))
    ms[0] = x0.mean
    Ls[0] = x0.chol
    elems = filter_elements(F, QL, H, c, cholR, ms, Ls)
    _, means, chols, _, _ = associative_scan(sqrt_filtering_operator, elems)
    means = np.concatenate([x0.mean[None], means])
    chols = np.concatenate([x0.chol[None], chols])
    om_, oL = get_obs(F, QL, H, c, cholR, means[:-1], chols[:-1])
    ress = whiten(om_, oL)
    ssq = float(np.sum(ress * ress) / N / d)
    nll = float(-np.sum(mvn_loglikelihood(om_, oL)))
    obj = objective(F, QL, means)
    return MVNSqrt(means, chols), nll, obj, ssq


# --------------------------------------------------------------------------
# smoother (pof/parallel_filtsmooth/smoother.py)
# --------------------------------------------------------------------------
def smoother_elements(F, QL, ms, Ls):
    """``get_elements`` + ``_sqrt_associative_params`` (``smoother.py:25-50``)."""
    nx = F.shape[-1]
    n = ms.shape[0] - 1
    Fb = np.broadcast_to(F, (n, nx, nx))
    QLb = np.broadcast_to(QL, (n, nx, nx))
    m, L = ms[:-1], Ls[:-1]
    Phi = np.concatenate(
        [
            np.concatenate([Fb @ L, QLb], axis=-1),
            np.concatenate([L, np.zeros((n, nx, nx))], axis=-1),
        ],
        axis=-2,
    )
    T = tria(Phi)
    Phi11 = T[:, :nx, :nx]
    Phi21 = T[:, nx:, :nx]
    Dm = T[:, nx:, nx:]
    # jlinalg.solve(Phi11^T, Phi21^T)^T (LU in the reference; Phi11 is triangular)
    E = mT(np.linalg.solve(mT(Phi11), mT(Phi21)))
    g = m - mv(E, mv(Fb, m))
    g = np.concatenate([g, ms[-1:]])
    E = np.concatenate([E, np.zeros((1, nx, nx))])
    Dm = np.concatenate([Dm, Ls[-1:]])
    return g, E, Dm


def sqrt_smoothing_operator(e1, e2):
    """(``smoother.py:53-63``); under ``reverse=True`` ``e1`` is the later aggregate."""
    g1, E1, D1 = e1
    g2, E2, D2 = e2
    return mv(E2, g1) + g2, E2 @ E1, tria(np.concatenate([E2 @ D1, D2], axis=-1))


def smoothing(F, QL, filt):
    """Parallel RTS smoother (``smoother.py:8-22``): returns ``(MVNSqrt, obj)``."""
    elems = smoother_elements(F, QL, filt.mean, filt.chol)
    means, _, chols = associative_scan(sqrt_smoothing_operator, elems, reverse=True)
    return MVNSqrt(means, chols), objective(F, QL, means)


def linear_filtsmooth(x0, F, QL, om):
    """(``pof/parallel_filtsmooth/__init__.py:5-10``): ``(MVNSqrt, nll, obj, ssq)``."""
    out, nll, _, ssq = linear_noiseless_filtering(x0, F, QL, om)
    out, obj = smoothing(F, QL, out)
    return out, nll, obj, ssq
