"""Pins of the oracle's alternative linearisations (``oracle/linearization.py``), CPU only.

Each check follows from the mathematics, not from the oracle's own formula:
the EK0 ``H`` is the un-preconditioned derivative projection; every affine
model reproduces the residual at the linearisation point where the
reference's ``b`` is ``r(m) - H m``; ``cholR`` factors the covariance it is
defined by (``H L L^T H^T`` for the uncertain EK1; the Schur complement
``Phi - Psi^T P^-1 Psi`` of the sigma-point joint covariance for the
unscented one); the rank-one Cholesky update against the dense update.
"""
import numpy as np
import pytest

from oracle import linearization as olin
from oracle import model as om
from oracle import problems as op


def _setup(name, q, **kw):
    p = getattr(op, name)(**kw)
    return om.set_up_solver(p, np.linspace(p.t0, p.tmax, 11), q)


def _state(D, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal(D)
    L = np.tril(rng.standard_normal((D, D))) * scale
    L[np.diag_indices(D)] = np.abs(L[np.diag_indices(D)]) + scale
    return m, L


@pytest.mark.parametrize("mult", [1.0, -1.0])
def test_cholesky_update_many_dense(mult):
    rng = np.random.default_rng(1)
    n = 5
    L = np.tril(rng.standard_normal((n, n)))
    L[np.diag_indices(n)] = np.abs(L[np.diag_indices(n)]) + 2.0
    V = rng.standard_normal((3, n)) * (0.3 if mult < 0 else 1.0)
    out = olin.cholesky_update_many(L, V, mult)
    want = L @ L.T + mult * V.T @ V
    np.testing.assert_allclose(out @ out.T, want, rtol=1e-12, atol=1e-12)
    assert np.all(np.triu(out, 1) == 0) and np.all(np.diag(out) > 0)


@pytest.mark.parametrize("name,q", [("logistic", 3), ("fitzhughnagumo", 2), ("lorenz63", 2)])
def test_ek0_and_uncertain(name, q):
    s = _setup(name, q)
    D = s.E0.shape[1]
    m, L = _state(D, 2)
    ek0 = olin.linearize_ek0(s, m)
    # the un-preconditioned projection: H x = x[1::q+1]
    np.testing.assert_array_equal(ek0.H @ np.arange(D, dtype=float), np.arange(D, dtype=float)[1::q + 1])
    np.testing.assert_allclose(ek0.H @ m + ek0.b, om.residual(s, m), rtol=1e-13, atol=1e-13)
    assert np.all(ek0.cholR == 0)
    unc = olin.uncertain_linearize(s, m, L)
    ek1 = om.linearize(s, m)
    np.testing.assert_array_equal(unc.H, ek1.H)
    np.testing.assert_array_equal(unc.b, ek1.b)
    HL = ek1.H @ L
    np.testing.assert_allclose(unc.cholR @ unc.cholR.T, HL @ HL.T, rtol=1e-12, atol=1e-13)


def test_regularized_blocks():
    s = _setup("fitzhughnagumo", 2)
    D = s.E0.shape[1]
    d = 2
    m, _ = _state(D, 3)
    lam = 0.25
    reg = olin.linearize_regularized(s, m, lam)
    ek1 = om.linearize(s, m)
    assert reg.H.shape == (d + D, D) and reg.b.shape == (d + D,) and reg.cholR.shape == (d + D, d + D)
    np.testing.assert_array_equal(reg.H[:d], ek1.H)
    np.testing.assert_array_equal(reg.H[d:], np.eye(D))
    np.testing.assert_array_equal(reg.b[:d], om.residual(s, m))
    np.testing.assert_array_equal(reg.b[d:], -m)
    np.testing.assert_allclose(reg.cholR[d:, d:] @ reg.cholR[d:, d:].T, np.eye(D) / lam, rtol=1e-15)
    assert np.all(reg.cholR[:d] == 0) and np.all(reg.cholR[:, :d] == 0)


@pytest.mark.parametrize("name,q,scale", [("logistic", 2, 0.05), ("lotkavolterra", 2, 0.1), ("rigid_body", 1, 0.1)])
def test_unscented_regression(name, q, scale):
    """Fields whose every component is nonlinear, so the downdated covariance is positive definite."""
    s = _setup(name, q)
    D = s.E0.shape[1]
    m, L = _state(D, 4, scale)
    ut = olin.linearize_unscented(s, m, L)
    # independent statement of the sigma-point moments
    n = D
    lam = 1.0 * (n + 3.0 + n) - n
    c = np.sqrt(n + lam) * L
    pts = np.vstack([m, m + c.T, m - c.T])
    w = np.full(2 * n + 1, 0.5 / (n + lam))
    w[0] = lam / (n + lam)
    fx = om.residual(s, pts)
    mf = w @ fx
    dx, df = pts - m, fx - mf
    Pxx = (dx.T * w) @ dx
    Pxf = (dx.T * w) @ df
    Pff = (df.T * w) @ df
    np.testing.assert_allclose(Pxx, L @ L.T, rtol=1e-12, atol=1e-14)
    F = np.linalg.solve(Pxx, Pxf).T
    np.testing.assert_allclose(ut.H, F, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ut.b, mf - F @ m, rtol=1e-9, atol=1e-12)
    schur = Pff - Pxf.T @ np.linalg.solve(Pxx, Pxf)
    np.testing.assert_allclose(ut.cholR @ ut.cholR.T, schur, rtol=1e-7, atol=1e-10 * np.abs(Pff).max())
    # small spread: the regression slope approaches the EK1 Jacobian
    assert np.abs(ut.H - om.linearize(s, m).H).max() < 10 * scale * np.abs(ut.H).max()


def test_unscented_singular_downdate_is_finite():
    """A linear component of the field (FitzHugh-Nagumo's w') makes the downdated covariance singular.
    The downdate then divides by rounding-level diagonals (or takes the square root of a rounding-level
    negative, NaN, which the reference zeroes, utils.py:174): the result is finite but rounding-dominated,
    so the GPU parity of the unscented linearisation uses fields whose every component is nonlinear."""
    s = _setup("fitzhughnagumo", 2)
    m, L = _state(s.E0.shape[1], 4, 0.1)
    ut = olin.linearize_unscented(s, m, L)
    assert np.all(np.isfinite(ut.cholR))
