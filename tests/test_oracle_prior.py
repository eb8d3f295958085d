"""Closed-form pin of the "prior" initial trajectory (``solve``'s default init).

The reference (``pof/initialization.py:75-89`` with ``pof/convenience.py:84-90``)
predicts the Taylor-mode state x0 (derivatives y^(j)(t0), zero covariance) in
ONE step to each grid time with the non-preconditioned IWP transitions, using
``steps = ts[1:]`` -- the absolute times, not ``ts[k] - ts[0]`` -- and returns
the result without multiplying by ``PI``.  So, per ODE dimension, with
h = ts[k] for k >= 1:

* mean:  m_k[i] = sum_{j >= i} x0[j] |h|^(j-i) / (j-i)!   (the Taylor
  polynomial of the solution's derivatives, extrapolated over h; the
  preconditioner's |dt| makes a negative time extrapolate over |h|);
* factor: L_k L_k^T = Q(h), Q_ij = |h|^(2q+1-i-j) / ((2q+1-i-j) (q-i)! (q-j)!)
  (the IWP process noise; the prediction of a zero covariance);
* k = 0: x0 itself with a zero factor.

These are written here from the mathematics, independently of the oracle's
transition code, and checked against the oracle (``oracle.model.prior_init``)
and the library's host entry point (``pof_initial_means``, no GPU).  Grids with
ts[0] != 0 pin the absolute steps and the missing PI: with ``ts[k] - ts[0]``,
or with PI applied, the values differ by orders of magnitude.  (A grid with an
interior time exactly 0 gives a zero step, hence NaN through P^-1, in the
reference as here; it is not a case the solver can take.)
"""
import ctypes
from math import factorial

import numpy as np
import pytest

from oracle import model as om
from oracle import problems as op
from pof import _lib

CASES = [
    ("logistic", {}, np.linspace(0.0, 10.0, 9), 3),
    ("logistic", {}, np.linspace(2.0, 3.0, 9), 2),
    ("fitzhughnagumo", {}, np.linspace(0.0, 100.0, 7), 3),
    ("fitzhughnagumo", {}, np.linspace(5.0, 9.0, 5), 1),
    ("lorenz63", {}, np.linspace(0.5, 0.9, 6), 4),
    ("rigid_body", {}, np.linspace(-1.0, -0.2, 5), 2),
]


def phi(h, q):
    # the Nordsieck scaling uses |dt| (transitions.py:53-68), so a negative time
    # predicts over |h|: P A P^-1 = |h|^(j-i) / (j-i)!
    A = np.zeros((q + 1, q + 1))
    for i in range(q + 1):
        for j in range(i, q + 1):
            A[i, j] = abs(h) ** (j - i) / factorial(j - i)
    return A


def qmat(h, q):
    Q = np.zeros((q + 1, q + 1))
    for i in range(q + 1):
        for j in range(q + 1):
            p = 2 * q + 1 - i - j
            Q[i, j] = abs(h) ** p / (p * factorial(q - i) * factorial(q - j))
    return Q


def expected(problem, q, ts):
    x0 = om.taylor_mode_init(problem, q).mean  # derivatives y^(j)(t0), interleaved per dimension
    d = problem.d
    N, D = len(ts), d * (q + 1)
    m = np.empty((N, D))
    C = np.zeros((N, D, D))
    m[0] = x0
    for k in range(1, N):
        m[k] = np.kron(np.eye(d), phi(ts[k], q)) @ x0
        C[k] = np.kron(np.eye(d), qmat(ts[k], q))
    return m, C


@pytest.mark.parametrize("name,kw,ts,q", CASES)
def test_oracle_prior_init_closed_form(name, kw, ts, q):
    p = getattr(op, name)(**kw)
    m, C = expected(p, q, ts)
    got = om.prior_init(p, q, ts)
    scale = np.abs(m).max(axis=1, keepdims=True)
    assert np.all(np.abs(got.mean - m) <= 1e-13 * scale)
    cc = np.einsum("nij,nkj->nik", got.chol, got.chol)
    cs = np.abs(C).max(axis=(1, 2), keepdims=True) + 1e-300
    assert np.all(np.abs(cc - C) <= 1e-12 * cs)
    # the trajectory the solver starts from: the same values, not multiplied by PI
    setup = om.set_up_solver(p, ts, q)
    traj = om.get_initial_trajectory(setup, "prior")
    assert np.array_equal(traj.mean, got.mean)


@pytest.mark.parametrize("name,kw,ts,q", CASES)
def test_library_prior_means_closed_form(name, kw, ts, q):
    """The library's host-side initial trajectory (pof_initial_means, used by the GPU solve)."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("libpof_hip.so is not built")
    import pof.ivp
    from pof.convenience import ProblemSpec

    p = getattr(op, name)(**kw)
    m, _ = expected(p, q, ts)
    ivp = getattr(pof.ivp, name)(**kw)
    spec = ProblemSpec(ivp.f, ivp.y0, ts, q, "prior")
    prob = spec.cstruct()
    out = np.empty((len(ts), spec.D))
    assert _lib.load().pof_initial_means(ctypes.byref(prob), _lib.hptr(out)) == _lib.POF_OK
    scale = np.abs(m).max(axis=1, keepdims=True)
    assert np.all(np.abs(out - m) <= 1e-13 * scale)


def test_relative_steps_would_fail_the_pin():
    """The check has teeth: ts[k] - ts[0] as the step, or PI applied, moves the values far off."""
    p = op.logistic()
    ts = np.linspace(2.0, 3.0, 9)
    q = 2
    m, _ = expected(p, q, ts)
    x0 = om.taylor_mode_init(p, q).mean
    wrong = np.array([phi(t - ts[0], q) @ x0 for t in ts[1:]])
    assert np.abs(wrong - m[1:]).max() > 1e-3 * np.abs(m).max()
    setup = om.set_up_solver(p, ts, q)
    assert np.abs(setup.PI @ m[1].T - m[1]).max() > 1e-3 * np.abs(m[1]).max()
