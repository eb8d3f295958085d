"""The reference's own host-side tests, run against this package's API with the
same calls and assertions: ``tests/test_transtions.py`` (IWP, projection
matrices, discretisation) and ``tests/test_ivp.py`` (the IVP objects; tornadox
is absent here, so its ``InitialValueProblem`` fields are checked instead).
The discretisation is also checked against the IWP closed forms, which the
reference test only constructs.  No GPU: the transitions are host functions of
the C ABI."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("dim", [1, 3, 5])
@pytest.mark.parametrize("order", [1, 3, 5])
def test_iwp(dim, order):
    from pof.transitions import IWP

    iwp = IWP(num_derivatives=order, wiener_process_dimension=dim)
    assert iwp.num_derivatives == order and iwp.wiener_process_dimension == dim


@pytest.mark.parametrize("dim", [1, 3, 5])
@pytest.mark.parametrize("order", [1, 3, 5])
def test_projection_matrix(dim, order):
    from pof.transitions import IWP, projection_matrix

    iwp = IWP(num_derivatives=order, wiener_process_dimension=dim)
    D = dim * (order + 1)
    E0 = projection_matrix(iwp, 0)
    E1 = projection_matrix(iwp, 1)
    assert E0.shape == (dim, D)
    assert E1.shape == (dim, D)
    x = np.random.default_rng(dim * 10 + order).uniform(size=D)
    assert all(E0 @ x == x[0 :: order + 1])
    assert all(E1 @ x == x[1 :: order + 1])


def test_discretize():
    from pof.transitions import IWP, TransitionModel, get_transition_model

    d, q = 2, 3
    iwp = IWP(num_derivatives=q, wiener_process_dimension=d)
    D = d * (q + 1)
    dt = 0.1
    tm = get_transition_model(iwp, dt)
    assert isinstance(tm, TransitionModel)
    assert tm.F.shape == (D, D) and tm.QL.shape == (D, D)
    # integrated Wiener process of order q, per ODE dimension (state [y, y', ..., y^(q)]):
    # F_ij = dt^(j-i) / (j-i)!, Q_ij = dt^(2q+1-i-j) / ((2q+1-i-j) (q-i)! (q-j)!)
    F1 = np.zeros((q + 1, q + 1))
    Q1 = np.zeros((q + 1, q + 1))
    for i in range(q + 1):
        for j in range(q + 1):
            if j >= i:
                F1[i, j] = dt ** (j - i) / math.factorial(j - i)
            p = 2 * q + 1 - i - j
            Q1[i, j] = dt**p / (p * math.factorial(q - i) * math.factorial(q - j))
    np.testing.assert_allclose(tm.F, np.kron(np.eye(d), F1), rtol=1e-14, atol=0)
    np.testing.assert_allclose(tm.QL @ tm.QL.T, np.kron(np.eye(d), Q1), rtol=1e-10, atol=1e-14 * Q1.max())


@pytest.mark.parametrize("name", ["logistic", "lotkavolterra"])
def test_ivp(name):
    import pof.ivp

    ivp = getattr(pof.ivp, name)()
    # tornadox.ivp.InitialValueProblem's fields: f, t0, tmax, y0 (and f(t, y) evaluates the field)
    assert ivp.t0 < ivp.tmax
    y0 = np.asarray(ivp.y0, dtype=np.float64)
    fy = np.asarray(ivp.f(None, y0))
    assert fy.shape == y0.shape and np.all(np.isfinite(fy))
