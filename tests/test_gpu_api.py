"""GPU tests of the sub-APIs around the solve: the IEKS iterator and the
linearisations.  Compared with the single solve (same kernels) and the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,N,q", [("logistic", 40, 3), ("fitzhughnagumo", 100, 3), ("vanderpol", 64, 2)])
def test_ieks_iterator_matches_solve(gpu, name, N, q):
    import torch

    import pof.ivp
    import pof.iterators
    import pof.solver

    kw = {"stiffness_constant": 1.0} if name == "vanderpol" else {}
    ivp = getattr(pof.ivp, name)(**kw)
    ts = np.linspace(ivp.t0, ivp.tmax, N)
    it, setup = pof.iterators.ieks_iterator(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init="constant")
    iterates = list(it)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init="constant", return_full_states=True)
    assert len(iterates) == info["iterations"]
    states, nll, obj, ssq = iterates[-1]
    assert torch.equal(states.mean, info["states"].mean)
    assert nll == info["nll"] and obj == info["obj"] and ssq == info["sigma_squared"]
    # one calibration per step here, two in solve
    L = states.chol.cpu().numpy()
    L2 = info["states"].chol.cpu().numpy()
    np.testing.assert_allclose(L * np.sqrt(ssq), L2, rtol=1e-12, atol=1e-300)
