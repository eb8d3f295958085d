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


@pytest.mark.parametrize("name,N,q,init", [("logistic", 40, 3, "constant"), ("fitzhughnagumo", 100, 3, "constant"),
                                           ("lorenz63", 65, 4, "prior")])
def test_ieks_iterator_iterates_match_oracle(gpu, name, N, q, init):
    """Every iterate the generator yields against the oracle's iterate of the same index
    (oracle.solver.solve with return_history: ieks_step of pof/step.py:33-45 restated):
    preconditioned means to 1e-9 relative, L L^T / ssq to 1e-6, nll and obj as in the solve
    parity tests, ssq to 1e-8 (stable for d = 1, FHN and Lorenz-63), and the same number
    of iterates as the oracle's iterations."""
    import torch

    import pof.ivp
    import pof.iterators
    from oracle import problems as op
    from oracle import solver as osolver

    kw = {"tmax": 0.5} if name == "lorenz63" else {}
    ivp = getattr(pof.ivp, name)(**kw)
    oivp = getattr(op, name)(**kw)
    ts = np.linspace(ivp.t0, ivp.tmax, N)
    it, _ = pof.iterators.ieks_iterator(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init=init)
    iterates = list(it)
    torch.cuda.synchronize()
    _, rinfo = osolver.solve(oivp, ts, q, init=init, maxiters=10_000, return_history=True)
    hist = rinfo["history"]
    print(f"case {name} N={N} q={q}: {len(iterates)} iterates, oracle {len(hist)}", flush=True)
    assert len(iterates) == len(hist)
    for k, ((st, nll, obj, ssq), (rs, rnll, robj, rssq)) in enumerate(zip(iterates, hist)):
        m = st.mean.cpu().numpy()
        assert np.abs(m - rs.mean).max() <= 1e-9 * np.abs(rs.mean).max(), k
        L = st.chol.cpu().numpy() / np.sqrt(ssq)
        R = rs.chol / np.sqrt(rssq)
        cc, rc = np.einsum("nij,nkj->nik", L, L), np.einsum("nij,nkj->nik", R, R)
        assert np.abs(cc - rc).max() <= 1e-6 * np.abs(rc).max(), k
        assert abs(ssq - rssq) <= 1e-8 * abs(rssq), (k, ssq, rssq)
        assert abs(obj - robj) <= 1e-6 * abs(robj) + 1e-12, k
        assert abs(nll - rnll) <= 1e-8 * abs(rnll) + 1e-8, k
