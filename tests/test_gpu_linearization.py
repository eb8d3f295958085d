"""GPU parity of the linearisations beyond EK1 and of the filters with observation noise.

* ``linearize_ek0``, ``uncertain_linearize``, ``linearize_regularized``
  (reference ``pof/observations.py:43-89``) and ``linearize_unscented``
  (``pof/linearization/unscented.py:19-70``) on random states, against the
  oracle (``oracle/linearization.py``).
* ``linear_filtsmooth`` (parallel and sequential) with a nonzero ``cholR``.
* IEKS iterations that linearise with EK0 / the uncertain EK1 and run the
  device filter, against the same loop on the oracle.

Tolerances (float64): ``H`` and ``b`` 1e-12 relative to their largest entry
(1e-10 for the unscented regression, which solves with ``L L^T``);
``cholR cholR^T`` 1e-10 (1e-8 unscented: the downdate subtracts nearly
equal covariances); filter means 1e-10, ``L L^T`` 1e-9.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def gram(c):
    return np.einsum("...ij,...kj->...ik", c, c)


def _setups(name, q, **kw):
    import pof.convenience as pc
    import pof.ivp
    from oracle import model as om
    from oracle import problems as op

    ivp = getattr(pof.ivp, name)(**kw)
    oivp = getattr(op, name)(**kw)
    ts = np.linspace(oivp.t0, oivp.tmax, 11)
    return pc.set_up_solver(f=ivp.f, y0=ivp.y0, ts=ts, order=q), om.set_up_solver(oivp, ts, q)


def _states(n, D, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, D))
    L = np.tril(rng.standard_normal((n, D, D))) * scale
    idx = np.arange(D)
    L[:, idx, idx] = np.abs(L[:, idx, idx]) + scale
    return m, L


@pytest.mark.parametrize("name,q", [("logistic", 3), ("fitzhughnagumo", 2), ("lorenz63", 2), ("rigid_body", 1)])
def test_ek0_uncertain_regularized(gpu, name, q):
    import torch

    import pof.observations as po
    from oracle import linearization as olin
    from pof.utils import MVNSqrt

    setup, os_ = _setups(name, q)
    D = os_.E0.shape[1]
    m, L = _states(7, D, 11)
    x = MVNSqrt(torch.as_tensor(m, device="cuda"), torch.as_tensor(L, device="cuda"))
    print(f"linearisations {name} q={q}", flush=True)
    ek0 = po.linearize_ek0(setup["om"], x)
    unc = po.uncertain_linearize(setup["om"], x)
    reg = po.linearize_regularized(setup["om"], x, 0.3)
    torch.cuda.synchronize()
    for k in range(m.shape[0]):
        r0 = olin.linearize_ek0(os_, m[k])
        assert np.array_equal(ek0.H[k].cpu().numpy(), r0.H)
        assert rel(ek0.b[k].cpu().numpy(), r0.b) < 1e-12
        assert np.all(ek0.cholR[k].cpu().numpy() == 0)
        r1 = olin.uncertain_linearize(os_, m[k], L[k])
        assert rel(unc.H[k].cpu().numpy(), r1.H) < 1e-12
        assert rel(unc.b[k].cpu().numpy(), r1.b) < 1e-12
        assert rel(gram(unc.cholR[k].cpu().numpy()), gram(r1.cholR)) < 1e-10
        assert np.all(np.triu(unc.cholR[k].cpu().numpy(), 1) == 0)
        r2 = olin.linearize_regularized(os_, m[k], 0.3)
        assert rel(reg.H[k].cpu().numpy(), r2.H) < 1e-12
        assert rel(reg.b[k].cpu().numpy(), r2.b) < 1e-12
        assert rel(reg.cholR[k].cpu().numpy(), r2.cholR) < 1e-15


@pytest.mark.parametrize("name,q,kappa", [("logistic", 2, None), ("lotkavolterra", 2, None), ("rigid_body", 1, None),
                                          ("lotkavolterra", 1, 0.5)])
def test_unscented(gpu, name, q, kappa):
    import torch

    from oracle import linearization as olin
    from pof.linearization.unscented import linearize_unscented
    from pof.utils import MVNSqrt

    setup, os_ = _setups(name, q)
    D = os_.E0.shape[1]
    m, L = _states(5, D, 12)
    print(f"unscented {name} q={q} kappa={kappa}", flush=True)
    out = linearize_unscented(setup["om"], MVNSqrt(torch.as_tensor(m, device="cuda"), torch.as_tensor(L, device="cuda")),
                              kappa=kappa)
    torch.cuda.synchronize()
    for k in range(m.shape[0]):
        r = olin.linearize_unscented(os_, m[k], L[k], kappa=kappa)
        assert rel(out.H[k].cpu().numpy(), r.H) < 1e-10
        assert rel(out.b[k].cpu().numpy(), r.b) < 1e-10
        assert rel(gram(out.cholR[k].cpu().numpy()), gram(r.cholR)) < 1e-8


@pytest.mark.parametrize("sequential,rotate_ql", [(False, False), (True, False), (False, True)])
def test_linear_filtsmooth_with_noise(gpu, sequential, rotate_ql):
    """rotate_ql: QL times an orthogonal matrix (same Q, no longer lower triangular), which the
    smoother-element factorisation must take without its banded shortcut."""
    import torch

    import pof.parallel_filtsmooth as pfs
    import pof.sequential_filtsmooth as sfs
    from oracle import model as om
    from oracle import parallel as opar
    from oracle import sequential as oseq
    from pof.observations import AffineModel
    from pof.transitions import TransitionModel
    from pof.utils import MVNSqrt

    rng = np.random.default_rng(21)
    d, q, Ns = 2, 2, 41
    D = d * (q + 1)
    F, QL = om.preconditioned_discretize(om.IWP(d, q))
    if rotate_ql:
        rot, _ = np.linalg.qr(rng.standard_normal((D, D)))
        QL = QL @ rot
        assert np.abs(np.triu(QL, 1)).max() > 1e-3
    H = rng.standard_normal((Ns, d, D))
    c = rng.standard_normal((Ns, d))
    R = np.tril(rng.standard_normal((Ns, d, d))) * 0.5
    m0 = rng.standard_normal(D)
    L0 = np.tril(rng.standard_normal((D, D)))
    dom = om.AffineModel(H, c, R)
    ofs = oseq.linear_filtsmooth if sequential else opar.linear_filtsmooth
    ref, rnll, robj, rssq = ofs(om.MVNSqrt(m0, L0), F, QL, dom)
    gdom = AffineModel(torch.as_tensor(H, device="cuda"), torch.as_tensor(c, device="cuda"),
                       torch.as_tensor(R, device="cuda"))
    gfs = sfs.linear_filtsmooth if sequential else pfs.linear_filtsmooth
    out, nll, obj, ssq = gfs(MVNSqrt(m0, L0), TransitionModel(F, QL), gdom)
    torch.cuda.synchronize()
    assert rel(out.mean.cpu().numpy(), ref.mean) < 1e-10
    assert rel(gram(out.chol.cpu().numpy()), gram(ref.chol)) < 1e-9
    assert abs(nll - rnll) <= 1e-9 * abs(rnll)
    assert abs(obj - robj) <= 1e-9 * abs(robj)
    # d = 2: ssq whitens with the sign-dependent factor; the noise makes the pivots well-defined here
    assert abs(ssq - rssq) <= 1e-8 * abs(rssq)


@pytest.mark.parametrize("method", ["ek0", "uncertain"])
def test_ieks_loop_with_other_linearisation(gpu, method):
    """Three iterations of linearise-at-the-previous-states + parallel filter/smoother (step.py:33-45
    with the linearisation swapped, as the reference's step.py comments show), against the oracle."""
    import torch

    import pof.convenience as pc
    import pof.ivp
    import pof.observations as po
    import pof.parallel_filtsmooth as pfs
    from oracle import linearization as olin
    from oracle import model as om
    from oracle import parallel as opar
    from oracle import problems as op
    from pof.observations import AffineModel
    from pof.utils import MVNSqrt

    ts = np.linspace(0.0, 10.0, 33)
    ivp = pof.ivp.logistic()
    setup = pc.set_up_solver(f=ivp.f, y0=ivp.y0, ts=ts, order=2)
    osetup = om.set_up_solver(op.logistic(), ts, 2)
    # start from the states of one EK1 step (a nonzero factor for the uncertain linearisation)
    import pof.step
    from oracle import solver as osolver

    st = pc.get_initial_trajectory(setup, method="constant")
    st, *_ = pof.step.ieks_step(states=st, setup=setup)
    ost = om.get_initial_trajectory(osetup, "constant")
    ost, *_ = osolver.ieks_step(osetup, ost)
    for it in range(3):
        lin = MVNSqrt(st.mean[1:].contiguous(), st.chol[1:].contiguous())
        if method == "ek0":
            dom = po.linearize_ek0(setup["om"], lin)
            odom = [olin.linearize_ek0(osetup, ost.mean[k + 1]) for k in range(len(ts) - 1)]
        else:
            dom = po.uncertain_linearize(setup["om"], lin)
            odom = [olin.uncertain_linearize(osetup, ost.mean[k + 1], ost.chol[k + 1]) for k in range(len(ts) - 1)]
        oH = np.stack([o.H for o in odom])
        ob = np.stack([o.b for o in odom])
        oR = np.stack([o.cholR for o in odom])
        st, nll, obj, ssq = pfs.linear_filtsmooth(setup["x0"], setup["dtm"], AffineModel(dom.H, dom.b, dom.cholR))
        ost, rnll, robj, rssq = opar.linear_filtsmooth(osetup.x0, osetup.F, osetup.QL, om.AffineModel(oH, ob, oR))
        torch.cuda.synchronize()
        print(f"{method} iteration {it}: nll {nll} vs {rnll}", flush=True)
        assert rel(st.mean.cpu().numpy(), ost.mean) < 1e-9
        assert rel(gram(st.chol.cpu().numpy()), gram(ost.chol)) < 1e-8
        assert abs(obj - robj) <= 1e-8 * abs(robj) + 1e-12
