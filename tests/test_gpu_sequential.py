"""GPU parity of the sequential paths against the FP64 CPU oracle.

* ``solve(sequential=True)``: each IEKS iteration runs the sequential
  square-root filter and RTS smoother (reference ``pof/step.py:39-40``,
  ``pof/sequential_filtsmooth``); one GPU worker steps through time per IVP.
* ``sequential_eks_solve``: single-pass EKS, EKF linearised at the predicted
  means (``pof/solver.py:76-96``, ``sequential_filtsmooth/filter.py:9-30``).
* ``init="coarse"``: EKS on 100 points looked up on the fine grid
  (``pof/initialization.py:112-129``, ``convenience.py:80-83``).
* ``ieks_step(sequential=True)`` and ``sequential_filtsmooth.linear_filtsmooth``.

Tolerances (float64): means 1e-9 relative to the largest entry (1e-6 for the
chaotic Lorenz-63 EKS, where rounding grows along the trajectory); ``L L^T``
1e-8 after dividing out the calibration; iteration counts equal; ``nll``
(+summed log-likelihood on this branch, as in the reference) and ``obj`` 1e-8.
``sigma_squared`` whitens with ``I^T`` (pof/utils.py:110-112), so for d >= 2
it depends on the column signs of the innovation factor ``I``, whose
Householder pivots can be rounding noise: the oracle's own value moves by
0.7% (rigid body) and 8% (Lorenz-63 EKS) under 1e-15 perturbations of y0.
It is compared to 1e-8 where the oracle is stable under those perturbations
and otherwise must lie within 1.5x the oracle's measured spread.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def cov(c):
    return np.einsum("nij,nkj->nik", c, c)


def close(a, b, rtol=1e-8):
    return abs(a - b) <= rtol * abs(b) + 1e-12


def eks_ssq_tolerance(oivp, ts, q, s_ref):
    from oracle import problems as op
    from oracle import solver as osolver

    spread = 0.0
    for eps in (1e-15, -1e-15, 3e-15, -3e-15, 1e-14, -1e-14):
        pp = op.Problem(oivp.name, oivp.params, oivp.y0 * (1 + eps), oivp.t0, oivp.tmax)
        _, inf = osolver.sequential_eks_solve(pp, ts, q, calibrate=False)
        spread = max(spread, abs(inf["sigma_squared"] - s_ref))
    return max(1e-8 * abs(s_ref), 1.5 * spread)


def _pair(name, **kw):
    import pof.ivp
    from oracle import problems as op

    return getattr(pof.ivp, name)(**kw), getattr(op, name)(**kw)


SEQ_CASES = [
    ("logistic", {}, 40, 3, "constant"),
    ("fitzhughnagumo", {}, 100, 3, "constant"),
    ("rigid_body", {}, 64, 2, "constant"),
    ("lotkavolterra", {}, 50, 2, "prior"),
    ("vanderpol", {"stiffness_constant": 1.0}, 64, 2, "constant"),
]


@pytest.mark.parametrize("name,kw,N,q,init", SEQ_CASES)
def test_sequential_solve_matches_oracle(gpu, name, kw, N, q, init):
    import torch

    import pof.solver
    from oracle import solver as osolver

    ivp, oivp = _pair(name, **kw)
    ts = np.linspace(oivp.t0, oivp.tmax, N)
    print(f"sequential solve {name} N={N} q={q} init={init}", flush=True)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init=init, maxiters=200, sequential=True)
    torch.cuda.synchronize()
    ref, rinfo = osolver.solve(oivp, ts, q, init=init, maxiters=200, is_sequential=True)
    assert info["iterations"] == rinfo["iterations"]
    assert rel(ys.mean.cpu().numpy(), ref.mean) < 1e-9
    s_gpu, s_ref = info["sigma_squared"], rinfo["sigma_squared"]
    assert close(s_gpu, s_ref)
    assert rel(cov(ys.chol.cpu().numpy()) / s_gpu**2, cov(ref.chol) / s_ref**2) < 1e-8
    assert close(info["obj"], rinfo["obj"], 1e-6)
    assert close(info["nll"], rinfo["nll"])


def test_sequential_equals_parallel_solution(gpu):
    """Parallel and sequential IEKS reach the same fixed point (same model, different association order)."""
    import torch

    import pof.ivp
    import pof.solver

    ivp = pof.ivp.fitzhughnagumo()
    ts = np.linspace(0.0, 20.0, 513)
    par, _ = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=3, init="constant")
    seq, _ = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=3, init="constant", sequential=True)
    torch.cuda.synchronize()
    assert rel(seq.mean.cpu().numpy(), par.mean.cpu().numpy()) < 1e-8


EKS_CASES = [
    ("fitzhughnagumo", {}, 300, 3, 1e-9),   # the README figure's sequential EKS
    ("logistic", {}, 50, 2, 1e-9),
    ("rigid_body", {}, 128, 3, 1e-9),
    ("lorenz63", {}, 200, 3, 1e-6),
]


@pytest.mark.parametrize("name,kw,N,q,tol", EKS_CASES)
@pytest.mark.parametrize("full,calibrate", [(False, True), (True, True), (True, False)])
def test_eks_matches_oracle(gpu, name, kw, N, q, tol, full, calibrate):
    import torch

    import pof.solver
    from oracle import solver as osolver

    ivp, oivp = _pair(name, **kw)
    ts = np.linspace(oivp.t0, oivp.tmax, N)
    print(f"eks {name} N={N} q={q} full={full} calibrate={calibrate}", flush=True)
    ys, info = pof.solver.sequential_eks_solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, return_full_states=full,
                                               calibrate=calibrate)
    torch.cuda.synchronize()
    ref, rinfo = osolver.sequential_eks_solve(oivp, ts, q, return_full_states=full, calibrate=calibrate)
    D = oivp.d * (q + 1)
    assert ys.mean.shape == ((N, D) if full else (N, oivp.d))
    assert rel(ys.mean.cpu().numpy(), ref.mean) < tol
    s_gpu, s_ref = info["sigma_squared"], rinfo["sigma_squared"]
    # calibrated factors carry sqrt(ssq) once: divide it out of L L^T
    cg, cr = (s_gpu, s_ref) if calibrate else (1.0, 1.0)
    assert rel(cov(ys.chol.cpu().numpy()) / cg, cov(ref.chol) / cr) < max(tol, 1e-8)
    assert abs(s_gpu - s_ref) <= eks_ssq_tolerance(oivp, ts, q, s_ref)
    assert close(info["nll"], rinfo["nll"], max(tol, 1e-8))
    assert close(info["obj"], rinfo["obj"], max(tol, 1e-6))
    assert info["calibrated"] == calibrate


def test_eks_readme_figure(gpu):
    """readmefig.svg of the reference: 300 means and +-2 sd bands of
    sequential_eks_solve(FHN, linspace(0, 100, 300), 3), decoded to ~5e-8."""
    from pathlib import Path

    import torch

    import pof.ivp
    import pof.solver

    z = np.load(Path(__file__).parent / "golden" / "readme_fhn.npz")
    ivp = pof.ivp.fitzhughnagumo()
    ts = np.linspace(0.0, 100.0, 300)
    ys, _ = pof.solver.sequential_eks_solve(f=ivp.f, y0=ivp.y0, ts=ts, order=3)
    torch.cuda.synchronize()
    m = ys.mean.cpu().numpy()
    c = ys.chol.cpu().numpy()
    tol = 2e-7
    np.testing.assert_allclose(m, z["seq_means"], atol=tol)
    std = np.sqrt(np.einsum("nij,nij->ni", c, c))
    b = np.sort(z["seq_bands"], axis=-1)
    np.testing.assert_allclose(b[..., 0], m - 2 * std, atol=tol)
    np.testing.assert_allclose(b[..., 1], m + 2 * std, atol=tol)


COARSE_CASES = [
    ("fitzhughnagumo", {}, 0.0, 100.0, 1000, 2),
    ("logistic", {}, 0.0, 10.0, 257, 3),
    ("lorenz63", {}, 0.0, 10.0, 400, 1),
    # ts[0] != 0: floor(ts / coarse_dt) runs past the coarse grid and clamps
    ("logistic", {}, 1.0, 11.0, 301, 2),
]


@pytest.mark.parametrize("name,kw,t0,t1,N,q", COARSE_CASES)
def test_coarse_init_matches_oracle(gpu, name, kw, t0, t1, N, q):
    import torch

    import pof.convenience as pc
    from oracle import model as om

    ivp, oivp = _pair(name, **kw)
    ts = np.linspace(t0, t1, N)
    print(f"coarse init {name} ts=[{t0}, {t1}] N={N} q={q}", flush=True)
    setup = pc.set_up_solver(f=ivp.f, y0=ivp.y0, ts=ts, order=q)
    st = pc.get_initial_trajectory(setup, method="coarse")
    torch.cuda.synchronize()
    ref = om.get_initial_trajectory(om.set_up_solver(oivp, ts, q), "coarse")
    tol = 1e-6 if name == "lorenz63" else 1e-9
    assert st.mean.shape == ref.mean.shape
    assert rel(st.mean.cpu().numpy(), ref.mean) < tol


@pytest.mark.parametrize("name,t1,N,q", [("rigid_body", 20.0, 200, 2), ("logistic", 10.0, 100, 2),
                                         ("fitzhughnagumo", 100.0, 500, 2)])
def test_solve_with_coarse_init(gpu, name, t1, N, q):
    """Converging cases only: from this init FHN with N=200, q=3 runs into maxiters and the
    oracle's own result moves by 12% under a 1e-15 perturbation of y0."""
    import torch

    import pof.solver
    from oracle import solver as osolver

    ivp, oivp = _pair(name)
    ts = np.linspace(0.0, t1, N)
    print(f"solve init=coarse {name} N={N} q={q}", flush=True)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init="coarse", maxiters=200)
    torch.cuda.synchronize()
    ref, rinfo = osolver.solve(oivp, ts, q, init="coarse", maxiters=200)
    assert info["iterations"] == rinfo["iterations"]
    # FHN over t in [0, 100] from the coarse trajectory amplifies rounding to ~1.5e-9
    assert rel(ys.mean.cpu().numpy(), ref.mean) < (1e-8 if name == "fitzhughnagumo" else 1e-9)


def test_coarse_init_in_a_busy_batch(gpu):
    """A coarse-init IVP admitted while another IVP is in flight: the EKS runs on its own workspace."""
    import torch

    import pof.batch
    import pof.ivp
    import pof.solver

    ivp = pof.ivp.logistic()
    b = pof.batch.Batch(1, 2, 1 << 14, 4)
    try:
        ta = b.admit(ivp.f, ivp.y0, np.linspace(0.0, 10.0, 300), init="constant", maxiters=200)
        b.launch()
        b.finish()
        tb = b.admit(ivp.f, ivp.y0, np.linspace(0.0, 10.0, 120), init="coarse", maxiters=200)
        res = {}
        while b.active:
            b.launch()
            for t, r in b.finish():
                res[t] = r
    finally:
        b.close()
    torch.cuda.synchronize()
    for t, N, init in ((ta, 300, "constant"), (tb, 120, "coarse")):
        alone, ainfo = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=np.linspace(0.0, 10.0, N), order=2, init=init,
                                        maxiters=200)
        ys, info = res[t]
        assert info["iterations"] == ainfo["iterations"]
        assert torch.equal(ys.mean, alone.mean)


def test_ieks_step_sequential_matches_oracle(gpu):
    import torch

    import pof.convenience as pc
    import pof.ivp
    import pof.step
    from oracle import model as om
    from oracle import problems as op
    from oracle import solver as osolver

    ts = np.linspace(0, 10, 40)
    ivp = pof.ivp.rigid_body()
    setup = pc.set_up_solver(f=ivp.f, y0=ivp.y0, ts=ts, order=2)
    states = pc.get_initial_trajectory(setup, method="constant")
    out, nll, obj, ssq = pof.step.ieks_step(states=states, setup=setup, sequential=True)
    torch.cuda.synchronize()
    osetup = om.set_up_solver(op.rigid_body(), ts, 2)
    ost = om.get_initial_trajectory(osetup, "constant")
    ref, rnll, robj, rssq = osolver.ieks_step(osetup, ost, is_sequential=True)
    assert rel(out.mean.cpu().numpy(), ref.mean) < 1e-10
    # factors carry sqrt(ssq); ssq itself is sign-sensitive for d = 3 (module docstring):
    # the oracle's own ssq of this step spans 0.35% under 1e-15..1e-14 perturbations of y0
    assert rel(cov(out.chol.cpu().numpy()) / ssq, cov(ref.chol) / rssq) < 1e-9
    assert abs(ssq - rssq) <= 0.01 * abs(rssq)
    assert close(nll, rnll, 1e-9) and close(obj, robj, 1e-9)


def test_sequential_linear_filtsmooth_random_model(gpu):
    import torch

    import pof.sequential_filtsmooth as sfs
    from oracle import model as om
    from oracle import sequential as oseq
    from pof.observations import AffineModel
    from pof.transitions import TransitionModel
    from pof.utils import MVNSqrt

    rng = np.random.default_rng(5)
    d, q, Ns = 2, 2, 37
    D = d * (q + 1)
    F, QL = om.preconditioned_discretize(om.IWP(d, q))
    H = rng.standard_normal((Ns, d, D))
    c = rng.standard_normal((Ns, d))
    m0 = rng.standard_normal(D)
    L0 = np.tril(rng.standard_normal((D, D)))
    x0 = om.MVNSqrt(m0, L0)
    dom = om.AffineModel(H, c, np.zeros((Ns, d, d)))
    ref, rnll, robj, rssq = oseq.linear_filtsmooth(x0, F, QL, dom)
    rfilt, _, _ = oseq.linear_noiseless_filter(x0, F, QL, dom)
    gdom = AffineModel(torch.as_tensor(H, device="cuda"), torch.as_tensor(c, device="cuda"), None)
    out, nll, obj, ssq = sfs.linear_filtsmooth(MVNSqrt(m0, L0), TransitionModel(F, QL), gdom)
    filt, fnll, fssq = sfs.linear_noiseless_filter(MVNSqrt(m0, L0), TransitionModel(F, QL), gdom)
    torch.cuda.synchronize()
    assert rel(out.mean.cpu().numpy(), ref.mean) < 1e-10
    assert rel(cov(out.chol.cpu().numpy()), cov(ref.chol)) < 1e-9
    assert rel(filt.mean.cpu().numpy(), rfilt.mean) < 1e-10
    assert rel(cov(filt.chol.cpu().numpy()), cov(rfilt.chol)) < 1e-9
    assert close(ssq, rssq, 1e-9) and close(nll, rnll, 1e-9) and close(obj, robj, 1e-9)
    assert fnll == nll and fssq == ssq


@pytest.mark.parametrize("sequential", [False, True])
def test_smoothing_on_given_filtered_states(gpu, sequential):
    """parallel_filtsmooth.smoothing / sequential_filtsmooth.smoothing over the oracle's filtered
    states of a random linear model (smoother.py:8-22, sequential_filtsmooth/smoother.py:8-28)."""
    import torch

    import pof.parallel_filtsmooth as pfs
    import pof.sequential_filtsmooth as sfs
    from oracle import model as om
    from oracle import parallel as opar
    from oracle import sequential as oseq
    from pof.transitions import TransitionModel
    from pof.utils import MVNSqrt

    rng = np.random.default_rng(31)
    d, q, Ns = 2, 2, 29
    D = d * (q + 1)
    F, QL = om.preconditioned_discretize(om.IWP(d, q))
    dom = om.AffineModel(rng.standard_normal((Ns, d, D)), rng.standard_normal((Ns, d)), np.zeros((Ns, d, d)))
    x0 = om.MVNSqrt(rng.standard_normal(D), np.tril(rng.standard_normal((D, D))))
    filt, _, _, _ = opar.linear_noiseless_filtering(x0, F, QL, dom)
    ref, robj = (oseq.smoothing if sequential else opar.smoothing)(F, QL, filt)
    fn = sfs.smoothing if sequential else pfs.smoothing
    dtm = TransitionModel(np.broadcast_to(F, (Ns, D, D)), np.broadcast_to(QL, (Ns, D, D)))
    out, obj = fn(dtm, MVNSqrt(torch.as_tensor(filt.mean, device="cuda"), torch.as_tensor(filt.chol, device="cuda")))
    torch.cuda.synchronize()
    assert rel(out.mean.cpu().numpy(), ref.mean) < 1e-10
    assert rel(cov(out.chol.cpu().numpy()), cov(ref.chol)) < 1e-9
    assert close(obj, robj, 1e-9)


@pytest.mark.parametrize("name,N,q", [("fitzhughnagumo", 200, 3), ("rigid_body", 64, 2)])
def test_extended_kalman_filter_and_filtsmooth(gpu, name, N, q):
    """sequential_filtsmooth.extended_kalman_filter / filtsmooth on the residual model of set_up_solver
    (filter.py:9-30, __init__.py:5-10) against the oracle."""
    import torch

    import pof.convenience as pc
    import pof.sequential_filtsmooth as sfs
    from oracle import model as om
    from oracle import sequential as oseq

    ivp, oivp = _pair(name)
    ts = np.linspace(oivp.t0, oivp.tmax, N)
    setup = pc.set_up_solver(f=ivp.f, y0=ivp.y0, ts=ts, order=q)
    os_ = om.set_up_solver(oivp, ts, q)
    filt, nll, ssq = sfs.extended_kalman_filter(setup["x0"], setup["dtm"], setup["om"])
    out, nll2, obj, ssq2 = sfs.filtsmooth(setup["x0"], setup["dtm"], setup["om"])
    torch.cuda.synchronize()
    rfilt, rnll, rssq = oseq.extended_kalman_filter(os_.x0, os_.F, os_.QL, os_, os_.Ns)
    rout, rnll2, robj, rssq2 = oseq.filtsmooth(os_.x0, os_.F, os_.QL, os_, os_.Ns)
    assert rel(filt.mean.cpu().numpy(), rfilt.mean) < 1e-9
    assert rel(cov(filt.chol.cpu().numpy()), cov(rfilt.chol)) < 1e-8
    assert close(nll, rnll, 1e-8)
    assert rel(out.mean.cpu().numpy(), rout.mean) < 1e-9
    assert rel(cov(out.chol.cpu().numpy()), cov(rout.chol)) < 1e-8
    assert close(obj, robj, 1e-6)
    assert nll2 == nll and ssq2 == ssq


def test_reference_equality_parallel_vs_sequential(gpu):
    """tests/test_filtsmooth.py:41-62 of the reference: q = 0, d = 1, all-zero model; the parallel and
    sequential filters and smoothers agree exactly."""
    import torch

    import pof.parallel_filtsmooth as pfs
    import pof.sequential_filtsmooth as sfs
    from pof.observations import AffineModel
    from pof.transitions import IWP, discretize_transitions
    from pof.utils import MVNSqrt

    iwp = IWP(num_derivatives=0, wiener_process_dimension=1)
    dtm = discretize_transitions(iwp, np.arange(10))
    Ns = dtm.F.shape[0]
    dom = AffineModel(torch.ones((Ns, 1, 1), dtype=torch.float64, device="cuda"),
                      torch.zeros((Ns, 1), dtype=torch.float64, device="cuda"), None)
    x0 = MVNSqrt(np.zeros(1), np.zeros((1, 1)))
    pf, *_ = pfs.linear_noiseless_filtering(x0, dtm, dom)
    ps, _ = pfs.smoothing(dtm, pf)
    sf, *_ = sfs.linear_noiseless_filter(x0, dtm, dom)
    ss, _ = sfs.smoothing(dtm, sf)
    torch.cuda.synchronize()
    assert pf.mean.shape == ps.mean.shape == sf.mean.shape == (10, 1)
    assert torch.equal(pf.mean, sf.mean) and torch.equal(pf.chol, sf.chol)
    assert torch.equal(ps.mean, ss.mean) and torch.equal(ps.chol, ss.chol)
