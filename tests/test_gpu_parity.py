"""GPU parity: the HIP path through the C ABI against the FP64 CPU oracle.

Tolerances (float64 everywhere): means agree to 1e-9 relative to the largest
mean entry (the oracle follows the same algorithm; differences are rounding);
iteration counts are equal.  The acceptance bar for means is 1e-6 relative.

Covariances: the returned factors are ``ssq * L`` (double calibration), and
only ``c c^T`` is unique.  ``L L^T`` is a function of the model alone and is
compared to 1e-6.  ``ssq`` (sigma_squared) is not: the reference whitens with
``L^T`` (pof/utils.py:110-112), so for d >= 2 it depends on the column signs
of the factors, and with noiseless observations many Householder pivots of
the rank-deficient factors are pure rounding noise.  The oracle itself moves
ssq by up to ~1% under a 1e-15 perturbation of y0 for some problems.  So ssq
must agree to 1e-8 where the oracle's own value is stable under that
perturbation (every d = 1 case, FHN, Lorenz-63) and otherwise lie within
twice the oracle's measured spread.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MEAN_TOL = 1e-9
COV_TOL = 1e-6


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def cov(c):
    return np.einsum("nij,nkj->nik", c, c)


def ssq_tolerance(oivp, ts, q, init, s_ref, fixed_iters=None):
    """1e-8 relative, or 1.5x the oracle's own spread under 1e-15..1e-14 perturbations of y0."""
    from oracle import problems as op
    from oracle import solver as osolver

    spread = 0.0
    for eps in (1e-15, -1e-15, 3e-15, -3e-15, 1e-14, -1e-14):
        pp = op.Problem(oivp.name, oivp.params, oivp.y0 * (1 + eps), oivp.t0, oivp.tmax)
        _, inf = osolver.solve(pp, ts, q, init=init, maxiters=200, fixed_iters=fixed_iters)
        spread = max(spread, abs(inf["sigma_squared"] - s_ref))
    return max(1e-8 * abs(s_ref), 1.5 * spread)


def _pair(name, **kw):
    import pof.ivp
    from oracle import problems as op

    return getattr(pof.ivp, name)(**kw), getattr(op, name)(**kw)


CASES = [
    ("logistic", {}, 16, 1, "constant"),
    ("logistic", {}, 16, 2, "constant"),
    ("logistic", {}, 16, 3, "constant"),
    ("logistic", {}, 32, 3, "prior"),
    ("logistic", {}, 21, 1, "prior"),
    ("fitzhughnagumo", {}, 100, 3, "constant"),
    ("rigid_body", {}, 128, 3, "constant"),
    ("vanderpol", {"stiffness_constant": 1.0}, 128, 2, "constant"),
    ("henonheiles", {"tmax": 10.0}, 32, 2, "constant"),
    ("lotkavolterra", {}, 50, 2, "prior"),
    # d = 3, q = 4: D = 15, the largest state of the batched mix (its own kernel instantiations)
    ("lorenz63", {"tmax": 0.5}, 129, 4, "constant"),
    ("lorenz63", {"tmax": 0.5}, 65, 4, "prior"),
    ("rigid_body", {}, 128, 4, "constant"),
]


@pytest.mark.parametrize("name,kw,N,q,init", CASES)
def test_solve_matches_oracle(gpu, name, kw, N, q, init):
    import torch

    import pof.solver
    from oracle import solver as osolver

    ivp, oivp = _pair(name, **kw)
    ts = np.linspace(oivp.t0, oivp.tmax, N)
    print(f"case {name} N={N} q={q} init={init}", flush=True)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init=init, maxiters=200)
    torch.cuda.synchronize()
    ref, rinfo = osolver.solve(oivp, ts, q, init=init, maxiters=200)
    assert info["iterations"] == rinfo["iterations"]
    assert rel(ys.mean.cpu().numpy(), ref.mean) < MEAN_TOL
    s_gpu, s_ref = info["sigma_squared"], rinfo["sigma_squared"]
    # covariance shape L L^T (sign-free): divide out ssq^2
    assert rel(cov(ys.chol.cpu().numpy()) / s_gpu**2, cov(ref.chol) / s_ref**2) < COV_TOL
    tol = ssq_tolerance(oivp, ts, q, init, s_ref)
    assert abs(s_gpu - s_ref) <= tol, (s_gpu, s_ref, tol)
    assert abs(info["obj"] - rinfo["obj"]) <= 1e-6 * abs(rinfo["obj"]) + 1e-12
    assert abs(info["nll"] - rinfo["nll"]) <= 1e-8 * abs(rinfo["nll"]) + 1e-8
    assert ys.mean.shape == (N, oivp.d)
    assert ys.chol.shape == (N, oivp.d, oivp.d * (q + 1))


@pytest.mark.parametrize("q", [3, 4])
def test_fixed_iterations_lorenz63(gpu, q):
    import torch

    import pof.ivp
    import pof.solver
    from oracle import problems as op
    from oracle import solver as osolver

    ts = np.linspace(0.0, 10.0, 257)
    ivp = pof.ivp.lorenz63()
    print(f"lorenz63 N=257 q={q} 4 fixed iterations", flush=True)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init="constant", fixed_iters=4)
    torch.cuda.synchronize()
    ref, rinfo = osolver.solve(op.lorenz63(), ts, q, init="constant", fixed_iters=4)
    assert info["iterations"] == 4.0
    # chaotic field, unconverged iterates: rounding is amplified; acceptance bar 1e-6
    assert rel(ys.mean.cpu().numpy(), ref.mean) < 1e-6
    s_gpu, s_ref = info["sigma_squared"], rinfo["sigma_squared"]
    assert rel(cov(ys.chol.cpu().numpy()) / s_gpu**2, cov(ref.chol) / s_ref**2) < COV_TOL
    assert abs(s_gpu - s_ref) <= ssq_tolerance(op.lorenz63(), ts, q, "constant", s_ref, fixed_iters=4)


def test_linear_filtsmooth_random_model(gpu):
    import torch

    from oracle import model as om
    from oracle import parallel as opar
    import pof.parallel_filtsmooth as pfs
    from pof.observations import AffineModel
    from pof.transitions import TransitionModel
    from pof.utils import MVNSqrt

    rng = np.random.default_rng(3)
    d, q, Ns = 2, 2, 37
    D = d * (q + 1)
    F, QL = om.preconditioned_discretize(om.IWP(d, q))
    H = rng.standard_normal((Ns, d, D))
    c = rng.standard_normal((Ns, d))
    m0 = rng.standard_normal(D)
    L0 = np.tril(rng.standard_normal((D, D)))
    x0 = om.MVNSqrt(m0, L0)
    ref, rnll, robj, rssq = opar.linear_filtsmooth(x0, F, QL, om.AffineModel(H, c, np.zeros((Ns, d, d))))
    out, nll, obj, ssq = pfs.linear_filtsmooth(
        MVNSqrt(m0, L0), TransitionModel(F, QL),
        AffineModel(torch.as_tensor(H, device="cuda"), torch.as_tensor(c, device="cuda"), None))
    torch.cuda.synchronize()
    assert rel(out.mean.cpu().numpy(), ref.mean) < 1e-10
    assert rel(cov(out.chol.cpu().numpy()), cov(ref.chol)) < 1e-9
    assert abs(ssq - rssq) <= 1e-9 * abs(rssq)
    assert abs(nll - rnll) <= 1e-9 * abs(rnll)
    assert abs(obj - robj) <= 1e-9 * abs(robj)


def test_reference_degenerate_equality(gpu):
    """tests/test_filtsmooth.py:41-62 of the reference: q = 0, d = 1, all-zero model."""
    import torch

    import pof.parallel_filtsmooth as pfs
    from oracle import model as om
    from pof.observations import AffineModel
    from pof.transitions import IWP, discretize_transitions
    from pof.utils import MVNSqrt

    iwp = IWP(num_derivatives=0, wiener_process_dimension=1)
    dtm = discretize_transitions(iwp, np.arange(10))
    Ns = dtm.F.shape[0]
    H = torch.ones((Ns, 1, 1), dtype=torch.float64, device="cuda")
    c = torch.zeros((Ns, 1), dtype=torch.float64, device="cuda")
    out, filt, *_ = pfs._run(MVNSqrt(np.zeros(1), np.zeros((1, 1))), dtm, AffineModel(H, c, None), True)
    torch.cuda.synchronize()
    assert out.mean.shape == (10, 1) and out.chol.shape == (10, 1, 1)
    assert torch.all(out.mean == 0) and torch.all(filt.mean == 0)
    # the filtered/smoothed factors are zero: every state is observed exactly
    assert torch.all(filt.chol == 0) and torch.all(out.chol == 0)


def test_linearize_matches_oracle(gpu):
    import torch

    import pof.convenience as pc
    import pof.ivp
    from oracle import model as om
    from oracle import problems as op

    ts = np.linspace(0, 10, 11)
    ivp = pof.ivp.lorenz63()
    setup = pc.set_up_solver(f=ivp.f, y0=ivp.y0, ts=ts, order=2)
    osetup = om.set_up_solver(op.lorenz63(), ts, 2)
    rng = np.random.default_rng(0)
    m = rng.standard_normal((7, 9))
    lin = pc.linearize_observation_model(setup["om"], torch.as_tensor(m, device="cuda"))
    ref = om.linearize(osetup, m)
    assert rel(lin.H.cpu().numpy(), ref.H) < 1e-13
    assert rel(lin.b.cpu().numpy(), ref.b) < 1e-12


def test_ieks_step_matches_oracle(gpu):
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
    out, nll, obj, ssq = pof.step.ieks_step(states=states, setup=setup)
    torch.cuda.synchronize()
    osetup = om.set_up_solver(op.rigid_body(), ts, 2)
    ost = om.get_initial_trajectory(osetup, "constant")
    ref, rnll, robj, rssq = osolver.ieks_step(osetup, ost)
    assert rel(states.mean.cpu().numpy(), ost.mean) < 1e-15
    assert rel(out.mean.cpu().numpy(), ref.mean) < 1e-10
    # the factor carries sqrt(ssq) (step.py:42-44): L L^T / ssq to 1e-8.  ssq itself depends on the
    # column signs of the observation factors for d = 3 (utils.py:110-112 whitens with L^T), and from
    # the constant trajectory some Householder pivots are exactly zero in exact arithmetic, so their
    # signs are rounding noise: two summation orders give ssq 0.19% apart here while a 1e-15 change
    # of y0 moves the oracle's by 1e-13 only.  Held to 1%; pinned where it is well conditioned (d = 1,
    # FHN) by the solve parity tests.
    assert rel(cov(out.chol.cpu().numpy()) / ssq, cov(ref.chol) / rssq) < 1e-8
    assert abs(ssq - rssq) <= 1e-2 * abs(rssq), (ssq, rssq)


@pytest.mark.parametrize("N", [2, 3, 4, 5])
@pytest.mark.parametrize("sequential", [False, True])
def test_tiny_grids_match_oracle(gpu, N, sequential):
    """The smallest grids (Ns = 1 .. 4: scans of one or two levels, a single combine or none)
    against the oracle, parallel and sequential; same tolerances as test_solve_matches_oracle."""
    import torch

    import pof.solver
    from oracle import solver as osolver

    ivp, oivp = _pair("logistic")
    ts = np.linspace(0.0, 1.0, N)
    print(f"tiny grid N={N} sequential={sequential}", flush=True)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=2, init="constant", maxiters=100,
                                sequential=sequential)
    torch.cuda.synchronize()
    ref, rinfo = osolver.solve(oivp, ts, 2, init="constant", maxiters=100, is_sequential=sequential)
    assert ys.mean.shape == (N, 1) and ys.chol.shape == (N, 1, 3)
    assert info["iterations"] == rinfo["iterations"]
    assert rel(ys.mean.cpu().numpy(), ref.mean) < 1e-9
    assert abs(info["obj"] - rinfo["obj"]) <= 1e-6 * abs(rinfo["obj"]) + 1e-12
