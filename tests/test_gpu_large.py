"""GPU parity of the large-state engine (batched LQ / GEMM primitives per scan level).

It runs the same IEKS as the per-worker kernels, so on problems both engines
take it must agree with them and with the oracle; on Lorenz-96 with d = 32,
q = 3 (D = 128, the size the per-worker kernels cannot hold) it is compared
with the oracle directly.

Tolerances (float64): iterations equal; means 1e-9 relative (1e-8 for
Lorenz-96, whose 128-wide products sum in a different order than numpy's);
``L L^T / ssq^2`` 1e-6.  ``sigma_squared`` depends on the column signs of the
factors for d >= 2 (as for the other engines): compared to the oracle within
1.5x the oracle's own spread under 1e-15 perturbations of y0, at least 1e-8.
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


def _pair(name, **kw):
    import pof.ivp
    from oracle import problems as op

    return getattr(pof.ivp, name)(**kw), getattr(op, name)(**kw)


def _ssq_tol(oivp, ts, q, init, s_ref, fixed_iters):
    from oracle import problems as op
    from oracle import solver as osolver

    # sigma_squared of d >= 2 depends on factor signs that rounding decides: its tolerance is
    # twice the oracle's own spread over six relative perturbations of y0 at 1e-15 .. 1e-14
    spread = 0.0
    for eps in (1e-15, -1e-15, 3e-15, -3e-15, 1e-14, -1e-14):
        pp = op.Problem(oivp.name, oivp.params, oivp.y0 * (1 + eps), oivp.t0, oivp.tmax)
        _, inf = osolver.solve(pp, ts, q, init=init, maxiters=200, fixed_iters=fixed_iters)
        spread = max(spread, abs(inf["sigma_squared"] - s_ref))
    return max(1e-8 * abs(s_ref), 2.0 * spread)


@pytest.mark.parametrize("name,kw,N,q,init", [
    ("fitzhughnagumo", {}, 100, 3, "constant"),
    ("logistic", {}, 32, 3, "prior"),
    ("rigid_body", {}, 64, 2, "constant"),
])
def test_large_engine_matches_per_worker_engine(gpu, name, kw, N, q, init):
    import torch

    import pof.solver

    ivp, oivp = _pair(name, **kw)
    ts = np.linspace(oivp.t0, oivp.tmax, N)
    print(f"large vs per-worker {name} N={N} q={q} init={init}", flush=True)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init=init, maxiters=200, engine="large")
    yr, rinfo = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init=init, maxiters=200)
    torch.cuda.synchronize()
    assert info["iterations"] == rinfo["iterations"]
    assert rel(ys.mean.cpu().numpy(), yr.mean.cpu().numpy()) < 1e-9
    s1, s2 = info["sigma_squared"], rinfo["sigma_squared"]
    assert rel(cov(ys.chol.cpu().numpy()) / s1**2, cov(yr.chol.cpu().numpy()) / s2**2) < 1e-6
    assert abs(info["obj"] - rinfo["obj"]) <= 1e-6 * abs(rinfo["obj"]) + 1e-12


def test_large_engine_fixed_iterations_lorenz63(gpu):
    import torch

    import pof.solver
    from oracle import solver as osolver

    ivp, oivp = _pair("lorenz63")
    ts = np.linspace(0.0, 10.0, 257)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=3, init="constant", fixed_iters=4, engine="large")
    torch.cuda.synchronize()
    ref, rinfo = osolver.solve(oivp, ts, 3, init="constant", fixed_iters=4)
    assert info["iterations"] == 4.0 and info["status"] == "fixed"
    assert rel(ys.mean.cpu().numpy(), ref.mean) < 1e-6


@pytest.mark.parametrize("d,q,N,fixed", [(32, 3, 17, 3), (8, 2, 33, None), (5, 4, 17, 3)])
def test_lorenz96_matches_oracle(gpu, d, q, N, fixed):
    """d = 32, q = 3 is the D = 128 configuration (short horizon, same dt as t in [0, 5] with 8192
    steps); d = 8, q = 2 (D = 24) is beyond the per-worker kernels too and runs to convergence;
    d = 5, q = 4 (D = 25) leaves partial 16-row blocks in the LQ panels and triangular solves and an
    odd vector length in the matrix-vector chains."""
    import torch

    import pof.solver
    from oracle import solver as osolver

    ivp, oivp = _pair("lorenz96", d=d)
    ts = np.linspace(0.0, (N - 1) * 5.0 / 8192, N)
    print(f"lorenz96 d={d} q={q} N={N} fixed={fixed}", flush=True)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init="constant", maxiters=200,
                                fixed_iters=fixed)
    torch.cuda.synchronize()
    ref, rinfo = osolver.solve(oivp, ts, q, init="constant", maxiters=200, fixed_iters=fixed)
    assert info["iterations"] == rinfo["iterations"]
    assert ys.mean.shape == (N, d) and ys.chol.shape == (N, d, d * (q + 1))
    assert rel(ys.mean.cpu().numpy(), ref.mean) < 1e-8
    s_gpu, s_ref = info["sigma_squared"], rinfo["sigma_squared"]
    assert rel(cov(ys.chol.cpu().numpy()) / s_gpu**2, cov(ref.chol) / s_ref**2) < 1e-6
    assert abs(s_gpu - s_ref) <= _ssq_tol(oivp, ts, q, "constant", s_ref, fixed)
    assert abs(info["obj"] - rinfo["obj"]) <= 1e-6 * abs(rinfo["obj"]) + 1e-12


def test_lorenz96_fp32_combine_products_fenced(gpu):
    """FP32 combine products (the scan combines' matrix products as FP32 MFMA GEMMs, everything
    else FP64) are refused by the public API: the means miss the 1e-6 bar against the FP64 oracle
    and sigma_squared, which scales every returned factor, is rounding noise in FP32.  The C ABI
    keeps them as an experiment with calibrate = 0 only; their means are compared here with the
    FP64 oracle directly (no HIP-to-HIP comparison), at the 1e-5 they reach, which is why the
    public API refuses them."""
    import ctypes

    import torch

    import pof.ivp
    import pof.solver
    from oracle import problems as op
    from oracle import solver as osolver
    from pof import _lib
    from pof.convenience import ProblemSpec

    ivp = pof.ivp.lorenz96(d=32)
    ts = np.linspace(0.0, 16 * 5.0 / 8192, 17)
    kw = dict(f=ivp.f, y0=ivp.y0, ts=ts, order=3, init="constant", fixed_iters=3)
    with pytest.raises(NotImplementedError):
        pof.solver.solve(products="float32", **kw)
    with pytest.raises(ValueError):
        pof.solver.solve(products="float16", **kw)
    lib = _lib.load()
    ctx = _lib.context(0)
    spec = ProblemSpec(ivp.f, ivp.y0, ts, 3, "constant")
    prob = spec.cstruct()
    N, d, D = len(ts), 32, 128
    means = torch.empty((N, d), dtype=torch.float64, device="cuda")
    out = _lib.Out()
    out.means_dev = means.data_ptr()
    opts = _lib.Opts()
    opts.maxiters, opts.fixed_iters, opts.products = 10, 3, _lib.PRODUCTS_F32
    opts.calibrate = 1
    rc = lib.pof_solve_large(ctx, ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(out), _lib.stream_ptr())
    assert rc == _lib.POF_E_UNSUPPORTED
    opts.calibrate = 0
    rc = lib.pof_solve_large(ctx, ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(out), _lib.stream_ptr())
    _lib.check(rc, ctx, "pof_solve_large")
    torch.cuda.synchronize()
    ref, _ = osolver.solve(op.lorenz96(d=32), ts, 3, init="constant", fixed_iters=3)
    err = rel(means.cpu().numpy(), ref.mean)
    print(f"fp32 combine products: means vs the FP64 oracle {err:.2e}", flush=True)
    assert err < 1e-5
