"""Stream ordering and failure clean-up of the C ABI.

* A call given the null stream (torch's default stream) must see what torch
  enqueued there before the call, although the library works on its own
  non-blocking stream: its input is produced here by a null-stream copy queued
  behind several large matrix products, with no synchronisation before the
  call, and the result must equal the one computed from synchronised input.
* A pof_solve_batch call that fails after admitting some IVPs must leave
  nothing in the context's reusable batches: the next solve of the same (D, d)
  succeeds, equals the solve before the failure, and never writes to the
  failed call's output buffers.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_ieks_step_orders_after_null_stream_producer(gpu):
    import torch

    import pof.convenience
    import pof.ivp
    import pof.step

    ivp = pof.ivp.lorenz63()
    ts = np.linspace(0.0, 2.0, 4097)
    setup = pof.convenience.set_up_solver(f=ivp.f, y0=ivp.y0, ts=ts, order=3)
    init = pof.convenience.get_initial_trajectory(setup, "constant").mean
    torch.cuda.synchronize()
    ref, nll_r, obj_r, ssq_r = pof.step.ieks_step(states=init, setup=setup)
    torch.cuda.synchronize()
    ref_m = ref.mean.clone()

    big = torch.randn(4096, 4096, dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    src = init.clone()
    dst = torch.zeros_like(init)
    print(f"caller stream handle {torch.cuda.current_stream().cuda_stream:#x}", flush=True)
    # about tens of ms of null-stream work ahead of the copy that produces the input
    for _ in range(8):
        big = (big @ big) * (1.0 / 4096.0)
    dst.copy_(src)
    out, nll, obj, ssq = pof.step.ieks_step(states=dst, setup=setup)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(big).all())
    assert torch.equal(out.mean, ref_m)
    assert (nll, obj, ssq) == (nll_r, obj_r, ssq_r)


def test_failed_batch_call_leaves_no_stale_ivp(gpu):
    import torch

    import pof.batch
    import pof.ivp
    import pof.solver
    from pof import _lib
    from pof.convenience import ProblemSpec

    ivp = pof.ivp.fhn_classic(a=0.2, b=0.2, c=3.0)
    ts = np.linspace(0.0, 20.0, 300)
    kw = dict(f=ivp.f, y0=ivp.y0, ts=ts, order=3, init="constant", maxiters=60)
    before, ib = pof.solver.solve(**kw)
    torch.cuda.synchronize()

    # two IVPs of the same (D, d); the second asks for FP32 products, which the
    # per-worker engines refuse after the first has been admitted
    specs = [ProblemSpec(ivp.f, ivp.y0, ts, 3, "constant"), ProblemSpec(ivp.f, ivp.y0 * 1.01, ts, 3, "constant")]
    probs = (_lib.Problem * 2)(*[s.cstruct() for s in specs])
    opts = (_lib.Opts * 2)(pof.batch._opts(True, 60, None), pof.batch._opts(True, 60, None))
    opts[1].products = _lib.PRODUCTS_F32
    sentinel = [torch.full((len(ts), 2), float("nan"), dtype=torch.float64, device=gpu) for _ in range(2)]
    outs = (_lib.Out * 2)()
    for o, t in zip(outs, sentinel):
        o.means_dev = t.data_ptr()
    torch.cuda.synchronize()
    ctx = _lib.context(0)
    rc = _lib.load().pof_solve_batch(ctx, probs, opts, outs, 2)
    assert rc == _lib.POF_E_UNSUPPORTED

    after, ia = pof.solver.solve(**kw)
    torch.cuda.synchronize()
    assert ia["iterations"] == ib["iterations"] and ia["status"] == ib["status"]
    assert torch.equal(after.mean, before.mean) and torch.equal(after.chol, before.chol)
    # the refused call's IVPs never ran: their output buffers are untouched
    for t in sentinel:
        assert bool(torch.isnan(t).all())
    # and a batch of the same group works again
    res = pof.batch.solve_batch([(ivp.f, ivp.y0, ts, 3, "constant")], maxiters=60)
    assert torch.equal(res[0][0].mean, before.mean)
