"""IEKS iterations replayed from a HIP graph (``pof_graphs_enable``, on by
default) against the same iterations launched kernel by kernel: results bitwise
equal (``torch.equal``) and the same info.

Replays alternate with the program's own work: a run to convergence copies the
reductions back and synchronises after every replay, a fixed-iteration run
replays back to back, and the batch test runs torch kernels and
synchronisations between steps while its IVPs retire at different iterations
(so graphs are captured, replayed, dropped and captured again).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_both(**kw):
    import torch

    import pof.solver
    from pof import _lib

    out = {}
    try:
        for on in (False, True):
            _lib.set_graphs(on)
            (m, L), info = pof.solver.solve(**kw)
            torch.cuda.synchronize()
            out[on] = (m.clone(), L.clone(), info)
    finally:
        _lib.set_graphs(True)
    return out


def _assert_same(out):
    import torch

    (m0, L0, i0), (m1, L1, i1) = out[False], out[True]
    assert torch.equal(m0, m1)
    assert torch.equal(L0, L1)
    for k in ("iterations", "nll", "obj", "sigma_squared", "status"):
        assert i0[k] == i1[k], k


def test_converging_solve_same_with_and_without_graphs(gpu):
    import pof.ivp

    ivp = pof.ivp.fitzhughnagumo()
    out = _solve_both(f=ivp.f, y0=ivp.y0, ts=np.linspace(ivp.t0, ivp.tmax, 100), order=3, init="constant")
    assert out[True][2]["iterations"] > 10  # most iterations were graph replays
    _assert_same(out)


def test_fixed_iterations_same_with_and_without_graphs(gpu):
    import pof.ivp

    ivp = pof.ivp.lorenz63()
    out = _solve_both(f=ivp.f, y0=ivp.y0, ts=np.linspace(0.0, 1.0, 4097), order=3, init="constant", fixed_iters=10)
    _assert_same(out)


@pytest.mark.parametrize("name,order,init,sequential", [
    ("rigid_body", 2, "prior", False),
    ("logistic", 2, "constant", True),
])
def test_other_solves_same_with_and_without_graphs(gpu, name, order, init, sequential):
    import pof.ivp

    ivp = getattr(pof.ivp, name)()
    out = _solve_both(f=ivp.f, y0=ivp.y0, ts=np.linspace(ivp.t0, ivp.tmax, 300), order=order, init=init,
                      sequential=sequential)
    _assert_same(out)


def test_batch_steps_same_with_and_without_graphs(gpu):
    import torch

    import pof.batch
    import pof.ivp
    from pof import _lib

    ivp = pof.ivp.fhn_classic()
    fixed = [3, 5, 8, 13]

    def run(on):
        _lib.set_graphs(on)
        b = pof.batch.Batch(d=2, order=3, entries=4096, slots=8)
        tickets = {}
        for i, k in enumerate(fixed):
            ts = np.linspace(0.0, 10.0, 50 + 17 * i)
            tickets[b.admit(ivp.f, ivp.y0 + 0.01 * i, ts, init="constant", fixed_iters=k)] = i
        x = torch.randn(256, 256, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
        res, steps = {}, 0
        while b.active:
            done = b.step()
            steps += 1
            for t, (ys, info) in done:
                res[tickets[t]] = (ys.mean.clone(), ys.chol.clone(), info, steps)
            y = (x @ x).sum()  # the program's own work between replays
            torch.cuda.synchronize()
            assert torch.isfinite(y)
        b.close()
        return res

    try:
        eager = run(False)
        replay = run(True)
    finally:
        _lib.set_graphs(True)
    assert sorted(eager) == sorted(replay) == list(range(len(fixed)))
    for i, k in enumerate(fixed):
        m0, L0, i0, s0 = eager[i]
        m1, L1, i1, s1 = replay[i]
        assert s0 == s1 == k
        assert torch.equal(m0, m1) and torch.equal(L0, L1), i
        for key in ("iterations", "nll", "obj", "sigma_squared", "status"):
            assert i0[key] == i1[key], (i, key)
