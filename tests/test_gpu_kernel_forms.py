"""The two forms of the scan kernels give bitwise the same results.

Each scan kernel of state dimension D <= 16 has a latency form (a matrix row
split over 2 or 4 lanes) and a throughput form (one lane per row that
reproduces the split form's reduction order: the same partial sums, chains and
combination order).  Which one a scan level runs is a performance choice
(pof_tuning_set, the autotuner's cache); results must not depend on it.  Here
every level runs in one form, then in the other, for each (D, d) of the
batched mix, and the outputs must be identical (torch.equal), as must the
batch results against solves alone whatever form each level took.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NEVER = 1 << 30


def _problem(d, q, N, seed=0):
    import pof.ivp

    f = {1: lambda: pof.ivp.logistic(r=1.7), 2: lambda: pof.ivp.fhn_classic(),
         3: lambda: pof.ivp.lorenz63(tmax=2.0), 4: lambda: pof.ivp.lorenz96(d=4, tmax=1.0)}[d]()
    rng = np.random.default_rng(seed)
    return f.f, f.y0 + 0.01 * rng.standard_normal(d), np.linspace(f.t0, f.tmax, N), q, "constant"


@pytest.mark.parametrize("d,q", [(1, 1), (1, 3), (1, 4), (2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3),
                                 (3, 4), (4, 1), (4, 3)])
def test_forms_are_bitwise_equal(gpu, d, q):
    import torch

    import pof.autotune
    import pof.solver

    D = d * (q + 1)
    f, y0, ts, q, init = _problem(d, q, 300)
    saved = pof.autotune.get_widths(D)
    out = []
    try:
        for w in (NEVER, 1):
            pof.autotune.set_widths(D, [w] * 4)
            print(f"d={d} q={q} D={D} width {w}", flush=True)
            out.append(pof.solver.solve(f=f, y0=y0, ts=ts, order=q, init=init, maxiters=25,
                                        return_full_states=True))
            torch.cuda.synchronize()
    finally:
        pof.autotune.set_widths(D, saved)
    (a, ia), (b, ib) = out
    assert ia["iterations"] == ib["iterations"]
    assert torch.equal(a.mean, b.mean) and torch.equal(a.chol, b.chol)
    assert torch.equal(ia["states"].mean, ib["states"].mean) and torch.equal(ia["states"].chol, ib["states"].chol)
    assert (ia["nll"], ia["obj"], ia["sigma_squared"]) == (ib["nll"], ib["obj"], ib["sigma_squared"])


def test_batch_in_throughput_form_equals_solves_alone(gpu):
    """A mixed batch whose every level runs in the throughput form, against each IVP solved alone
    in the latency form: bitwise equal."""
    import torch

    import pof.autotune
    import pof.batch
    import pof.solver

    probs = [_problem(d, q, 100 + 37 * i, seed=i) for i, (d, q) in enumerate(
        [(1, 2), (2, 3), (3, 3), (3, 4), (2, 1), (1, 4), (3, 2), (4, 1)])]
    Ds = sorted({p[0].d * (p[3] + 1) for p in probs})
    saved = {D: pof.autotune.get_widths(D) for D in Ds}
    try:
        for D in Ds:
            pof.autotune.set_widths(D, [NEVER] * 4)
        alone = [pof.solver.solve(f=f, y0=y0, ts=ts, order=q, init=init, maxiters=30) for f, y0, ts, q, init in probs]
        for D in Ds:
            pof.autotune.set_widths(D, [1] * 4)
        res = pof.batch.solve_batch(probs, maxiters=30)
        torch.cuda.synchronize()
    finally:
        for D, w in saved.items():
            pof.autotune.set_widths(D, w)
    for (ys1, i1), (ys, i) in zip(alone, res):
        assert i["iterations"] == i1["iterations"]
        assert torch.equal(ys.mean, ys1.mean) and torch.equal(ys.chol, ys1.chol)
