"""pof_linearize_model_noalloc (no device allocation) against pof_linearize_model.

Both run the same kernel on the same inputs; the variant only takes its device
buffer (P and the per-state work slices) from the library's file-scope scratch
array instead of a hipMalloc per call.  Outputs must be bitwise equal, on two
calls in a row with the second on fewer states (the scratch keeps the first
call's contents), for every linearisation method.  Sizes beyond the scratch
are refused with POF_E_ARG before anything is launched.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

METHODS = ["ek1", "ek0", "uncertain", "regularized", "unscented"]


def _states(n, D, seed, gpu):
    import torch

    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, D))
    L = np.tril(0.05 * rng.standard_normal((n, D, D)), -1) + 0.3 * np.eye(D)[None]
    return (torch.as_tensor(m, device=gpu).contiguous(), torch.as_tensor(L, device=gpu).contiguous())


def _run(fn, ctx, prob, method, m, L, params, gpu):
    import torch

    from pof import _lib

    n, D = m.shape
    d = prob.d
    dO = d + D if method == "regularized" else d
    H = torch.full((n, dO, D), np.nan, dtype=torch.float64, device=gpu)
    c = torch.full((n, dO), np.nan, dtype=torch.float64, device=gpu)
    R = torch.full((n, dO, dO), np.nan, dtype=torch.float64, device=gpu)
    p = (ctypes.c_double * 4)(*params)
    torch.cuda.synchronize()
    rc = fn(ctx, ctypes.byref(prob), _lib.LIN_METHODS[method], p, _lib.dptr(m), _lib.dptr(L), n, _lib.dptr(H),
            _lib.dptr(c), _lib.dptr(R), _lib.stream_ptr())
    _lib.check(rc, ctx, method)
    torch.cuda.synchronize()
    return H, c, R


@pytest.mark.parametrize("method", METHODS)
def test_noalloc_matches_allocating_entry_point(gpu, method):
    import torch

    import pof.ivp
    from pof import _lib
    from pof.convenience import ProblemSpec

    lib = _lib.load()
    ctx = _lib.context(0)
    ivp = pof.ivp.lorenz63()
    q = 4
    D = 3 * (q + 1)
    prob = ProblemSpec(ivp.f, ivp.y0, np.linspace(0.0, 1.0, 11), q, "constant").cstruct()
    params = [0.5, 1.0, 0.0, np.nan]
    for call, (n, seed) in enumerate([(3000, 1), (700, 2)]):
        print(f"method {method} call {call} n {n}", flush=True)
        m, L = _states(n, D, seed, gpu)
        ref = _run(lib.pof_linearize_model, ctx, prob, method, m, L, params, gpu)
        got = _run(lib.pof_linearize_model_noalloc, ctx, prob, method, m, L, params, gpu)
        for a, b in zip(got, ref):
            assert bool(torch.isfinite(b).all())
            assert torch.equal(a, b)


def test_noalloc_refuses_sizes_beyond_its_scratch(gpu):
    import torch

    import pof.ivp
    from pof import _lib
    from pof.convenience import ProblemSpec

    lib = _lib.load()
    ctx = _lib.context(0)
    ivp = pof.ivp.lorenz63()
    prob = ProblemSpec(ivp.f, ivp.y0, np.linspace(0.0, 1.0, 11), 4, "constant").cstruct()
    # 362 doubles of work per state at D = 15, d = 3: 30,000 states need more than the 64 MiB array;
    # the size is checked on the host before any launch
    n, D, d = 30000, 15, 3
    m = torch.zeros((n, D), dtype=torch.float64, device=gpu)
    L = torch.zeros((n, D, D), dtype=torch.float64, device=gpu)
    H = torch.zeros((n, d, D), dtype=torch.float64, device=gpu)
    c = torch.zeros((n, d), dtype=torch.float64, device=gpu)
    R = torch.zeros((n, d, d), dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    rc = lib.pof_linearize_model_noalloc(ctx, ctypes.byref(prob), _lib.LIN_METHODS["unscented"], None,
                                         _lib.dptr(m), _lib.dptr(L), n, _lib.dptr(H), _lib.dptr(c), _lib.dptr(R),
                                         _lib.stream_ptr())
    assert rc == _lib.POF_E_ARG
    assert "scratch" in lib.pof_last_error(ctx).decode()
    assert float(H.abs().max()) == 0.0
