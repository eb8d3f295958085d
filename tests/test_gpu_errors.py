"""Error paths of the GPU entry points: each refusal is an error status with a
message (raised by pof._lib.check as NotImplementedError for an unsupported
request, ValueError for an invalid argument), and nothing is left running on
the device."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_state_dimension_beyond_every_engine(gpu):
    import pof.ivp
    import pof.solver

    ivp = pof.ivp.lorenz96(d=32)  # q = 4: D = 160, beyond the large-state engine's 128 and every per-worker kernel
    with pytest.raises(NotImplementedError, match="D=160"):
        pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=np.linspace(0.0, 0.1, 5), order=4, init="constant")
    # an order beyond the Taylor-mode initialisation is refused before that, as an invalid argument
    with pytest.raises(ValueError, match="Taylor-mode"):
        pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=np.linspace(0.0, 0.1, 5), order=20, init="constant")


def test_large_engine_refuses_the_sequential_branch(gpu):
    import pof.ivp
    import pof.solver

    ivp = pof.ivp.lorenz96(d=32)
    with pytest.raises(NotImplementedError, match="parallel IEKS only"):
        pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=np.linspace(0.0, 0.01, 5), order=3, init="constant",
                         sequential=True)


def test_batch_refuses_admission_while_an_iteration_is_in_flight(gpu):
    import pof.batch
    import pof.ivp

    ivp = pof.ivp.logistic()
    b = pof.batch.Batch(d=1, order=2, entries=20000, slots=4)
    b.admit(ivp.f, ivp.y0, np.linspace(0.0, 1.0, 50), init="constant", maxiters=20)
    b.launch()
    with pytest.raises(ValueError, match="between pof_batch_launch and pof_batch_finish"):
        b.admit(ivp.f, ivp.y0, np.linspace(0.0, 1.0, 30), init="constant")
    b.finish()
    while b.active:
        b.step()
    b.close()


def test_batch_refuses_a_problem_of_another_dimension(gpu):
    import pof.batch
    import pof.ivp

    b = pof.batch.Batch(d=2, order=3, entries=20000, slots=4)
    ivp = pof.ivp.logistic()
    with pytest.raises(ValueError, match="D="):
        b.admit(ivp.f, ivp.y0, np.linspace(0.0, 1.0, 30), init="constant")
    b.close()


def test_fp32_products_refused_outside_the_large_engine(gpu):
    """The per-worker engines are FP64 only: a batch admission asking for FP32 products is refused."""
    import torch

    import pof.batch
    import pof.ivp
    from pof import _lib
    from pof.convenience import ProblemSpec

    ivp = pof.ivp.logistic()
    b = pof.batch.Batch(d=1, order=2, entries=20000, slots=4)
    spec = ProblemSpec(ivp.f, ivp.y0, np.linspace(0.0, 1.0, 20), 2, "constant")
    prob = spec.cstruct()
    opts = pof.batch._opts(True, 10, None)
    opts.products = _lib.PRODUCTS_F32
    means = torch.empty((20, 1), dtype=torch.float64, device="cuda")
    out = _lib.Out()
    out.means_dev = means.data_ptr()
    ticket = ctypes.c_int64()
    rc = b._lib.pof_batch_admit(b._h, ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(out),
                                ctypes.byref(ticket))
    assert rc == _lib.POF_E_UNSUPPORTED
    assert "large-state engine only" in b._lib.pof_last_error(b._ctx).decode()
    assert b.active == 0
    b.close()
