"""The C ABI of libpof_hip.so without a GPU.

Checks that the library loads, that it exports every function declared in
include/pof.h (and that the ctypes binding covers exactly that set), that
its host-side entry points (set-up, transitions, initial trajectories,
vector fields) agree with the oracle, that its error codes follow the
header, and that the Python front end fails loudly instead of falling back
to anything when no GPU is present.  No kernel is launched here.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import model as om
from oracle import problems as op
from pof import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pof.h"

# (D, d) pairs that have kernel instantiations (csrc/pof_inst_d*.hip)
ENGINES = {
    1: [1, 2, 3, 4, 5, 6],
    2: [4, 6, 8, 10, 12],
    3: [6, 9, 12, 15, 18],
    4: [8, 12, 16, 20],
}


@pytest.fixture(scope="module")
def lib():
    if not _lib.LIB_PATH.exists():
        pytest.skip("libpof_hip.so is not built")
    return _lib.load()


def header_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(pof_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_symbols():
    assert header_functions() == _lib.exported_symbols()


def test_library_exports_every_header_symbol(lib):
    for name in header_functions():
        assert hasattr(lib, name), name
    # also visible to a plain dlsym, as a C caller sees it
    raw = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in header_functions():
        assert ctypes.cast(getattr(raw, name), ctypes.c_void_p).value


def test_supported_engine_table(lib):
    for d in range(1, 6):
        for D in range(1, 41):
            expect = int(D in ENGINES.get(d, []))
            assert lib.pof_supported(D, d) == expect, (D, d)


@pytest.mark.parametrize("d,q,dt", [(1, 1, 0.1), (2, 3, 0.01), (3, 3, 10.0 / 65536), (2, 4, 0.5), (4, 2, 1e-3)])
def test_transitions_match_oracle(lib, d, q, dt):
    D = d * (q + 1)
    F, QL, P, PI = (np.empty((D, D)), np.empty((D, D)), np.empty(D), np.empty(D))
    rc = lib.pof_transitions(d, q, dt, _lib.hptr(F), _lib.hptr(QL), _lib.hptr(P), _lib.hptr(PI))
    assert rc == _lib.POF_OK
    iwp = om.IWP(d, q)
    Fo, QLo = om.preconditioned_discretize(iwp)
    Po, PIo = om.nordsieck_preconditioner(iwp, dt)
    np.testing.assert_allclose(F, Fo, rtol=0, atol=1e-15)
    # Cholesky of flip(hilbert(q + 1)) (cond up to 4.8e5 at q = 4): rounding-level agreement
    np.testing.assert_allclose(QL, QLo, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(P, np.diag(Po), rtol=1e-14)
    np.testing.assert_allclose(PI, np.diag(PIo), rtol=1e-14)


def _problem(kind, y0, ts, q, init, params=()):
    p = _lib.Problem()
    p.kind = _lib.KINDS[kind]
    p.d = len(y0)
    p.q = q
    p.init = init
    p.n_times = len(ts)
    p.ts_host = _lib.hptr(ts)
    p.y0_host = _lib.hptr(y0)
    for i, v in enumerate(params):
        p.params[i] = v
    return p


CASES = [
    ("logistic", op.logistic(), 3),
    ("fitzhughnagumo", op.fitzhughnagumo(), 3),
    ("lorenz63", op.lorenz63(), 3),
    ("rigid_body", op.rigid_body(), 2),
    ("vanderpol", op.vanderpol(), 4),
    ("henonheiles", op.henonheiles(), 2),
    ("lotkavolterra", op.lotkavolterra(), 1),
]


@pytest.mark.parametrize("kind,prob,q", CASES, ids=[c[0] for c in CASES])
def test_setup_matches_oracle(lib, kind, prob, q):
    ts = np.linspace(prob.t0, prob.tmax, 40)
    y0 = np.ascontiguousarray(prob.y0)
    p = _problem(kind, y0, ts, q, _lib.INIT_CONSTANT, prob.params)
    D = prob.d * (q + 1)
    F, QL, P, PI, m0, L0 = (np.empty((D, D)), np.empty((D, D)), np.empty(D), np.empty(D), np.empty(D),
                            np.empty((D, D)))
    rc = lib.pof_setup(ctypes.byref(p), *(_lib.hptr(a) for a in (F, QL, P, PI, m0, L0)))
    assert rc == _lib.POF_OK
    s = om.set_up_solver(prob, ts, q)
    np.testing.assert_allclose(F, s.F, atol=1e-15)
    np.testing.assert_allclose(QL, s.QL, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(P, np.diag(s.P), rtol=1e-14)
    np.testing.assert_allclose(m0, s.x0.mean, rtol=1e-12, atol=1e-12 * np.abs(s.x0.mean).max())
    np.testing.assert_array_equal(L0, s.x0.chol)


@pytest.mark.parametrize("kind,prob,q", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("init", ["constant", "prior"])
def test_initial_means_match_oracle(lib, kind, prob, q, init):
    ts = np.linspace(prob.t0, prob.tmax, 25)
    y0 = np.ascontiguousarray(prob.y0)
    code = _lib.INIT_CONSTANT if init == "constant" else _lib.INIT_PRIOR
    p = _problem(kind, y0, ts, q, code, prob.params)
    D = prob.d * (q + 1)
    m = np.empty((len(ts), D))
    assert lib.pof_initial_means(ctypes.byref(p), _lib.hptr(m)) == _lib.POF_OK
    ref = om.get_initial_trajectory(om.set_up_solver(prob, ts, q), init).mean
    np.testing.assert_allclose(m, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("name", sorted(_lib.KINDS))
def test_vector_fields_match_oracle(lib, name):
    prob = getattr(op, name)() if name != "lorenz96" else op.lorenz96(d=8)
    rng = np.random.default_rng(_lib.KINDS[name])
    params = np.zeros(_lib.MAX_PARAMS)
    params[: len(prob.params)] = prob.params
    for _ in range(5):
        y = np.ascontiguousarray(prob.y0 + 0.1 * rng.standard_normal(prob.d) * (1 + np.abs(prob.y0)))
        f = np.empty(prob.d)
        assert lib.pof_eval_field(_lib.KINDS[name], prob.d, _lib.hptr(params), _lib.hptr(y), _lib.hptr(f)) == 0
        ref = prob.f(None, y)
        np.testing.assert_allclose(f, ref, rtol=1e-13, atol=1e-13 * max(1.0, np.abs(ref).max()))


def test_error_codes(lib):
    ts = np.linspace(0.0, 1.0, 5)
    y0 = np.array([0.01])
    m = np.empty((5, 4))
    # unknown field kind
    p = _problem("logistic", y0, ts, 3, _lib.INIT_CONSTANT, (1.0, 1.0))
    p.kind = 99
    assert lib.pof_initial_means(ctypes.byref(p), _lib.hptr(m)) == _lib.POF_E_ARG
    # unknown initial-trajectory method
    p = _problem("logistic", y0, ts, 3, 7, (1.0, 1.0))
    assert lib.pof_initial_means(ctypes.byref(p), _lib.hptr(m)) == _lib.POF_E_INIT
    # fewer than two time points
    p = _problem("logistic", y0, ts, 3, _lib.INIT_CONSTANT, (1.0, 1.0))
    p.n_times = 1
    assert lib.pof_initial_means(ctypes.byref(p), _lib.hptr(m)) == _lib.POF_E_ARG
    # dimension mismatch between kind and d
    f = np.empty(2)
    yy = np.zeros(2)
    par = np.zeros(_lib.MAX_PARAMS)
    assert lib.pof_eval_field(_lib.KINDS["logistic"], 2, _lib.hptr(par), _lib.hptr(yy), _lib.hptr(f)) == _lib.POF_E_ARG
    # negative order or dimension (q = 0 is the reference's degenerate test model and is accepted)
    F = np.empty((1, 1))
    assert lib.pof_transitions(1, -1, 0.1, _lib.hptr(F), _lib.hptr(F), None, None) == _lib.POF_E_ARG
    assert lib.pof_transitions(0, 1, 0.1, _lib.hptr(F), _lib.hptr(F), None, None) == _lib.POF_E_ARG
    assert lib.pof_transitions(1, 0, 0.1, _lib.hptr(F), _lib.hptr(F), None, None) == _lib.POF_OK
    assert F[0, 0] == 1.0
    # no context: error string of a null context is still a valid C string
    assert isinstance(lib.pof_last_error(None), (bytes, type(None)))


def test_python_front_end_fails_loudly_without_gpu(lib):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import pof.ivp
    import pof.solver

    ivp = pof.ivp.logistic()
    with pytest.raises(RuntimeError, match="GPU"):
        pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=np.linspace(0, 1, 5), order=2, init="constant")


def test_non_device_field_is_rejected():
    import pof.solver

    with pytest.raises(TypeError):
        pof.solver.solve(f=lambda t, y: -y, y0=np.ones(1), ts=np.linspace(0, 1, 5), order=2, init="constant")


def test_unknown_init_raises_exception():
    import pof.ivp
    import pof.solver

    ivp = pof.ivp.logistic()
    with pytest.raises(Exception, match="not found"):
        pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=np.linspace(0, 1, 5), order=2, init="nope")
