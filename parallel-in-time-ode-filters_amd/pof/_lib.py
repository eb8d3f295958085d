"""ctypes binding of ``libpof_hip.so`` (declared in ``include/pof.h``).

The shared library is built in-tree by ``pof._build.build()`` and lives next
to this file.  There is no fallback: if the library or a GPU is missing,
every entry point raises.
"""
import ctypes
import os
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libpof_hip.so")

POF_OK = 0
POF_E_ARG = -1
POF_E_UNSUPPORTED = -2
POF_E_HIP = -3
POF_E_INIT = -4
POF_E_FULL = -5

KINDS = {
    "logistic": 1,
    "lotkavolterra": 2,
    "vanderpol": 3,
    "fitzhughnagumo": 4,
    "fhn_classic": 5,
    "rober": 6,
    "rigid_body": 7,
    "seir": 8,
    "threebody": 9,
    "henonheiles": 10,
    "lorenz63": 11,
    "lorenz96": 12,
}
INIT_PRIOR = 0
INIT_CONSTANT = 1
INIT_COARSE = 2
PRODUCTS_F64 = 0
PRODUCTS_F32 = 1
LIN_METHODS = {"ek1": 0, "ek0": 1, "uncertain": 2, "regularized": 3, "unscented": 4}
MAX_PARAMS = 8
STATUS = {0: "converged", 1: "maxiter", 2: "nan", 3: "fixed"}
KERNEL_CLASSES = ["filter_leaf", "filter_up", "filter_down", "filter_post", "smooth_up", "smooth_down",
                  "smooth_post", "reduce"]


class Problem(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("d", ctypes.c_int32),
        ("q", ctypes.c_int32),
        ("init", ctypes.c_int32),
        ("n_times", ctypes.c_int64),
        ("ts_host", ctypes.POINTER(ctypes.c_double)),
        ("y0_host", ctypes.POINTER(ctypes.c_double)),
        ("params", ctypes.c_double * MAX_PARAMS),
    ]


class Opts(ctypes.Structure):
    _fields_ = [
        ("maxiters", ctypes.c_int32),
        ("fixed_iters", ctypes.c_int32),
        ("calibrate", ctypes.c_int32),
        ("sequential", ctypes.c_int32),
        ("products", ctypes.c_int32),
    ]


class Out(ctypes.Structure):
    _fields_ = [
        ("means_dev", ctypes.c_void_p),
        ("chols_dev", ctypes.c_void_p),
        ("full_means_dev", ctypes.c_void_p),
        ("full_chols_dev", ctypes.c_void_p),
        ("iterations", ctypes.c_double),
        ("nll", ctypes.c_double),
        ("obj", ctypes.c_double),
        ("sigma_squared", ctypes.c_double),
        ("calibrated", ctypes.c_int32),
        ("status", ctypes.c_int32),
    ]


_P = ctypes.POINTER
_d = ctypes.c_double
_vp = ctypes.c_void_p
_dp = _P(_d)

# name -> (restype, argtypes); the exported symbol set of include/pof.h
SIGNATURES = {
    "pof_ctx_create": (ctypes.c_int, [ctypes.c_int, _P(_vp)]),
    "pof_ctx_destroy": (None, [_vp]),
    "pof_last_error": (ctypes.c_char_p, [_vp]),
    "pof_supported": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "pof_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pof_profile_read": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _P(ctypes.c_longlong), _P(ctypes.c_longlong)]),
    "pof_graphs_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pof_tuning_set": (ctypes.c_int, [_vp, ctypes.c_int, _P(ctypes.c_int32)]),
    "pof_tuning_get": (ctypes.c_int, [_vp, ctypes.c_int, _P(ctypes.c_int32)]),
    "pof_solve": (ctypes.c_int, [_vp, _P(Problem), _P(Opts), _P(Out), _vp]),
    "pof_solve_large": (ctypes.c_int, [_vp, _P(Problem), _P(Opts), _P(Out), _vp]),
    "pof_setup": (ctypes.c_int, [_P(Problem), _dp, _dp, _dp, _dp, _dp, _dp]),
    "pof_transitions": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _d, _dp, _dp, _dp, _dp]),
    "pof_initial_means": (ctypes.c_int, [_P(Problem), _dp]),
    "pof_coarse_means": (ctypes.c_int, [_vp, _P(Problem), _dp]),
    "pof_eks_solve": (ctypes.c_int, [_vp, _P(Problem), ctypes.c_int, _P(Out), _vp]),
    "pof_eval_field": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
    "pof_linearize": (ctypes.c_int, [_vp, _P(Problem), _vp, ctypes.c_int64, _vp, _vp, _vp]),
    "pof_linearize_model": (ctypes.c_int, [_vp, _P(Problem), ctypes.c_int, _dp, _vp, _vp, ctypes.c_int64, _vp, _vp,
                                           _vp, _vp]),
    "pof_linearize_model_noalloc": (ctypes.c_int, [_vp, _P(Problem), ctypes.c_int, _dp, _vp, _vp, ctypes.c_int64, _vp,
                                                   _vp, _vp, _vp]),
    "pof_ieks_step": (ctypes.c_int, [_vp, _P(Problem), _vp, ctypes.c_int, _vp, _vp, _dp, _dp, _dp, _vp]),
    "pof_linear_filtsmooth": (
        ctypes.c_int,
        [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp, _dp, _vp, _vp, _vp, ctypes.c_int, _vp,
         _vp, _vp, _vp, _dp, _dp, _dp, _vp],
    ),
    "pof_smoothing": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64, _dp, _dp, _vp, _vp, ctypes.c_int, _vp, _vp, _dp,
                                     _vp]),
    "pof_ekf": (ctypes.c_int, [_vp, _P(Problem), _dp, _dp, _dp, _dp, _vp, _vp, _dp, _dp, _vp]),
    "pof_batch_entries_for": (ctypes.c_int64, [ctypes.c_int64]),
    "pof_batch_bytes": (ctypes.c_int64, [ctypes.c_int, ctypes.c_int64, ctypes.c_int32]),
    "pof_batch_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int32, _P(_vp)]),
    "pof_batch_destroy": (None, [_vp]),
    "pof_batch_admit": (ctypes.c_int, [_vp, _P(Problem), _P(Opts), _P(Out), _P(ctypes.c_int64)]),
    "pof_batch_launch": (ctypes.c_int, [_vp]),
    "pof_batch_ready": (ctypes.c_int, [_vp]),
    "pof_batch_finish": (ctypes.c_int, [_vp, ctypes.c_int32, _P(ctypes.c_int64), _P(ctypes.c_int32)]),
    "pof_batch_active": (ctypes.c_int, [_vp]),
    "pof_batch_free_entries": (ctypes.c_int64, [_vp]),
    "pof_solve_batch": (ctypes.c_int, [_vp, _P(Problem), _P(Opts), _P(Out), ctypes.c_int32]),
}

_lock = threading.Lock()
_lib = None
_ctx = {}


def load():
    """Load the shared library (no GPU needed) and bind every exported symbol."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


class PofError(RuntimeError):
    pass


def check(rc, ctx=None, what=""):
    if rc == POF_OK:
        return
    msg = ""
    if ctx is not None:
        m = load().pof_last_error(ctx)
        msg = m.decode() if m else ""
    if rc == POF_E_INIT:
        raise Exception(msg or f"{what}: unknown init method")
    if rc == POF_E_UNSUPPORTED:
        raise NotImplementedError(f"{what}: {msg}")
    if rc == POF_E_ARG:
        raise ValueError(f"{what}: invalid argument {msg}")
    raise PofError(f"{what} failed (status {rc}): {msg}")


def context(device=None):
    """Per-(thread, device) library context on the current HIP device."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("pof needs a ROCm GPU: torch.cuda.is_available() is False")
    if device is None:
        device = torch.cuda.current_device()
    key = (threading.get_ident(), int(device))
    c = _ctx.get(key)
    if c is None:
        lib = load()
        h = ctypes.c_void_p()
        rc = lib.pof_ctx_create(int(device), ctypes.byref(h))
        if rc != POF_OK:
            raise PofError(f"pof_ctx_create(device={device}) failed with status {rc}")
        c = h
        _ctx[key] = c
        # the autotuner's cached kernel-form widths (pof/tuning_gfx950.json), if any
        from . import autotune

        autotune.apply_cache(c)
    return c


def set_graphs(on, device=None):
    """Replay unchanged IEKS iterations from a HIP graph (``on=True``, the default) or launch every kernel."""
    ctx = context(device)
    check(load().pof_graphs_enable(ctx, 1 if on else 0), ctx, "pof_graphs_enable")


def stream_ptr(device=None):
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def dptr(t):
    """Device pointer of a contiguous float64 CUDA tensor (or None)."""
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous() and t.dtype.is_floating_point, "expected a contiguous CUDA tensor"
    return ctypes.c_void_p(t.data_ptr())


def hptr(a):
    """Host pointer of a contiguous float64 numpy array."""
    return a.ctypes.data_as(_dp)


def exported_symbols():
    return sorted(SIGNATURES)
