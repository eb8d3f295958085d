"""Autotuner of the scan kernels' forms (no counterpart in the reference).

Every scan kernel of state dimension D <= 16 (filter up-sweep and down-sweep,
smoother up-sweep and down-sweep) exists in two forms with bitwise equal
results: a latency form, a matrix row split over 2 or 4 lanes and one combine
per wavefront, fastest when a scan level holds too few combines to fill the
chip; and a throughput form, one lane per row and several combines per
wavefront, fewer instructions per combine when a level is wide.  The library
picks the throughput form for a level with at least ``widths[k]`` work items
(summed over a batch's IVPs); ``pof_tuning_set`` sets the four widths per D.

:func:`tune` measures, per (D, d) kernel instantiation, the per-class kernel
time (HIP events around every launch, ``pof_profile_*``) of two synthetic
workloads -- one long IVP (Ns = 2^15, levels 16384 .. 1) and a batch of 96
IVPs of Ns = 2^12 (levels 196608 .. 96) -- for each candidate width, and keeps
for each kernel class the width with the least total time.  The result is
written to ``tuning_gfx950.json`` next to this file; :func:`apply_cache` (run
when a library context is created) installs it.
"""
import ctypes
import json
from pathlib import Path

import numpy as np

from . import _lib

CACHE = Path(__file__).resolve().with_name("tuning_gfx950.json")
NEVER = 1 << 30
CANDIDATES = (NEVER, 65536, 16384, 4096, 1024)
CLASSES = ("filter_up", "filter_down", "smooth_up", "smooth_down")
# (D, d) pairs whose latency form splits rows (D <= 16), with the order q = D / d - 1 <= 4 that a
# device problem of dimension d takes
_FIELDS = {1: "logistic", 2: "fhn_classic", 3: "lorenz63", 4: "lorenz96"}


def set_widths(D, widths, ctx=None):
    """Level widths from which the four scan kernels of dimension D run in their throughput form."""
    ctx = ctx if ctx is not None else _lib.context()
    arr = (ctypes.c_int32 * 4)(*[int(w) for w in widths])
    _lib.check(_lib.load().pof_tuning_set(ctx, int(D), arr), ctx, "pof_tuning_set")


def get_widths(D, ctx=None):
    ctx = ctx if ctx is not None else _lib.context()
    arr = (ctypes.c_int32 * 4)()
    _lib.check(_lib.load().pof_tuning_get(ctx, int(D), arr), ctx, "pof_tuning_get")
    return list(arr)


def read_cache(path=CACHE):
    """{D: [4 widths]} from the cache file, or {} without one."""
    p = Path(path)
    if not p.exists():
        return {}
    data = json.loads(p.read_text())
    return {int(D): [int(w) for w in v] for D, v in data.get("widths", {}).items()}


def apply_cache(ctx, path=CACHE):
    """Install the cached widths on a library context (called when it is created)."""
    lib = _lib.load()
    for D, widths in read_cache(path).items():
        arr = (ctypes.c_int32 * 4)(*widths)
        _lib.check(lib.pof_tuning_set(ctx, D, arr), ctx, "pof_tuning_set")


def _problems(D, d):
    from . import ivp as _ivp

    q = D // d - 1
    name = _FIELDS[d]
    if name == "lorenz96":
        f = _ivp.lorenz96(d=4)
    else:
        f = getattr(_ivp, name)()
    long = [(f.f, f.y0, np.linspace(f.t0, f.tmax, (1 << 15) + 1), q, "constant")]
    batch = [(f.f, f.y0 * (1.0 + 1e-3 * i), np.linspace(f.t0, f.tmax, (1 << 12) + 1), q, "constant")
             for i in range(96)]
    return long, batch


def _class_ms(ctx, run):
    import torch

    from . import roofline

    lib = _lib.load()
    n = len(roofline.KERNEL_CLASSES)
    ms = (ctypes.c_double * n)()
    la = (ctypes.c_longlong * n)()
    wk = (ctypes.c_longlong * n)()
    _lib.check(lib.pof_profile_enable(ctx, 1), ctx, "pof_profile_enable")
    run()
    torch.cuda.synchronize()
    lib.pof_profile_read(ctx, n, ms, la, wk)
    _lib.check(lib.pof_profile_enable(ctx, 0), ctx, "pof_profile_enable")
    return {k: ms[roofline.KERNEL_CLASSES.index(k)] for k in CLASSES}


def tune(pairs=None, candidates=CANDIDATES, iters=2, write=True, path=CACHE, log=None):
    """Measure the candidate widths per (D, d) and kernel class; returns (and writes) the cache.

    ``pairs``: (D, d) kernel instantiations to tune (default: every pair with D <= 16 and
    q = D / d - 1 in 1..4 that has a device problem of dimension d).
    """
    import torch

    from . import batch as _batch
    from . import solver as _solver

    ctx = _lib.context()
    if pairs is None:
        pairs = [(d * (q + 1), d) for d in (1, 2, 3, 4) for q in (1, 2, 3, 4) if d * (q + 1) <= 16]
    # one width per D, over every d of that D: measurements add up per D
    totals = {}
    for D, d in pairs:
        if not _lib.load().pof_supported(D, d):
            continue
        long, batch = _problems(D, d)

        def run():
            for f, y0, ts, q, init in long:
                _solver.solve(f=f, y0=y0, ts=ts, order=q, init=init, fixed_iters=iters)
            _batch.solve_batch(batch, fixed_iters=iters)

        run()  # warm-up: allocation of the batches
        torch.cuda.synchronize()
        for w in candidates:
            set_widths(D, [w] * 4, ctx)
            ms = _class_ms(ctx, run)
            for k in CLASSES:
                totals.setdefault(D, {}).setdefault(k, {}).setdefault(w, 0.0)
                totals[D][k][w] += ms[k]
            if log:
                log(f"D={D} d={d} width={w}: " + ", ".join(f"{k} {ms[k]:.3f} ms" for k in CLASSES))
    widths = {}
    for D, per in totals.items():
        widths[D] = [min(per[k], key=lambda w: per[k][w]) for k in CLASSES]
        set_widths(D, widths[D], ctx)
    cache = {
        "what": "scan-level widths from which each scan kernel runs in its throughput form, per state dimension "
                "D: [filter_up, filter_down, smooth_up, smooth_down]; 1073741824 = never",
        "device": torch.cuda.get_device_name(0),
        "workloads": "per (D, d): one IVP of Ns = 2^15 and a batch of 96 IVPs of Ns = 2^12, "
                     f"{iters} fixed iterations each, kernel time per class from HIP events",
        "candidates": list(candidates),
        "widths": {str(D): widths[D] for D in sorted(widths)},
        "measured_ms": {str(D): {k: {str(w): round(v, 4) for w, v in per[k].items()} for k in CLASSES}
                        for D, per in sorted(totals.items())},
    }
    if write:
        Path(path).write_text(json.dumps(cache, indent=1) + "\n")
    return cache


if __name__ == "__main__":
    import sys

    c = tune(log=lambda s: print(s, file=sys.stderr, flush=True))
    print(json.dumps(c["widths"]))
