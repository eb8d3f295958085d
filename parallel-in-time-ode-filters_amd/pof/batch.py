"""Many IVPs per launch: batched solves and a continuously batched group.

Not in the reference, which solves one IVP per ``pof.solver.solve`` call.
Every IVP keeps the arithmetic of a single solve (same kernels, same scan
order, its own reductions and convergence test), so its result is identical
whether it is solved alone or in a batch.

* :func:`solve_batch` solves a list of IVPs of any mix of ``(d, order)``
  (one device group per state dimension, groups iterated concurrently).
* :class:`Batch` is one group (fixed ``d`` and ``order``) whose IVPs can be
  admitted and retired between iterations: the building block of
  :mod:`pof.serve`.
"""
import ctypes
from typing import NamedTuple

import numpy as np

from . import _lib
from .convenience import ProblemSpec
from .utils import MVNSqrt


class Problem(NamedTuple):
    """One IVP of a batch: the arguments of ``pof.solver.solve``."""

    f: object
    y0: object
    ts: object
    order: int
    init: str = "prior"


def _opts(calibrate, maxiters, fixed_iters):
    o = _lib.Opts()
    o.maxiters = int(maxiters)
    o.fixed_iters = int(fixed_iters or 0)
    o.calibrate = 1 if calibrate else 0
    return o


def _outputs(spec, dev, full):
    import torch

    N, d, D = spec.ts.shape[0], spec.d, spec.D
    means = torch.empty((N, d), dtype=torch.float64, device=dev)
    chols = torch.empty((N, d, D), dtype=torch.float64, device=dev)
    fm = torch.empty((N, D), dtype=torch.float64, device=dev) if full else None
    fL = torch.empty((N, D, D), dtype=torch.float64, device=dev) if full else None
    return means, chols, fm, fL


def _fill_out(out, tensors):
    means, chols, fm, fL = tensors
    out.means_dev = means.data_ptr()
    out.chols_dev = chols.data_ptr()
    out.full_means_dev = fm.data_ptr() if fm is not None else None
    out.full_chols_dev = fL.data_ptr() if fL is not None else None


def _result(out, tensors):
    means, chols, fm, fL = tensors
    info = {
        "iterations": float(out.iterations),
        "nll": float(out.nll),
        "obj": float(out.obj),
        "sigma_squared": float(out.sigma_squared),
        "calibrated": bool(out.calibrated),
        "status": _lib.STATUS.get(out.status, str(out.status)),
    }
    if fm is not None:
        info["states"] = MVNSqrt(fm, fL)
    return MVNSqrt(means, chols), info


def solve_batch(problems, *, calibrate=True, maxiters=10_000, fixed_iters=None, return_full_states=False):
    """Solve many IVPs together; returns ``[(MVNSqrt(means, chol_covs), info), ...]`` in input order.

    Each entry of ``problems`` is a :class:`Problem` (or a tuple / dict with
    the same fields).  Results equal ``pof.solver.solve`` on each IVP alone.
    """
    import torch

    specs = []
    for p in problems:
        if isinstance(p, dict):
            p = Problem(**p)
        elif not isinstance(p, Problem):
            p = Problem(*p)
        specs.append(ProblemSpec(p.f, p.y0, p.ts, p.order, p.init))
    n = len(specs)
    if n == 0:
        return []
    dev = torch.device("cuda", torch.cuda.current_device())
    probs = (_lib.Problem * n)()
    opts = (_lib.Opts * n)()
    outs = (_lib.Out * n)()
    tensors = []
    for i, s in enumerate(specs):
        probs[i] = s.cstruct()
        opts[i] = _opts(calibrate, maxiters, fixed_iters)
        t = _outputs(s, dev, return_full_states)
        _fill_out(outs[i], t)
        tensors.append(t)
    torch.cuda.synchronize(dev)  # output buffers are written on the library's streams
    ctx = _lib.context(dev.index)
    _lib.check(_lib.load().pof_solve_batch(ctx, probs, opts, outs, n), ctx, "pof_solve_batch")
    return [_result(outs[i], tensors[i]) for i in range(n)]


class BatchFull(Exception):
    """The group has no free slot or workspace range for this IVP right now."""


class Batch:
    """One continuously batched group of IVPs with state dimension ``D = d (order + 1)``.

    ``entries`` is the workspace size (an IVP with ``len(ts)`` points needs
    ``2 len(ts)``; see :func:`entries_for`), ``slots`` the maximum number of
    IVPs in flight.  Usage::

        b = Batch(d=3, order=3, entries=..., slots=...)
        t = b.admit(f, y0, ts)            # any time between steps
        done = b.step()                   # one IEKS iteration for every admitted IVP
        for ticket, (ys, info) in done: ...
    """

    def __init__(self, d, order, entries, slots, device=None):
        import torch

        self.d = int(d)
        self.order = int(order)
        self.D = self.d * (self.order + 1)
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self._lib = _lib.load()
        self._ctx = _lib.context(self.device.index)
        h = ctypes.c_void_p()
        _lib.check(self._lib.pof_batch_create(self._ctx, self.D, self.d, int(entries), int(slots), ctypes.byref(h)),
                   self._ctx, "pof_batch_create")
        self._h = h
        self._pending = {}
        self._launched = False

    def close(self):
        if self._h is not None:
            self._lib.pof_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover - interpreter shutdown
            pass

    @property
    def active(self):
        return int(self._lib.pof_batch_active(self._h))

    @property
    def free_entries(self):
        return int(self._lib.pof_batch_free_entries(self._h))

    def admit(self, f, y0, ts, init="prior", *, calibrate=True, maxiters=10_000, fixed_iters=None,
              return_full_states=False):
        """Queue one IVP for the next iteration; returns its ticket.  Raises :class:`BatchFull`."""
        spec = ProblemSpec(f, y0, ts, self.order, init)
        if spec.D != self.D:
            raise ValueError(f"IVP has D={spec.D}, the batch D={self.D}")
        prob = spec.cstruct()
        opts = _opts(calibrate, maxiters, fixed_iters)
        out = _lib.Out()
        tensors = _outputs(spec, self.device, return_full_states)
        _fill_out(out, tensors)
        import torch

        # the output buffers come from torch's allocator on torch's stream; the batch writes them
        # on its own stream (waiting for the whole device would stall the other groups)
        torch.cuda.current_stream(self.device).synchronize()
        ticket = ctypes.c_int64()
        rc = self._lib.pof_batch_admit(self._h, ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(out),
                                       ctypes.byref(ticket))
        if rc == _lib.POF_E_FULL:
            raise BatchFull(self._lib.pof_last_error(self._ctx).decode())
        _lib.check(rc, self._ctx, "pof_batch_admit")
        # keep the host inputs and the output structs alive until the IVP finishes
        self._pending[ticket.value] = (spec, out, tensors)
        return ticket.value

    def launch(self):
        """Enqueue one iteration for every admitted, unfinished IVP (returns at once)."""
        _lib.check(self._lib.pof_batch_launch(self._h), self._ctx, "pof_batch_launch")
        self._launched = True

    @property
    def launched(self):
        return self._launched

    def ready(self):
        """True if :meth:`finish` would not wait (the launched iteration has completed)."""
        rc = self._lib.pof_batch_ready(self._h)
        if rc < 0:
            _lib.check(rc, self._ctx, "pof_batch_ready")
        return rc == 1

    def finish(self):
        """Wait for the launched iteration; returns ``[(ticket, (MVNSqrt, info)), ...]`` of finished IVPs."""
        if not self._launched:
            return []
        self._launched = False
        cap = max(len(self._pending), 1)
        done = (ctypes.c_int64 * cap)()
        nd = ctypes.c_int32()
        _lib.check(self._lib.pof_batch_finish(self._h, cap, done, ctypes.byref(nd)), self._ctx, "pof_batch_finish")
        res = []
        for i in range(min(nd.value, cap)):
            spec, out, tensors = self._pending.pop(done[i])
            res.append((done[i], _result(out, tensors)))
        return res

    def step(self):
        """launch() + finish()."""
        self.launch()
        return self.finish()


def entries_for(n_times):
    """Workspace entries an IVP with ``n_times`` time points occupies."""
    return 2 * int(n_times)


def bytes_for(D, entries, slots):
    """Device bytes of a :class:`Batch`."""
    return int(_lib.load().pof_batch_bytes(int(D), int(entries), int(slots)))
