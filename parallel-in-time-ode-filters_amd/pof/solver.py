"""Public API: the parallel IEKS ODE solve (reference ``pof/solver.py``)."""
import ctypes

from . import _lib
from .convenience import ProblemSpec
from .utils import MVNSqrt


def solve(*, f, y0, ts, order, init="prior", calibrate=True, maxiters=10_000, sequential=False,
          fixed_iters=None, return_full_states=False, engine="auto", products="float64"):
    """Parallel-in-time IEKS solve on the GPU (``pof/solver.py:11-73``).

    Same keyword arguments and result as the reference:
    ``(MVNSqrt(means[len(ts), d], chol_covs[len(ts), d, D]), info)`` with
    ``info = {"iterations", "nll", "obj", "sigma_squared", "calibrated"}``.
    ``f`` must come from ``pof.ivp`` (a device vector field).  Extra:
    ``fixed_iters`` runs exactly that many iterations without the convergence
    test (benchmark mode); ``return_full_states`` also returns the
    preconditioned states in ``info["states"]``.  ``info["status"]`` tells
    how the loop ended.  ``sequential=True`` runs each iteration as the
    sequential filter + smoother (``step.py:39-40``), one GPU worker per IVP.
    ``engine="large"`` forces the large-state engine (batched LQ / GEMM
    primitives per scan level), which ``"auto"`` picks by itself for state
    dimensions beyond the per-worker kernels (Lorenz-96, D = 128).
    ``products="float32"`` (FP32 combine products on the large-state
    engine) is refused with ``NotImplementedError``: its means miss the FP64
    oracle by about 1.3e-6 relative at D = 128 (above the 1e-6 bar), and
    ``sigma_squared``, which scales every returned factor, is rounding noise
    in FP32.
    """
    import torch

    spec = ProblemSpec(f, y0, ts, order, init)
    N, d, D = spec.ts.shape[0], spec.d, spec.D
    dev = torch.device("cuda", torch.cuda.current_device())
    means = torch.empty((N, d), dtype=torch.float64, device=dev)
    chols = torch.empty((N, d, D), dtype=torch.float64, device=dev)
    fm = torch.empty((N, D), dtype=torch.float64, device=dev) if return_full_states else None
    fL = torch.empty((N, D, D), dtype=torch.float64, device=dev) if return_full_states else None
    opts = _lib.Opts()
    opts.maxiters = int(maxiters)
    opts.fixed_iters = int(fixed_iters or 0)
    opts.calibrate = 1 if calibrate else 0
    opts.sequential = 1 if sequential else 0
    if products == "float32":
        raise NotImplementedError(
            "products='float32': FP32 combine products miss the 1e-6 bar on the means (about 1.3e-6 at D = 128) "
            "and do not resolve sigma_squared, which scales the returned factors; use products='float64'")
    if products != "float64":
        raise ValueError(f"products={products!r}: expected 'float64'")
    opts.products = _lib.PRODUCTS_F64
    out = _lib.Out()
    out.means_dev = means.data_ptr()
    out.chols_dev = chols.data_ptr()
    out.full_means_dev = fm.data_ptr() if fm is not None else None
    out.full_chols_dev = fL.data_ptr() if fL is not None else None
    prob = spec.cstruct()
    ctx = _lib.context(dev.index)
    if engine not in ("auto", "large"):
        raise ValueError(f"engine={engine!r}: expected 'auto' or 'large'")
    fn = _lib.load().pof_solve_large if engine == "large" else _lib.load().pof_solve
    _lib.check(fn(ctx, ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(out), _lib.stream_ptr(dev)), ctx,
               "pof_solve")
    info = {
        "iterations": float(out.iterations),
        "nll": float(out.nll),
        "obj": float(out.obj),
        "sigma_squared": float(out.sigma_squared),
        "calibrated": bool(out.calibrated),
        "status": _lib.STATUS.get(out.status, str(out.status)),
    }
    if return_full_states:
        info["states"] = MVNSqrt(fm, fL)
    return MVNSqrt(means, chols), info


def sequential_eks_solve(*, f, y0, ts, order, return_full_states=False, calibrate=True):
    """Single-pass sequential EKS (``pof/solver.py:76-96``) on the GPU.

    Square-root EKF linearised at each predicted mean, then the RTS smoother;
    factors scaled by ``sqrt(sigma_squared)`` when ``calibrate``.  Returns
    ``(MVNSqrt(E0 m, E0 L), info)``, or ``P (m, L)`` (``[len(ts), D]``,
    ``[len(ts), D, D]``) with ``return_full_states``.  ``info["nll"]`` is the
    summed log-likelihood, as in the reference.
    """
    import torch

    spec = ProblemSpec(f, y0, ts, order, "constant")
    N, d, D = spec.ts.shape[0], spec.d, spec.D
    dev = torch.device("cuda", torch.cuda.current_device())
    out = _lib.Out()
    if return_full_states:
        means = torch.empty((N, D), dtype=torch.float64, device=dev)
        chols = torch.empty((N, D, D), dtype=torch.float64, device=dev)
        out.full_means_dev = means.data_ptr()
        out.full_chols_dev = chols.data_ptr()
    else:
        means = torch.empty((N, d), dtype=torch.float64, device=dev)
        chols = torch.empty((N, d, D), dtype=torch.float64, device=dev)
        out.means_dev = means.data_ptr()
        out.chols_dev = chols.data_ptr()
    prob = spec.cstruct()
    ctx = _lib.context(dev.index)
    _lib.check(
        _lib.load().pof_eks_solve(ctx, ctypes.byref(prob), 1 if calibrate else 0, ctypes.byref(out),
                                  _lib.stream_ptr(dev)),
        ctx, "pof_eks_solve",
    )
    info = {
        "nll": float(out.nll),
        "obj": float(out.obj),
        "sigma_squared": float(out.sigma_squared),
        "calibrated": bool(calibrate),
    }
    return MVNSqrt(means, chols), info
