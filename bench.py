"""Benchmark: parallel IEKS ODE solves on one MI355X.

Workload (one "step" = one solve): Lorenz-63 (sigma, rho, beta) = (10, 28, 8/3),
y0 = (1, 1, 1), t in [0, 10], Ns = 65536 uniform steps (len(ts) = 65537),
IWP order q = 3 (state dimension D = 12), float64 throughout, constant
initial trajectory, exactly 10 IEKS iterations per solve.  Inputs are tiny
(y0, ts); everything else lives in HBM.

Also measured (same line): "converged", the same solve run to the
convergence criterion (at most 200 iterations); "accuracy", the error of the
10-iteration solve against the FP64 oracle on the first 512 steps of the
grid; and (objects "batched", "served" and "precision") a
batch of 2048 mixed IVPs solved together vs the plain one-IVP-per-launch
scan; a Poisson request stream served with continuous batching vs a FIFO
plain per-request scan (p50/p99 latency, solves/s); and Lorenz-96 with
d = 32, q = 3 (D = 128) on the large-state engine (ms per iteration at
t in [0, 5] with 8192 steps, errors against the oracle on a short prefix).

Prints one JSON line: solves/s over K timed solves, the roofline of the
dominant kernel class (algorithmic flops per launch over its mean launch
time, measured with HIP events on the launch stream in an instrumented
solve after the timed region), and the CPU baseline: the FP64 NumPy
restatement of the reference algorithm (the oracle) timed on the host on
one full-size IEKS iteration (solves/s = 1 / (10 x that time)).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT / "parallel-in-time-ode-filters_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

METRIC = "ODE solves/s + p50/p99 latency, batched vs plain GPU scan at 1e-6 rel err"
NS = 65536
Q = 3
ITERS = 10
TRAFFIC_SUMMARY = ROOT / "profiles" / "ieks_lorenz63_Ns65536_q3_f64_hbm_traffic.md"


def gpu_run(steps, warmup):
    import ctypes

    import numpy as np
    import torch

    import pof.ivp
    import pof.solver
    from pof import _lib, roofline

    torch.cuda.set_device(0)
    ivp = pof.ivp.lorenz63()
    ts = np.linspace(0.0, 10.0, NS + 1)

    def solve():
        return pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=Q, init="constant", fixed_iters=ITERS)

    for _ in range(warmup):
        solve()
    torch.cuda.synchronize()
    # each solve returns after its last iteration (it reads info on the host),
    # so the per-step wall times sum to the timed region
    step_s = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts0 = time.perf_counter()
        ys, info = solve()
        step_s.append(time.perf_counter() - ts0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    finite = bool(torch.isfinite(ys.mean).all())

    # instrumented solve: per-kernel-class event timing on the launch stream
    lib = _lib.load()
    ctx = _lib.context(0)
    _lib.check(lib.pof_profile_enable(ctx, 1), ctx, "pof_profile_enable")
    solve()
    torch.cuda.synchronize()
    n = len(roofline.KERNEL_CLASSES)
    ms = (ctypes.c_double * n)()
    launches = (ctypes.c_longlong * n)()
    work = (ctypes.c_longlong * n)()
    lib.pof_profile_read(ctx, n, ms, launches, work)
    _lib.check(lib.pof_profile_enable(ctx, 0), ctx, "pof_profile_enable")
    d, D = 3, 3 * (Q + 1)
    per_class = {}
    for i, name in enumerate(roofline.KERNEL_CLASSES):
        if launches[i] == 0:
            continue
        fl = roofline.flops_per_unit(name, D, d) * work[i]
        by = roofline.bytes_per_unit(name, D, d) * work[i]
        per_class[name] = {
            "ms": ms[i],
            "launches": int(launches[i]),
            "mean_launch_us": 1e3 * ms[i] / launches[i],
            "tflops": fl / (ms[i] * 1e-3) / 1e12 if ms[i] > 0 else None,
            "gbs": by / (ms[i] * 1e-3) / 1e9 if ms[i] > 0 else None,
        }
    dom = max(per_class, key=lambda k: per_class[k]["ms"])
    c = per_class[dom]
    flops_per_launch = roofline.flops_per_unit(dom, D, d) * work[roofline.KERNEL_CLASSES.index(dom)] / c["launches"]
    achieved = flops_per_launch / (c["mean_launch_us"] * 1e-6) / 1e12
    total_ms = sum(v["ms"] for v in per_class.values())
    # HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of this workload (null if absent)
    from pof.profiling import read_traffic

    traffic = read_traffic(str(TRAFFIC_SUMMARY), f"k_{dom}<{D}")
    iter_flops = roofline.iteration_flops(NS, D, d)
    iter_bytes = roofline.iteration_bytes(NS, D)
    t_it = float(np.median(step_s)) / ITERS
    t_hbm = iter_bytes / (roofline.PEAK_HBM_GBS * 1e9)
    t_fp64 = iter_flops / (roofline.PEAK_FP64_TFLOPS * 1e12)
    return {
        "seconds": dt,
        "ms_per_iteration_median": 1e3 * t_it,
        "iteration_roofline": {
            "what": "whole IEKS iteration at algorithmic counts: bytes s Ns (3D^2 + 4D), flops Ns (74.3 D^3 + "
                    "4/3 (D + d)^3); time = median over the timed steps / iterations per step",
            "bytes": iter_bytes,
            "flops": iter_flops,
            "ms": 1e3 * t_it,
            "achieved_gbs": iter_bytes / t_it / 1e9,
            "achieved_tflops": iter_flops / t_it / 1e12,
            "hbm_frac": t_hbm / t_it,
            "fp64_frac": t_fp64 / t_it,
            "bound": "fp64" if t_fp64 > t_hbm else "hbm",
        },
        "finite": finite,
        "iterations": info["iterations"],
        "sigma_squared": info["sigma_squared"],
        "per_class": per_class,
        "dominant": dom,
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": roofline.PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / roofline.PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "kernel": dom,
            "flops_per_launch": flops_per_launch,
            "mean_launch_us": c["mean_launch_us"],
            "algorithmic_bytes_per_launch": roofline.bytes_per_unit(dom, D, d)
            * work[roofline.KERNEL_CLASSES.index(dom)] / c["launches"],
            "traffic_source": TRAFFIC_SUMMARY.name if traffic is not None else None,
        },
        "instrumented_iteration_ms": total_ms / ITERS,
        "iteration_tflops_algorithmic": iter_flops / (total_ms / ITERS * 1e-3) / 1e12,
    }


def plain_checks(prefix=512, maxiters=200):
    """The plain workload's other figures: run to the convergence criterion
    (solver.py:36-57; capped at `maxiters`) on the full grid, and the error of
    the 10-iteration solve against the FP64 oracle on the first `prefix` steps
    of the same grid (same dt), where the oracle finishes in a fraction of a second."""
    import numpy as np
    import torch

    import pof.ivp
    import pof.solver
    from oracle import problems as op
    from oracle import solver as osol
    from pof import workloads

    ivp = pof.ivp.lorenz63()
    ts = np.linspace(0.0, 10.0, NS + 1)
    pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=Q, init="constant", maxiters=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=Q, init="constant", maxiters=maxiters)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    conv = {
        "iterations": info["iterations"],
        "status": info.get("status"),
        "seconds": t,
        "ms_per_iteration": 1e3 * t / max(info["iterations"], 1.0),
        "maxiters": maxiters,
        "finite": bool(torch.isfinite(ys.mean).all()),
    }
    tp = ts[: prefix + 1]
    yp, ip = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=tp, order=Q, init="constant", fixed_iters=ITERS)
    torch.cuda.synchronize()
    ref, rinfo = osol.solve(op.lorenz63(), tp, Q, init="constant", fixed_iters=ITERS)
    elem, norm = workloads.max_rel_err(yp.mean.cpu().numpy(), ref.mean)
    seq, _ = osol.solve(op.lorenz63(), tp, Q, init="constant", fixed_iters=ITERS, is_sequential=True)
    elem_s, norm_s = workloads.max_rel_err(yp.mean.cpu().numpy(), seq.mean)
    c = yp.chol.cpu().numpy() / ip["sigma_squared"]
    cr = ref.chol / rinfo["sigma_squared"]
    cc = np.einsum("nij,nkj->nik", c, c)
    ccr = np.einsum("nij,nkj->nik", cr, cr)
    acc = {
        "check": f"first {prefix} steps of the grid (dt = 10/{NS}), {ITERS} fixed iterations, vs the FP64 oracle",
        "max_rel_err_means": elem,
        "normwise_err_means": norm,
        "cov_rel_err": float(np.linalg.norm(cc - ccr) / np.linalg.norm(ccr)),
        "max_rel_err_means_vs_sequential": elem_s,
        "normwise_err_means_vs_sequential": norm_s,
    }
    return conv, acc


def cpu_baseline():
    """The oracle (NumPy FP64 parallel-scan IEKS) on one full-size iteration, one thread."""
    import numpy as np

    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    from oracle import model as om
    from oracle import problems as op
    from oracle import solver as osol

    p = op.lorenz63()
    ts = np.linspace(0.0, 10.0, NS + 1)
    setup = om.set_up_solver(p, ts, Q)
    states = om.get_initial_trajectory(setup, "constant")

    def run():
        t0 = time.perf_counter()
        osol.ieks_step(setup, states)
        return time.perf_counter() - t0

    def runs():
        run()  # warm-up
        return sorted(run() for _ in range(3))

    if threadpool_limits is not None:
        with threadpool_limits(limits=1):
            ts_ = runs()
    else:
        ts_ = runs()
    t = ts_[1]
    return {
        "value": 1.0 / (ITERS * t),
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": f"one IEKS iteration of the Lorenz-63 Ns={NS} q={Q} f64 workload, median of 3 after one warm-up "
                  f"({t:.2f} s; min {ts_[0]:.2f} s), BLAS/LAPACK on 1 thread; solves/s = 1 / ({ITERS} x iteration time)",
        "iteration_s_min": ts_[0],
        "iteration_s_median": t,
        "host_cores": len(os.sched_getaffinity(0)),
    }


def batched_run(n_ivps, maxiters, plain_sample, oracle_sample, seed=4):
    """Mixed-IVP batch (one launch per scan level per (d, q) group) vs the plain per-IVP scan.

    The batch is solved once untimed (allocation, warm-up) and once timed.  The
    plain one-IVP-per-launch time is the sum over a random sample of the IVPs
    solved one at a time (each checked to equal its batch result bitwise),
    scaled to the whole batch.
    """
    import numpy as np
    import torch

    import pof.batch
    import pof.solver
    from pof import workloads

    probs = workloads.mixed_ivps(n_ivps, seed=seed)
    pof.batch.solve_batch(probs, maxiters=maxiters)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = pof.batch.solve_batch(probs, maxiters=maxiters)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    idx = rng.choice(n_ivps, size=min(plain_sample, n_ivps), replace=False)
    t_plain = 0.0
    same = True
    for i in idx:
        f, y0, ts, q, init = probs[i]
        t0 = time.perf_counter()
        ys, _ = pof.solver.solve(f=f, y0=y0, ts=ts, order=q, init=init, maxiters=maxiters)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        same = same and bool(torch.equal(ys.mean, res[i][0].mean)) and bool(torch.equal(ys.chol, res[i][0].chol))
    plain_total = t_plain / len(idx) * n_ivps
    # accuracy against the FP64 oracle on the cheapest converged members
    from oracle import problems as op
    from oracle import solver as osol

    conv = [i for i in range(n_ivps) if res[i][1]["status"] == "converged"]
    conv.sort(key=lambda i: len(probs[i][2]) * (probs[i][0].d * (probs[i][3] + 1)) ** 3)
    errs = []
    for i in conv[:oracle_sample]:
        f, y0, ts, q, init = probs[i]
        (om, _), _ = osol.solve(op.Problem(f.name, f.params, y0, ts[0], ts[-1]), ts, q, init=init,
                                maxiters=maxiters)
        errs.append(workloads.max_rel_err(res[i][0].mean.cpu().numpy(), om))
    status = [r[1]["status"] for r in res]
    its = np.array([r[1]["iterations"] for r in res])
    return {
        "workload": f"{n_ivps} IVPs: logistic / FitzHugh-Nagumo classic / Lorenz-63 with random parameters and y0, "
                    f"Ns log-uniform [256, 16384], q uniform 1..4, f64, maxiters {maxiters}, constant init, seed {seed}",
        "solves_per_s": n_ivps / t_batch,
        "seconds": t_batch,
        "plain_solves_per_s": n_ivps / plain_total,
        "plain_seconds_estimated": plain_total,
        "plain_sample": int(len(idx)),
        "speedup_vs_plain": plain_total / t_batch,
        "same_as_alone": same,
        "iterations_mean": float(its.mean()),
        "status_counts": {k: status.count(k) for k in sorted(set(status))},
        "max_rel_err_vs_oracle": max(e[0] for e in errs) if errs else None,
        "normwise_err_vs_oracle": max(e[1] for e in errs) if errs else None,
        "oracle_sample": len(errs),
    }


def _progress(label, every):
    """on_done hook: one stderr line per `every` completed requests."""

    def hook(rec, n_done):
        if n_done % every == 0:
            print(f"{label}: {n_done} requests done at {rec.done:.1f} s", file=sys.stderr, flush=True)

    return hook


def served_run(n_req, rate, maxiters, seed=5, verbose=False, oracle_sample=2):
    """Poisson request stream served with continuous batching vs a FIFO plain per-request scan."""
    import numpy as np
    import torch

    import pof.serve
    import pof.solver
    from pof import workloads

    reqs = workloads.mixed_ivps(n_req, seed=seed, ns_min=256, ns_max=65536)
    arr = workloads.poisson_arrivals(n_req, rate, seed=seed)
    srv = pof.serve.Server(budget_bytes=96 << 30, maxiters=maxiters)
    srv.run(reqs[: min(24, n_req)], np.zeros(min(24, n_req)))  # warm-up: allocate the groups
    t0 = time.perf_counter()
    recs = srv.run(reqs, arr, on_done=_progress("served", 500) if verbose else None)
    wall = time.perf_counter() - t0
    st = pof.serve.latency_stats(recs)

    def solve_one(r):
        f, y0, ts, q, init = r
        ys, _ = pof.solver.solve(f=f, y0=y0, ts=ts, order=q, init=init, maxiters=maxiters)
        return ys.mean.cpu().numpy()

    t0 = time.perf_counter()
    precs = pof.serve.fifo_plain(reqs, arr, solve_one, on_done=_progress("fifo plain", 500) if verbose else None)
    pwall = time.perf_counter() - t0
    pst = pof.serve.latency_stats(precs)
    same = all(np.array_equal(a.means, b.means) for a, b in zip(recs, precs))
    srv.close()
    # accuracy against the FP64 oracle on the cheapest converged requests
    from oracle import problems as op
    from oracle import solver as osol

    conv = [i for i, r in enumerate(recs) if r.status == "converged"]
    conv.sort(key=lambda i: len(reqs[i][2]) * (reqs[i][0].d * (reqs[i][3] + 1)) ** 3)
    errs = []
    for i in conv[:oracle_sample]:
        f, y0, ts, q, init = reqs[i]
        (om, _), _ = osol.solve(op.Problem(f.name, f.params, y0, ts[0], ts[-1]), ts, q, init=init,
                                maxiters=maxiters)
        errs.append(workloads.max_rel_err(recs[i].means, om))
    return {
        "workload": f"{n_req} requests, Poisson arrivals at {rate:.1f}/s (seed {seed}), same IVP mix with "
                    f"Ns log-uniform [256, 65536], maxiters {maxiters}",
        "solves_per_s": n_req / wall,
        "p50_ms": st["p50_ms"],
        "p99_ms": st["p99_ms"],
        "plain_fifo_solves_per_s": n_req / pwall,
        "plain_fifo_p50_ms": pst["p50_ms"],
        "plain_fifo_p99_ms": pst["p99_ms"],
        "same_as_alone": bool(same),
        "max_rel_err_vs_oracle": max(e[0] for e in errs) if errs else None,
        "normwise_err_vs_oracle": max(e[1] for e in errs) if errs else None,
        "oracle_sample": len(errs),
    }


def precision_run(ns=8192, iters=3, check_ns=32, check_iters=3, fp32_products=True):
    """Lorenz-96 (d = 32, F = 8, y0 = F + 0.01 N(0, 1) from seed 0), t in [0, 5] with 8192 steps, q = 3:
    D = 128 on the large-state engine (batched LQ / GEMM primitives per scan level), float64.

    Timed: one solve of `iters` fixed iterations after an untimed one.  Errors:
    the first `check_ns` steps (same dt) after `check_iters` iterations against
    the oracle: max relative mean error (pof.workloads.max_rel_err), the
    covariance error ||c c^T - c_ref c_ref^T|| / ||c_ref c_ref^T|| after dividing
    out sigma_squared^2 (sign-free; sigma_squared itself depends on factor signs
    for d >= 2), and the relative sigma_squared difference.  The same figures
    with the scan combines' products in FP32 (``products="float32"``) go into
    ``fp32_products``.
    """
    import numpy as np
    import torch

    import pof.ivp
    import pof.solver
    from oracle import problems as op
    from oracle import solver as osol
    from pof import workloads

    ivp = pof.ivp.lorenz96(d=32)
    ts = np.linspace(0.0, 5.0 * ns / 8192, ns + 1)
    tsc = ts[: check_ns + 1]
    ref, rinfo = osol.solve(op.lorenz96(d=32), tsc, 3, init="constant", fixed_iters=check_iters)
    cr = ref.chol / rinfo["sigma_squared"]
    ccr = np.einsum("nij,nkj->nik", cr, cr)

    def measure(products):
        def run(n_it, grid):
            return pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=grid, order=3, init="constant", fixed_iters=n_it,
                                    engine="large", products=products)

        run(1, ts)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ys, _ = run(iters, ts)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        finite = bool(torch.isfinite(ys.mean).all())
        yc, ic = run(check_iters, tsc)
        torch.cuda.synchronize()
        elem, norm = workloads.max_rel_err(yc.mean.cpu().numpy(), ref.mean)
        c = yc.chol.cpu().numpy() / ic["sigma_squared"]
        cc = np.einsum("nij,nkj->nik", c, c)
        cov_err = float(np.linalg.norm(cc - ccr) / np.linalg.norm(ccr))
        ssq_diff = abs(ic["sigma_squared"] - rinfo["sigma_squared"]) / abs(rinfo["sigma_squared"])
        return t, finite, elem, norm, cov_err, ssq_diff

    t, finite, elem, norm, cov_err, ssq_diff = measure("float64")
    try:
        if not fp32_products:
            raise RuntimeError("not run (--no-fp32-products)")
        t32, finite32, elem32, norm32, cov32, ssq32 = measure("float32")
        fp32 = {
            "what": "scan combines' matrix products as FP32 MFMA GEMMs (FP32 accumulate); factorisations, "
                    "triangular solves, elements and observation statistics FP64",
            "ms_per_iteration": 1e3 * t32 / iters,
            "finite": finite32,
            "max_rel_err_means": elem32,
            "normwise_err_means": norm32,
            "cov_rel_err": cov32,
            "sigma_squared_rel_diff": ssq32,
        }
    except RuntimeError as e:  # an optional experiment: its failure must not cost the line
        fp32 = {"error": str(e)}
    return {
        "workload": f"Lorenz-96 d=32 F=8 seed 0, t=[0,{5.0 * ns / 8192:g}] Ns={ns} q=3 (D=128), f64, constant init, "
                    f"{iters} fixed iterations, large-state engine",
        "ms_per_iteration": 1e3 * t / iters,
        "seconds": t,
        "finite": finite,
        "check": f"first {check_ns} steps, {check_iters} fixed iterations, vs the FP64 oracle",
        "max_rel_err_means": elem,
        "normwise_err_means": norm,
        "cov_rel_err": cov_err,
        "sigma_squared_rel_diff": ssq_diff,
        "fp32_products": fp32,
    }


def copy_probe():
    """A plain device copy of known size (2^25 float64 = 256 MiB read + 256 MiB written per launch).

    Run under ``rocprofv3 --pmc FETCH_SIZE`` / ``WRITE_SIZE`` to calibrate those
    counters against a kernel whose HBM traffic is known.
    """
    import torch

    n = 1 << 25
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="only run the HBM counter calibration copy")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched and served measurements")
    ap.add_argument("--no-precision", action="store_true", help="skip the Lorenz-96 D = 128 measurement")
    ap.add_argument("--no-checks", action="store_true",
                    help="skip the run-to-convergence solve and the oracle prefix check of the plain workload")
    ap.add_argument("--precision-only", action="store_true",
                    help="only the Lorenz-96 D = 128 measurement (printed as its own JSON object)")
    ap.add_argument("--precision-ns", type=int, default=8192)
    ap.add_argument("--no-fp32-products", action="store_true",
                    help="skip the Lorenz-96 run with FP32 combine products")
    ap.add_argument("--batch-ivps", type=int, default=2048)
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="requests/s (default: 70%% of the measured batched capacity, scaled to the longer Ns)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.precision_only:
        import torch

        torch.cuda.set_device(0)
        print(json.dumps({"precision": precision_run(args.precision_ns, fp32_products=not args.no_fp32_products)}),
              flush=True)
        return
    g = gpu_run(args.steps, args.warmup)
    if args.verbose:
        print(json.dumps(g["per_class"], indent=1), file=sys.stderr)
    converged = accuracy = None
    if not args.no_checks:
        converged, accuracy = plain_checks()
        if args.verbose:
            print(json.dumps({"converged": converged, "accuracy": accuracy}), file=sys.stderr, flush=True)
    batched = served = None
    if not args.no_batch:
        batched = batched_run(args.batch_ivps, maxiters=100, plain_sample=32, oracle_sample=4)
        if args.verbose:
            print(json.dumps(batched), file=sys.stderr, flush=True)
        # mean Ns of log-uniform [256, 65536] over [256, 16384]: 3.04x more work per request
        rate = args.serve_rate or 0.7 * batched["solves_per_s"] / 3.04
        served = (served_run(args.serve_requests, rate, maxiters=100, verbose=args.verbose)
                  if args.serve_requests > 0 else None)
        if args.verbose:
            print(json.dumps(served), file=sys.stderr, flush=True)
    precision = (None if args.no_precision
                 else precision_run(args.precision_ns, fp32_products=not args.no_fp32_products))
    if args.verbose and precision:
        print(json.dumps(precision), file=sys.stderr, flush=True)
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    ms_step = 1e3 * g["seconds"] / args.steps
    out = {
        "metric": METRIC,
        "value": args.steps / g["seconds"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain parallel IEKS, Lorenz-63 t=[0,10] Ns={NS} q={Q} (D=12), f64, {ITERS} fixed "
                        f"iterations per solve, constant init",
            "Ns": NS,
            "q": Q,
            "iterations_per_solve": ITERS,
            "batch": 1,
        },
        "ms_per_iteration": ms_step / ITERS,
        "ms_per_iteration_median": g["ms_per_iteration_median"],
        "roofline": g["roofline"],
        "iteration_roofline": g["iteration_roofline"],
        "cpu_baseline": cpu,
        "kernel_classes_ms_per_solve": {k: round(v["ms"], 4) for k, v in g["per_class"].items()},
        "finite": g["finite"],
        "converged": converged,
        "accuracy": accuracy,
        "batched": batched,
        "served": served,
        "precision": precision,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
