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
one full-size IEKS iteration in each of as many processes as the host has
cores (solves/s = processes / (10 x the mean iteration time)).
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
            "peak_note": "MI355X's FP64 peak (78.6 TFLOP/s) is the same on the matrix cores and the vector ALUs; "
                         "the D = 12 combines form their tile products on the matrix cores and the Householder "
                         "pivots on the vector ALUs",
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
    # the configuration is unstable for the parallel method itself: after two iterations the parallel
    # iterates are far from the sequential ones (the same linear model, another association order)
    y2, _ = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=Q, init="constant", fixed_iters=2)
    s2, _ = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=Q, init="constant", fixed_iters=2, sequential=True)
    torch.cuda.synchronize()
    conv["stability"] = {
        "what": "max |mean| after 2 fixed iterations on the full grid: parallel scan vs the sequential filter and "
                "smoother (the same linear models; they differ only by rounding and its amplification)",
        "parallel_max_abs_mean": float(y2.mean.abs().max()),
        "sequential_max_abs_mean": float(s2.mean.abs().max()),
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


def known_answers():
    """The reference's own full-size run at the benchmark's state dimension and length: rigid body
    (d = 3), N = 65536 (65535 steps), q = 1 and 3 (D = 6 and 12), constant init, maxiters = 1000,
    against experiments/3_work_precision_diagram/data/rigidbody_Tesla_V100-SXM2-32GB.csv (FP64 on a
    V100): iteration counts, and rmse_final / rmse_traj against a DOP853 reference solution
    (rtol 1e-13).  For q = 3 both rmse sit at the reference solution's error floor (~1e-11)."""
    import csv

    import numpy as np
    import torch
    from scipy.integrate import solve_ivp

    import pof.ivp
    import pof.solver
    from oracle import problems as op

    rows = {}
    with open(ROOT / "tests" / "golden" / "work_precision_pins_large.csv") as f:
        for r in csv.DictReader(f):
            if r["problem"] == "rigidbody" and r["N"] == "65536":
                rows[int(r["q"])] = r
    ivp = pof.ivp.rigid_body()
    p = op.rigid_body()
    ts = np.linspace(ivp.t0, ivp.tmax, 65536)
    sol = solve_ivp(lambda _, y: p.f(None, y), p.t_span, p.y0, method="DOP853", rtol=1e-13, atol=1e-15,
                    dense_output=True)
    yref = sol.sol(ts).T
    out = {"workload": "rigid body, t in [0, 20], N = 65536, constant init, maxiters 1000, f64",
           "source": "rigidbody_Tesla_V100-SXM2-32GB.csv:11 (reference, FP64, V100)"}
    for q in (1, 3):
        t0 = time.perf_counter()
        ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init="constant", maxiters=1000)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        m = ys.mean.cpu().numpy()
        r = rows[q]
        out[f"q{q}"] = {
            "D": 3 * (q + 1),
            "iterations": info["iterations"],
            "iterations_reference": float(r["iterations"]),
            "rmse_final": float(np.linalg.norm(m[-1] - yref[-1])),
            "rmse_final_reference": float(r["rmse_final"]),
            "rmse_traj": float(np.linalg.norm(m - yref, axis=1).mean()),
            "rmse_traj_reference": float(r["rmse_traj"]),
            "seconds": dt,
            "reference_seconds_v100": float(r["runtime"]),
        }
    return out


def small_state_run(ns=65536, iters=10):
    """Plain parallel IEKS at the smallest state of the north-star HBM target: logistic q = 3
    (d = 1, D = 4), Ns = 2^16, 10 fixed iterations.  Reports ms per iteration and the
    element-once HBM fraction (bytes s Ns (3 D^2 + 4 D) per iteration at 8 TB/s), and the
    per-class kernel times of one instrumented solve."""
    import ctypes

    import numpy as np
    import torch

    import pof.ivp
    import pof.solver
    from pof import _lib, roofline

    ivp = pof.ivp.logistic()
    ts = np.linspace(0.0, 10.0, ns + 1)

    def solve():
        return pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=3, init="constant", fixed_iters=iters)

    solve()
    torch.cuda.synchronize()
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        solve()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    t_it = float(np.median(times)) / iters
    D, d = 4, 1
    by = roofline.iteration_bytes(ns, D)
    fl = roofline.iteration_flops(ns, D, d)
    lib = _lib.load()
    ctx = _lib.context(0)
    n = len(roofline.KERNEL_CLASSES)
    ms = (ctypes.c_double * n)()
    la = (ctypes.c_longlong * n)()
    wk = (ctypes.c_longlong * n)()
    lib.pof_profile_enable(ctx, 1)
    solve()
    torch.cuda.synchronize()
    lib.pof_profile_read(ctx, n, ms, la, wk)
    lib.pof_profile_enable(ctx, 0)
    return {
        "workload": f"plain parallel IEKS, logistic t=[0,10] Ns={ns} q=3 (D=4), f64, {iters} fixed iterations",
        "ms_per_iteration": 1e3 * t_it,
        "element_once_bytes": by,
        "achieved_gbs": by / t_it / 1e9,
        "hbm_frac": by / t_it / (roofline.PEAK_HBM_GBS * 1e9),
        "fp64_frac": fl / t_it / (roofline.PEAK_FP64_TFLOPS * 1e12),
        "kernel_classes_ms_per_iteration": {k: round(ms[i] / iters, 4) for i, k in enumerate(roofline.KERNEL_CLASSES)
                                            if la[i] > 0},
        "launches_per_iteration": sum(la) / iters,
    }


def _lorenz63_iteration_job(_):
    """One full-size oracle IEKS iteration of the Lorenz-63 workload in a worker process (seconds)."""
    import numpy as np

    from oracle import model as om
    from oracle import problems as op
    from oracle import solver as osol

    setup = om.set_up_solver(op.lorenz63(), np.linspace(0.0, 10.0, NS + 1), Q)
    states = om.get_initial_trajectory(setup, "constant")
    t0 = time.perf_counter()
    osol.ieks_step(setup, states)
    return time.perf_counter() - t0


def cpu_baseline():
    """The oracle (NumPy FP64 parallel-scan IEKS) on the host: one full-size iteration of the
    workload in each of as many processes as there are host cores, all at once (BLAS on one
    thread each), so solves/s = cores / (10 x the mean iteration time) is the host's throughput on
    independent solves.  The single-process time of one iteration is reported beside it."""
    import multiprocessing as mp

    import numpy as np

    cores = len(os.sched_getaffinity(0))
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in env:
        os.environ[k] = "1"
    try:
        ctx = mp.get_context("spawn")
        with ctx.Pool(1) as pool:
            single = pool.map(_lorenz63_iteration_job, [0])[0]
        t0 = time.perf_counter()
        with ctx.Pool(cores) as pool:
            per = pool.map(_lorenz63_iteration_job, range(cores), chunksize=1)
        wall = time.perf_counter() - t0
    finally:
        for k, v in env.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    t = float(np.mean(per))
    return {
        "value": cores / (ITERS * t),
        "unit": "solves/s",
        "cores": cores,
        "kind": "port",
        "sample": f"one IEKS iteration of the Lorenz-63 Ns={NS} q={Q} f64 workload in each of {cores} processes at "
                  f"once (BLAS on 1 thread each; mean {t:.2f} s per iteration, {wall:.1f} s wall); solves/s = "
                  f"{cores} / ({ITERS} x mean iteration time)",
        "iteration_s_mean": t,
        "iteration_s_single_process": single,
        "single_core_solves_per_s": 1.0 / (ITERS * single),
    }


def batched_run(n_ivps, maxiters, plain_sample, seed=4, cpu=True):
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
    # the same batch with every scan level in the latency form (the kernels' forms before tuning)
    from pof import autotune

    tuned = autotune.read_cache()
    t_default = None
    if tuned:
        for D in tuned:
            autotune.set_widths(D, [autotune.NEVER] * 4)
        t0 = time.perf_counter()
        res_d = pof.batch.solve_batch(probs, maxiters=maxiters)
        torch.cuda.synchronize()
        t_default = time.perf_counter() - t0
        same_forms = all(bool(torch.equal(a[0].mean, b[0].mean)) for a, b in zip(res, res_d))
        for D, w in tuned.items():
            autotune.set_widths(D, w)
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
    oracle = stratified_oracle_check(probs, [r[1] for r in res], [r[0].mean for r in res], maxiters)
    status = [r[1]["status"] for r in res]
    its = np.array([r[1]["iterations"] for r in res])
    return {
        "workload": f"{n_ivps} IVPs: logistic / FitzHugh-Nagumo classic / Lorenz-63 with random parameters and y0, "
                    f"Ns log-uniform [256, 16384], q uniform 1..4, f64, maxiters {maxiters}, constant init, seed {seed}",
        "solves_per_s": n_ivps / t_batch,
        "seconds": t_batch,
        "autotuned": {
            "what": "scan kernel forms from the autotuner's cache (pof/tuning_gfx950.json) against every level in "
                    "the latency form",
            "tuned_seconds": t_batch,
            "default_seconds": t_default,
            "speedup": (t_default / t_batch) if t_default else None,
            "results_bitwise_equal": same_forms if tuned else None,
        },
        "plain_solves_per_s": n_ivps / plain_total,
        "plain_seconds_estimated": plain_total,
        "plain_sample": int(len(idx)),
        "speedup_vs_plain": plain_total / t_batch,
        "same_as_alone": same,
        "iterations_mean": float(its.mean()),
        "status_counts": {k: status.count(k) for k in sorted(set(status))},
        "oracle_check": oracle,
        "max_normwise_err_vs_oracle": oracle["max_normwise_err_well_conditioned"],
        "algorithmic_tflops": batch_flops(probs, [r[1]["iterations"] for r in res]) / t_batch / 1e12,
        "cpu_baseline": batch_cpu_baseline(probs, [r[1]["iterations"] for r in res]) if cpu else None,
    }


def batch_flops(probs, iterations):
    """Algorithmic flops of a batch: sum over IVPs of iterations x Ns (74.3 D^3 + 4/3 (D + d)^3)."""
    from pof import roofline

    tot = 0.0
    for (f, _, ts, q, _), it in zip(probs, iterations):
        tot += it * roofline.iteration_flops(len(ts) - 1, f.d * (q + 1), f.d)
    return tot


def stratified_oracle_check(probs, infos, means, maxiters, prefix=512, prefix_iters=3):
    """Compare batch members with the FP64 oracle, one per (d, q) group and status.

    For every (d, q) group (D = d (q + 1) up to 15) it takes the shortest member of each final
    status (converged, maxiter), so the maxiter members and the q = 4 Lorenz-63 (D = 15)
    members are covered.  A converged member is compared as solved (iterations exact, means).
    A maxiter member (not converged; the oracle's run to maxiters would take minutes) is
    compared on a prefix of its grid: the first `prefix` steps, `prefix_iters` fixed iterations,
    solved on the GPU and by the oracle alike.

    Each comparison also runs the oracle once more with y0 scaled by 1 + 1e-15 and reports that
    spread (normwise): where the configuration amplifies rounding (an unstable parallel scan, an
    iteration path that wanders before converging) the oracle moves by as much as the GPU differs
    from it, and the GPU's error there measures the conditioning, not the implementation.  The
    members whose own spread is below 1e-8 are the well-conditioned ones; the bar (1e-6) applies to
    them.
    """
    import numpy as np
    import torch

    import pof.solver
    from oracle import problems as op
    from oracle import solver as osol
    from pof import workloads

    pick = {}
    for i, ((f, _, ts, q, _), info) in enumerate(zip(probs, infos)):
        if info["status"] not in ("converged", "maxiter"):
            continue
        key = (f.name, f.d, q, info["status"])
        if key not in pick or len(ts) < len(probs[pick[key]][2]):
            pick[key] = i
    rows = []
    worst = 0.0
    for key in sorted(pick):
        i = pick[key]
        f, y0, ts, q, init = probs[i]
        oprob = op.Problem(f.name, f.params, y0, ts[0], ts[-1])
        y0p = np.asarray(y0, dtype=np.float64) * (1.0 + 1e-15)
        if key[3] == "converged":
            (om, _), oinfo = osol.solve(oprob, ts, q, init=init, maxiters=maxiters)
            (op_, _), pinfo = osol.solve(op.Problem(f.name, f.params, y0p, ts[0], ts[-1]), ts, q, init=init,
                                         maxiters=maxiters)
            m = means[i].cpu().numpy()
            what = f"converged, {int(infos[i]['iterations'])} iterations (oracle {int(oinfo['iterations'])}, " \
                   f"perturbed oracle {int(pinfo['iterations'])})"
            same_it = oinfo["iterations"] == infos[i]["iterations"]
        else:
            tp = ts[: prefix + 1]
            yp, _ = pof.solver.solve(f=f, y0=y0, ts=tp, order=q, init=init, fixed_iters=prefix_iters)
            torch.cuda.synchronize()
            (om, _), _ = osol.solve(op.Problem(f.name, f.params, y0, tp[0], tp[-1]), tp, q, init=init,
                                    fixed_iters=prefix_iters)
            (op_, _), _ = osol.solve(op.Problem(f.name, f.params, y0p, tp[0], tp[-1]), tp, q, init=init,
                                     fixed_iters=prefix_iters)
            m = yp.mean.cpu().numpy()
            what = f"{key[3]}: first {len(tp) - 1} steps, {prefix_iters} fixed iterations"
            same_it = True
        elem, norm = workloads.max_rel_err(m, om)
        _, spread = workloads.max_rel_err(op_, om)
        well = spread < 1e-8
        if well:
            worst = max(worst, norm)
        rows.append({"problem": f.name, "d": f.d, "q": q, "D": f.d * (q + 1), "Ns": len(ts) - 1, "check": what,
                     "iterations_equal": bool(same_it), "max_rel_err": elem, "normwise_err": norm,
                     "oracle_spread_normwise": spread, "well_conditioned": bool(well)})
    wc = [r for r in rows if r["well_conditioned"]]
    return {"members": rows, "groups": len({(r["d"], r["q"]) for r in rows}),
            "well_conditioned": len(wc),
            "max_normwise_err_well_conditioned": worst,
            "max_rel_err_well_conditioned": max((r["max_rel_err"] for r in wc), default=None),
            "all_iterations_equal_well_conditioned": all(r["iterations_equal"] for r in wc),
            "ill_conditioned": [f"{r['problem']} q={r['q']} Ns={r['Ns']}: GPU {r['normwise_err']:.1e} vs oracle "
                                f"spread {r['oracle_spread_normwise']:.1e}" for r in rows if not r["well_conditioned"]]}


def _oracle_solve_job(args):
    """Oracle IEKS iterations of one IVP in a worker process (the CPU baseline of the batch):
    the seconds per iteration, over `iters` fixed iterations after the set-up."""
    name, params, y0, ts, q, iters = args
    from oracle import model as om
    from oracle import problems as op
    from oracle import solver as osol

    setup = om.set_up_solver(op.Problem(name, params, y0, ts[0], ts[-1]), ts, q)
    states = om.get_initial_trajectory(setup, "constant")
    t0 = time.perf_counter()
    for _ in range(iters):
        states = osol.ieks_step(setup, states)[0]
    return (time.perf_counter() - t0) / iters


def batch_cpu_baseline(probs, iterations, sample=48, iters=2, seed=1):
    """The oracle on a random sample of the batch's IVPs, one process per host core (BLAS on one
    thread each): each sampled IVP runs `iters` IEKS iterations and its solve time is that time per
    iteration times the iteration count its GPU solve took; solves/s = sample / (summed solve times
    / cores), the pool's throughput at full occupancy."""
    import multiprocessing as mp

    import numpy as np

    rng = np.random.default_rng(seed)
    idx = rng.choice(len(probs), size=min(sample, len(probs)), replace=False)
    jobs = []
    for i in idx:
        f, y0, ts, q, _ = probs[i]
        jobs.append((f.name, tuple(f.params), np.asarray(y0, dtype=np.float64), ts, q, iters))
    cores = len(os.sched_getaffinity(0))
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in env:
        os.environ[k] = "1"
    try:
        ctx = mp.get_context("spawn")
        t0 = time.perf_counter()
        with ctx.Pool(cores) as pool:
            per = pool.map(_oracle_solve_job, jobs, chunksize=1)
        wall = time.perf_counter() - t0
    finally:
        for k, v in env.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    solve_s = sum(t * max(float(iterations[i]), 1.0) for t, i in zip(per, idx))
    return {
        "value": len(jobs) / (solve_s / cores),
        "unit": "solves/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{len(jobs)} IVPs drawn at random (seed {seed}) from the batch; each ran {iters} oracle IEKS "
                  f"iterations in a pool of {cores} processes (BLAS on 1 thread each, {wall:.1f} s wall), its "
                  f"solve time taken as seconds per iteration x the iteration count of its GPU solve; "
                  f"{solve_s:.0f} s of CPU work for the {len(jobs)} solves",
    }


def _progress(label, every):
    """on_done hook: one stderr line per `every` completed requests."""

    def hook(rec, n_done):
        if n_done % every == 0:
            print(f"{label}: {n_done} requests done at {rec.done:.1f} s", file=sys.stderr, flush=True)

    return hook


def served_run(n_req, rate, maxiters, seed=5, verbose=False):
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
    import torch as _t

    oracle = stratified_oracle_check(reqs, [{"status": r.status, "iterations": r.iterations} for r in recs],
                                     [_t.as_tensor(r.means) for r in recs], maxiters)
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
        "oracle_check": oracle,
        "max_normwise_err_vs_oracle": oracle["max_normwise_err_well_conditioned"],
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
            if products == "float64":
                return pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=grid, order=3, init="constant", fixed_iters=n_it,
                                        engine="large")
            return fp32_solve(ivp, grid, n_it)

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

    def fp32_solve(ivp_, grid, n_it):
        """The FP32-product experiment through the C ABI (the Python API refuses it): calibrate = 0,
        means only (sigma_squared, and with it every factor, is rounding noise in FP32)."""
        import ctypes

        from pof import _lib
        from pof.convenience import ProblemSpec
        from pof.utils import MVNSqrt

        spec = ProblemSpec(ivp_.f, ivp_.y0, grid, 3, "constant")
        prob = spec.cstruct()
        means = torch.empty((len(grid), spec.d), dtype=torch.float64, device="cuda")
        chols = torch.empty((len(grid), spec.d, spec.D), dtype=torch.float64, device="cuda")
        out = _lib.Out()
        out.means_dev, out.chols_dev = means.data_ptr(), chols.data_ptr()
        opts = _lib.Opts()
        opts.maxiters, opts.fixed_iters, opts.calibrate, opts.products = 10, n_it, 0, _lib.PRODUCTS_F32
        ctx = _lib.context(0)
        _lib.check(_lib.load().pof_solve_large(ctx, ctypes.byref(prob), ctypes.byref(opts), ctypes.byref(out),
                                               _lib.stream_ptr()), ctx, "pof_solve_large")
        return MVNSqrt(means, chols), {"sigma_squared": float(out.sigma_squared)}

    t, finite, elem, norm, cov_err, ssq_diff = measure("float64")
    try:
        if not fp32_products:
            raise RuntimeError("not run (--no-fp32-products)")
        t32, finite32, elem32, norm32, cov32, ssq32 = measure("float32")
        fp32 = {
            "what": "experiment, refused by the public API: scan combines' matrix products as FP32 MFMA GEMMs "
                    "(FP32 accumulate); factorisations, triangular solves, elements and observation statistics "
                    "FP64; calibrate = 0 (sigma_squared is rounding noise in FP32, so cov_rel_err and "
                    "sigma_squared_rel_diff compare uncalibrated, meaningless factors)",
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
        batched = batched_run(args.batch_ivps, maxiters=100, plain_sample=256, cpu=not args.no_cpu_baseline)
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
    known = None if args.no_checks else known_answers()
    small = None if args.no_checks else small_state_run()
    if args.verbose:
        print(json.dumps({"known_answers": known, "small_state": small}), file=sys.stderr, flush=True)
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
        "known_answers": known,
        "small_state": small,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
