"""Full-size known answers: the reference's own work-precision runs at N = 8192 and 65536.

``tests/golden/work_precision_pins_large.csv`` holds, verbatim, the rows of the
reference's ``experiments/3_work_precision_diagram/data/*_Tesla_V100-SXM2-32GB.csv``
at those sizes: FP64 ``solve(init="constant", maxiters=1000)`` on
``ts = linspace(t0, tmax, N)`` (``run_benchmark.py:63-80,143-148,179``), with the
iteration count and the errors ``rmse_final = |y_N - y*(T)|``,
``rmse_traj = mean_k |y_k - y*(t_k)|`` against a Kvaerno5 reference solution at
``atol = 1e-16``, ``rtol = 1e-13``.  The GPU solve runs the same configuration
here; the reference solution is recomputed with scipy's DOP853 at
``rtol = 1e-13``, as in tests/test_oracle.py.  At N = 65536 (65535 steps) and
q = 3 the state dimension is 12, the size of the Lorenz-63 benchmark, and
every scan level of a non-power-of-two length runs.

Pinned: the iteration count, exactly, wherever the reference's parallel and
sequential IEKS agree on it (where they differ the count depends on rounding,
and only convergence is required); ``rmse_final`` / ``rmse_traj`` to 1e-5
relative plus 1e-10 absolute (the two reference solutions, Kvaerno5 there and
DOP853 here, differ by a few 1e-11) where they exceed 1e-8; below that they sit
at that error floor and only their order of magnitude is checked.  The
configurations whose parallel IEKS diverges in the reference (Henon-Heiles,
q = 3: 1001 iterations, errors ~1e6..1e8) diverge or converge by rounding (the
oracle converges at N = 8192): either is accepted, a converged solve only with
errors below 1e-6.
"""
import csv
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"

# Rows whose iteration count depends on rounding although the reference's parallel and
# sequential counts agree: the oracle reproduces the CSV's 25 iterations for Henon-Heiles q = 2
# at N = 65536, but its objective passes through ~1e31 on the way (the iterates wander by ten
# orders of magnitude before converging), so a different summation order converges at another
# iteration (37 on the GPU).  Only convergence and the converged solution's errors are pinned.
ROUNDING_SENSITIVE = {("henonheiles", 65536, 2)}


def _rows():
    with open(GOLDEN / "work_precision_pins_large.csv") as f:
        return list(csv.DictReader(f))


ROWS = _rows()


def _problem(name):
    import pof.ivp

    return {
        "logistic": lambda: pof.ivp.logistic(),
        "fhn": lambda: pof.ivp.fitzhughnagumo(),
        "rigidbody": lambda: pof.ivp.rigid_body(),
        "vdp0": lambda: pof.ivp.vanderpol(stiffness_constant=1.0),
        "henonheiles": lambda: pof.ivp.henonheiles(tmax=10.0),
    }[name]()


_REF = {}


def _reference(name, ts):
    """DOP853 solution at ts (rtol 1e-13, atol 1e-15), cached per problem and grid."""
    from scipy.integrate import solve_ivp

    from oracle import problems as op

    key = (name, len(ts))
    if key not in _REF:
        p = {
            "logistic": op.logistic,
            "fhn": op.fitzhughnagumo,
            "rigidbody": op.rigid_body,
            "vdp0": lambda: op.vanderpol(stiffness_constant=1.0),
            "henonheiles": lambda: op.henonheiles(tmax=10.0),
        }[name]()
        sol = solve_ivp(lambda _, y: p.f(None, y), p.t_span, p.y0, method="DOP853", rtol=1e-13, atol=1e-15,
                        dense_output=True)
        _REF[key] = sol.sol(ts).T
    return _REF[key]


@pytest.mark.parametrize("row", ROWS, ids=[f"{r['problem']}-N{r['N']}-q{r['q']}" for r in ROWS])
def test_full_size_work_precision_row(gpu, row):
    import torch

    import pof.solver

    ivp = _problem(row["problem"])
    N, q = int(row["N"]), int(row["q"])
    ts = np.linspace(ivp.t0, ivp.tmax, N)
    print(f"case {row['problem']} N={N} q={q}", flush=True)
    ys, info = pof.solver.solve(f=ivp.f, y0=ivp.y0, ts=ts, order=q, init="constant", maxiters=1000)
    torch.cuda.synchronize()
    want_it = float(row["iterations"])
    m = ys.mean.cpu().numpy()
    if want_it == 1001.0:
        # The reference's parallel IEKS diverges here (its sequential IEKS converges), but whether the
        # parallel iteration diverges is decided by rounding: the oracle converges for Henon-Heiles
        # q = 3 at N = 8192, in 28, 34 or 50 iterations for y0, y0 (1 + 1e-15) and y0 (1 - 1e-15).
        # Either outcome is accepted; a converged solve must be accurate.
        if info["status"] == "maxiter":
            assert info["iterations"] == 1001.0
            assert not np.isfinite(m).all() or np.abs(m).max() > 1e3
            return
        assert info["status"] == "converged", info["status"]
        yref = _reference(row["problem"], ts)
        rf = float(np.linalg.norm(m[-1] - yref[-1]))
        rt = float(np.linalg.norm(m - yref, axis=1).mean())
        print(f"  converged in {info['iterations']} iterations: rmse_final {rf:.3e} rmse_traj {rt:.3e}", flush=True)
        assert rf < 1e-6 and rt < 1e-6, (rf, rt)
        return
    if row["iterations"] == row["seq_iterations"] and (row["problem"], N, q) not in ROUNDING_SENSITIVE:
        assert info["iterations"] == want_it, (info["iterations"], want_it)
    else:
        # rounding-sensitive in the reference itself (FHN q = 3 at N = 65536: parallel 80, sequential
        # 62; measured here 113): a converged solve
        assert info["status"] == "converged" and info["iterations"] <= 1000
    yref = _reference(row["problem"], ts)
    rf = float(np.linalg.norm(m[-1] - yref[-1]))
    rt = float(np.linalg.norm(m - yref, axis=1).mean())
    print(f"  iterations {info['iterations']} (ref {want_it}) rmse_final {rf:.6e} (ref {row['rmse_final']}) "
          f"rmse_traj {rt:.6e} (ref {row['rmse_traj']})", flush=True)
    for got, want in ((rf, float(row["rmse_final"])), (rt, float(row["rmse_traj"]))):
        if want > 1e-8:
            assert abs(got - want) <= 1e-5 * want + 1e-10, (got, want)
        else:
            assert got < 1e-8, (got, want)
