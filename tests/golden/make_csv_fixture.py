"""Extract the IEKS pins of the reference's work-precision CSVs into a small fixture.

Source: ``experiments/3_work_precision_diagram/data/<problem>_Tesla_V100-SXM2-32GB.csv``
of the reference (FP64 runs of ``solve(init="constant", maxiters=1000)`` on
``ts = linspace(t0, tmax, N)``, ``run_benchmark.py:63-80,143-148,179``).
Only the columns ``Ns`` (= len(ts)), ``IEKS(q)_iterations``,
``IEKS(q)_rmse_final``, ``IEKS(q)_rmse_traj``, ``sIEKS(q)_iterations``,
``IEKS(q)_runtime`` (seconds on the V100, min of 3 timed runs) (q = 1, 2, 3)
are kept, values copied verbatim:

* ``work_precision_pins.csv``: rows with N <= 256, the sizes the CPU oracle
  runs in seconds (checked against the oracle on the CPU);
* ``work_precision_pins_large.csv``: rows with N in {8192, 65536}, full-size
  known answers for the GPU solve (N = 65536 is the length of the Lorenz-63
  benchmark grid, at the same state dimension D = 12 for q = 3).

Usage: ``python tests/golden/make_csv_fixture.py /root/reference``
"""
import csv
import sys
from pathlib import Path

PROBLEMS = ["logistic", "fhn", "rigidbody", "vdp0", "henonheiles"]
MAX_N = 256
LARGE_N = (8192, 65536)


def main(ref_root):
    write(ref_root, "work_precision_pins.csv", lambda N: N <= MAX_N)
    write(ref_root, "work_precision_pins_large.csv", lambda N: N in LARGE_N)


def write(ref_root, name, keep):
    src_dir = Path(ref_root) / "experiments" / "3_work_precision_diagram" / "data"
    out = Path(__file__).with_name(name)
    rows = []
    for prob in PROBLEMS:
        with open(src_dir / f"{prob}_Tesla_V100-SXM2-32GB.csv") as f:
            for rec in csv.DictReader(f):
                N = int(rec["Ns"])
                if not keep(N):
                    continue
                for q in (1, 2, 3):
                    rows.append(
                        {
                            "problem": prob,
                            "N": N,
                            "q": q,
                            "iterations": rec[f"IEKS({q})_iterations"],
                            "rmse_final": rec[f"IEKS({q})_rmse_final"],
                            "rmse_traj": rec[f"IEKS({q})_rmse_traj"],
                            "seq_iterations": rec[f"sIEKS({q})_iterations"],
                            "runtime": rec[f"IEKS({q})_runtime"],
                        }
                    )
    with open(out, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)
    print(f"wrote {len(rows)} rows to {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
