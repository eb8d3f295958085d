"""Summaries of rocprofv3 output for the IEKS kernels.

* ``stats``: the ``--kernel-trace --stats`` per-kernel table as Markdown.
* ``mfma``: MFMA busy fraction and FP64 MFMA flops per kernel class from one
  ``--pmc MfmaUtil MfmaFlopsF64 GRBM_GUI_ACTIVE`` run.
* ``traffic``: HBM bytes per launch of every kernel from two counter runs
  (``--pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE``, KiB per dispatch), each
  scaled by what the same counter reads for a plain device copy of known
  size (``bench.py --copy-probe``: 2^25 float64 read and written per launch).

``read_traffic`` returns the bytes per launch of one kernel from a committed
traffic summary, for ``bench.py``'s roofline record.

Usage::

    python -m pof.profiling stats RUN_kernel_stats.csv --command "..." > summary.md
    python -m pof.profiling traffic --fetch F.csv --write W.csv --copy-fetch CF.csv \
        --copy-write CW.csv --copy-bytes 268435456 --command "..." > traffic.md
"""
import argparse
import collections
import csv
import re
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(name):
    """Kernel name without ``void`` and the argument list (``(anonymous namespace)::`` kept)."""
    name = name.replace("void ", "").replace("(anonymous namespace)::", "anon::")
    return name.split("(")[0]


def counter_means(path):
    """{kernel name: (dispatches, mean counter value)} of a rocprofv3 counter_collection.csv."""
    agg = collections.defaultdict(lambda: [0, 0.0])
    with open(path) as fh:
        for r in csv.DictReader(fh):
            a = agg[r["Kernel_Name"]]
            a[0] += 1
            a[1] += float(r["Counter_Value"])
    return {k: (n, s / n) for k, (n, s) in agg.items()}


def copy_scale(path, copy_bytes):
    """Known bytes / counter bytes of the calibration copy (largest copy kernel of the run)."""
    means = counter_means(path)
    cands = [v for k, v in means.items() if COPY_KERNEL in k]
    if not cands:
        raise SystemExit(f"no {COPY_KERNEL} dispatch in {path}")
    kib = max(v[1] for v in cands)
    return copy_bytes / (kib * 1024.0)


def traffic_rows(fetch, write, copy_fetch, copy_write, copy_bytes, match="pof::"):
    sf = copy_scale(copy_fetch, copy_bytes)
    sw = copy_scale(copy_write, copy_bytes)
    f = counter_means(fetch)
    w = counter_means(write)
    rows = []
    for k in sorted(f, key=lambda k: -f[k][0] * f[k][1]):
        if match not in k or k not in w:
            continue
        n, fk = f[k]
        _, wk = w[k]
        fb = fk * 1024.0 * sf
        wb = wk * 1024.0 * sw
        rows.append((short_name(k), n, fb, wb, fb + wb))
    return sf, sw, rows


def traffic_markdown(fetch, write, copy_fetch, copy_write, copy_bytes, command=""):
    sf, sw, rows = traffic_rows(fetch, write, copy_fetch, copy_write, copy_bytes)
    out = ["# HBM traffic per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE)", ""]
    if command:
        out += ["Commands:", "", "```", command, "```", ""]
    out += [
        f"Calibration: for a device copy that reads and writes {int(copy_bytes)} bytes per launch, "
        f"true bytes = FETCH_SIZE x {sf:.4f} and WRITE_SIZE x {sw:.4f} (counters in KiB); every row "
        f"below is scaled by these factors and averaged over the launches of the profiled run.",
        "",
        f"fetch_scale={sf:.6f} write_scale={sw:.6f}",
        "",
        "| kernel | launches | fetch MB/launch | write MB/launch | bytes_per_launch |",
        "|---|---|---|---|---|",
    ]
    for name, n, fb, wb, tb in rows:
        out.append(f"| `{name}` | {n} | {fb / 1e6:.3f} | {wb / 1e6:.3f} | {tb:.0f} |")
    return "\n".join(out) + "\n"


def stats_markdown(path, command=""):
    out = ["# Kernel time (rocprofv3 --kernel-trace --stats)", ""]
    if command:
        out += ["Command:", "", "```", command, "```", ""]
    out += ["| kernel | calls | total ms | average us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    with open(path) as fh:
        for r in csv.DictReader(fh):
            out.append(
                f"| `{short_name(r['Name'])[:90]}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                f"{float(r['Percentage']):.2f} |"
            )
    return "\n".join(out) + "\n"


def kernel_class(name):
    """rocBLAS / Tensile GEMM kernels (``Cijk_*``) form one class per data type (``_DB_`` FP64,
    ``_SB_`` FP32); other kernels by short name."""
    k = short_name(name)
    if not k.startswith("Cijk_"):
        return k
    if "_DB_" in k:
        return "rocBLAS DGEMM (Cijk_*, FP64 MFMA)"
    if "_SB_" in k:
        return "rocBLAS SGEMM (Cijk_*, FP32 MFMA)"
    return "rocBLAS GEMM (Cijk_*)"


def mfma_rows(pmc_csv, stats_csv=None):
    """Per kernel class: dispatches, duration-weighted MFMA busy fraction and FP64 MFMA flops.

    ``pmc_csv`` is a counter_collection.csv of ``--pmc MfmaUtil MfmaFlopsF64 GRBM_GUI_ACTIVE``
    (rocprofv3's derived counters: MFMA busy cycles over active cycles x SIMDs, in percent;
    512 flops per F64 MFMA op).  Each dispatch's MfmaUtil is weighted by its GRBM_GUI_ACTIVE
    cycles.  With ``stats_csv`` (kernel stats of the same command) the flops are also given as
    TFLOP/s over the class's kernel time.
    """
    per = collections.defaultdict(dict)  # dispatch -> counters
    names = {}
    with open(pmc_csv) as fh:
        for r in csv.DictReader(fh):
            did = r.get("Dispatch_Id") or r.get("Correlation_Id")
            per[did][r["Counter_Name"]] = float(r["Counter_Value"])
            names[did] = r["Kernel_Name"]
    agg = collections.defaultdict(lambda: [0, 0.0, 0.0, 0.0])  # n, sum util*gui, sum gui, flops
    for did, c in per.items():
        a = agg[kernel_class(names[did])]
        g = c.get("GRBM_GUI_ACTIVE", 0.0)
        a[0] += 1
        a[1] += c.get("MfmaUtil", 0.0) * g
        a[2] += g
        a[3] += c.get("MfmaFlopsF64", 0.0)
    ms = {}
    if stats_csv:
        with open(stats_csv) as fh:
            for r in csv.DictReader(fh):
                k = kernel_class(r["Name"])
                ms[k] = ms.get(k, 0.0) + float(r["TotalDurationNs"]) / 1e6
    rows = []
    for k, (n, uw, g, fl) in sorted(agg.items(), key=lambda x: -x[1][2]):
        rows.append((k, n, uw / g if g else 0.0, g, fl, ms.get(k)))
    tot_g = sum(r[3] for r in rows)
    overall = sum(r[2] * r[3] for r in rows) / tot_g if tot_g else 0.0
    return rows, overall


def mfma_markdown(pmc_csv, stats_csv=None, command=""):
    rows, overall = mfma_rows(pmc_csv, stats_csv)
    out = ["# MFMA utilisation (rocprofv3 --pmc MfmaUtil MfmaFlopsF64 GRBM_GUI_ACTIVE)", ""]
    if command:
        out += ["Commands:", "", "```", command, "```", ""]
    out += [f"Whole run, weighted by active cycles: MFMA busy {overall:.2f}% of SIMD cycles.", ""]
    out += ["| kernel class | dispatches | MFMA busy % | active cycles share % | FP64 MFMA GFLOP "
            "| TFLOP/s over kernel time |",
            "|---|---|---|---|---|---|"]
    tot_g = sum(r[3] for r in rows) or 1.0
    for k, n, u, g, fl, ms in rows[:20]:
        tf = f"{fl / (ms * 1e-3) / 1e12:.2f}" if ms else ""
        out.append(f"| `{k[:80]}` | {n} | {u:.2f} | {100 * g / tot_g:.2f} | {fl / 1e9:.2f} | {tf} |")
    return "\n".join(out) + "\n"


ISSUE_INSTS = ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VALU_FMA_F64",
               "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_TRANS_F64"]


def _per_dispatch(path):
    per = collections.defaultdict(dict)
    names = {}
    with open(path) as fh:
        for r in csv.DictReader(fh):
            did = r.get("Dispatch_Id") or r.get("Correlation_Id")
            per[did][r["Counter_Name"]] = float(r["Counter_Value"])
            names[did] = r["Kernel_Name"]
    return per, names


def issue_rows(busy_csv, insts_csv, match="pof::"):
    """Per kernel class: dispatches, VALU / SALU busy (rocprofv3's VALUBusy and SALUBusy, in percent,
    weighted by each dispatch's GRBM_GUI_ACTIVE cycles) and instructions per wave from the raw SQ
    counters of a second pass (``ISSUE_INSTS``)."""
    per, names = _per_dispatch(busy_csv)
    busy = collections.defaultdict(lambda: [0, 0.0, 0.0, 0.0])  # n, sum valu*g, sum salu*g, sum g
    for did, c in per.items():
        if match not in names[did]:
            continue
        a = busy[short_name(names[did])]
        g = c.get("GRBM_GUI_ACTIVE", 0.0)
        a[0] += 1
        a[1] += c.get("VALUBusy", 0.0) * g
        a[2] += c.get("SALUBusy", 0.0) * g
        a[3] += g
    per, names = _per_dispatch(insts_csv)
    ins = collections.defaultdict(lambda: collections.defaultdict(float))
    for did, c in per.items():
        if match not in names[did]:
            continue
        a = ins[short_name(names[did])]
        for k in ISSUE_INSTS:
            a[k] += c.get(k, 0.0)
    rows = []
    for k, (n, vg, sg, g) in sorted(busy.items(), key=lambda x: -x[1][3]):
        c = ins.get(k, {})
        w = c.get("SQ_WAVES", 0.0) or 1.0
        f64 = sum(c.get(x, 0.0) for x in ISSUE_INSTS[4:])
        valu = c.get("SQ_INSTS_VALU", 0.0)
        rows.append({
            "kernel": k, "dispatches": n, "valu_busy": vg / g if g else 0.0, "salu_busy": sg / g if g else 0.0,
            "cycles_share": g, "valu_per_wave": valu / w, "salu_per_wave": c.get("SQ_INSTS_SALU", 0.0) / w,
            "lds_per_wave": c.get("SQ_INSTS_LDS", 0.0) / w, "f64_per_wave": f64 / w,
            "f64_share": f64 / valu if valu else 0.0,
        })
    tot = sum(r["cycles_share"] for r in rows) or 1.0
    for r in rows:
        r["cycles_share"] /= tot
    return rows


def busy_by_grid(busy_csv, kernels):
    """{(kernel, grid size): (dispatches, VALUBusy weighted by GRBM_GUI_ACTIVE)} for the kernels whose
    short name is in ``kernels``: one entry per scan level width."""
    agg = collections.defaultdict(lambda: [0, 0.0, 0.0])
    cyc = {}
    with open(busy_csv) as fh:
        rows = list(csv.DictReader(fh))
    per = collections.defaultdict(dict)
    meta = {}
    for r in rows:
        did = r.get("Dispatch_Id") or r.get("Correlation_Id")
        per[did][r["Counter_Name"]] = float(r["Counter_Value"])
        meta[did] = (short_name(r["Kernel_Name"]), int(float(r.get("Grid_Size") or 0)))
    for did, c in per.items():
        k, g = meta[did]
        if k not in kernels:
            continue
        a = agg[(k, g)]
        gg = c.get("GRBM_GUI_ACTIVE", 0.0)
        a[0] += 1
        a[1] += c.get("VALUBusy", 0.0) * gg
        a[2] += gg
    for key, (n, vg, g) in agg.items():
        cyc[key] = (n, vg / g if g else 0.0)
    return cyc


def issue_markdown(busy_csv, insts_csv, command="", levels=()):
    rows = issue_rows(busy_csv, insts_csv)
    out = ["# Instruction issue (rocprofv3 --pmc VALUBusy SALUBusy GRBM_GUI_ACTIVE, then SQ instruction counts)", ""]
    if command:
        out += ["Commands:", "", "```", command, "```", ""]
    out += ["| kernel | dispatches | active cycles share % | VALU busy % | SALU busy % | VALU / wave | SALU / wave "
            "| LDS / wave | FP64 arith / wave | FP64 share of VALU % |",
            "|---|---|---|---|---|---|---|---|---|---|"]
    for r in rows[:16]:
        out.append(f"| `{r['kernel'][:60]}` | {r['dispatches']} | {100 * r['cycles_share']:.1f} | {r['valu_busy']:.1f} "
                   f"| {r['salu_busy']:.1f} | {r['valu_per_wave']:.0f} | {r['salu_per_wave']:.0f} "
                   f"| {r['lds_per_wave']:.0f} | {r['f64_per_wave']:.0f} | {100 * r['f64_share']:.1f} |")
    if levels:
        lv = busy_by_grid(busy_csv, set(levels))
        out += ["", "VALU busy % by launch width (one scan level per width):", "",
                "| kernel | grid (threads) | dispatches | VALU busy % |", "|---|---|---|---|"]
        for (k, g) in sorted(lv, key=lambda x: (x[0], -x[1])):
            n, v = lv[(k, g)]
            out.append(f"| `{k}` | {g} | {n} | {v:.1f} |")
    return "\n".join(out) + "\n"


def read_traffic(md_path, kernel):
    """Bytes per launch of the first kernel whose name contains ``kernel``, or None."""
    try:
        text = open(md_path).read()
    except OSError:
        return None
    for line in text.splitlines():
        m = re.match(r"\| `([^`]*)` \| (\d+) \| [\d.]+ \| [\d.]+ \| (\d+) \|", line)
        if m and kernel in m.group(1):
            return float(m.group(3))
    return None


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pof.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-bytes", type=float, default=float(8 << 25))
    t.add_argument("--command", default="")
    m = sub.add_parser("mfma")
    m.add_argument("csv")
    m.add_argument("--stats", default=None)
    m.add_argument("--command", default="")
    i = sub.add_parser("issue")
    i.add_argument("--busy", required=True)
    i.add_argument("--insts", required=True)
    i.add_argument("--command", default="")
    i.add_argument("--levels", nargs="*", default=[], help="kernel short names to split by launch width")
    a = ap.parse_args(argv)
    if a.cmd == "issue":
        sys.stdout.write(issue_markdown(a.busy, a.insts, a.command, a.levels))
    elif a.cmd == "stats":
        sys.stdout.write(stats_markdown(a.csv, a.command))
    elif a.cmd == "mfma":
        sys.stdout.write(mfma_markdown(a.csv, a.stats, a.command))
    else:
        sys.stdout.write(traffic_markdown(a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes, a.command))


if __name__ == "__main__":
    main()
