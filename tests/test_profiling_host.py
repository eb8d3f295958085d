"""The rocprofv3 summarisers of ``pof.profiling`` on small synthetic counter files
(host logic only): cycle-weighted busy fractions, instructions per wave, the
split by launch width, and the traffic read-back the bench uses."""
import csv

import pytest

FIELDS = ["Correlation_Id", "Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(path, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        w.writeheader()
        for did, grid, name, counters in rows:
            for k, v in counters.items():
                w.writerow({"Correlation_Id": did, "Dispatch_Id": did, "Grid_Size": grid, "Kernel_Name": name,
                            "Counter_Name": k, "Counter_Value": v})


def test_issue_rows_weight_by_active_cycles(tmp_path):
    from pof import profiling

    up = "void pof::k_filter_up<12>(pof::GroupArgs, pof::LaunchMap, int)"
    busy = tmp_path / "busy.csv"
    insts = tmp_path / "insts.csv"
    # two dispatches of one kernel: 80% busy over 300 cycles and 20% over 100 -> 65% weighted
    _write(busy, [(1, 4096, up, {"VALUBusy": 80.0, "SALUBusy": 4.0, "GRBM_GUI_ACTIVE": 300.0}),
                  (2, 256, up, {"VALUBusy": 20.0, "SALUBusy": 8.0, "GRBM_GUI_ACTIVE": 100.0}),
                  (3, 256, "__amd_rocclr_copyBuffer", {"VALUBusy": 1.0, "SALUBusy": 1.0, "GRBM_GUI_ACTIVE": 5.0})])
    _write(insts, [(1, 4096, up, {"SQ_WAVES": 64, "SQ_INSTS_VALU": 6400, "SQ_INSTS_SALU": 640, "SQ_INSTS_LDS": 320,
                                  "SQ_INSTS_VALU_FMA_F64": 3000, "SQ_INSTS_VALU_MUL_F64": 100,
                                  "SQ_INSTS_VALU_ADD_F64": 100, "SQ_INSTS_VALU_TRANS_F64": 0}),
                   (2, 256, up, {"SQ_WAVES": 4, "SQ_INSTS_VALU": 400, "SQ_INSTS_SALU": 40, "SQ_INSTS_LDS": 20,
                                 "SQ_INSTS_VALU_FMA_F64": 200, "SQ_INSTS_VALU_MUL_F64": 0,
                                 "SQ_INSTS_VALU_ADD_F64": 0, "SQ_INSTS_VALU_TRANS_F64": 0})])
    rows = profiling.issue_rows(str(busy), str(insts))
    assert len(rows) == 1  # the runtime's copy kernel is not a pof:: kernel
    r = rows[0]
    assert r["kernel"] == "pof::k_filter_up<12>" and r["dispatches"] == 2
    assert r["valu_busy"] == pytest.approx(65.0)
    assert r["salu_busy"] == pytest.approx((4.0 * 300 + 8.0 * 100) / 400)
    assert r["valu_per_wave"] == pytest.approx(100.0)
    assert r["salu_per_wave"] == pytest.approx(10.0)
    assert r["lds_per_wave"] == pytest.approx(5.0)
    assert r["f64_per_wave"] == pytest.approx(3400 / 68)
    assert r["f64_share"] == pytest.approx(3400 / 6800)
    lv = profiling.busy_by_grid(str(busy), {"pof::k_filter_up<12>"})
    assert lv == {("pof::k_filter_up<12>", 4096): (1, 80.0), ("pof::k_filter_up<12>", 256): (1, 20.0)}
    md = profiling.issue_markdown(str(busy), str(insts), "cmd", ["pof::k_filter_up<12>"])
    assert "| `pof::k_filter_up<12>` | 2 | 100.0 | 65.0 |" in md and "| 4096 | 1 | 80.0 |" in md


def test_read_traffic_from_committed_summary():
    from pathlib import Path

    from pof import profiling

    md = Path(__file__).resolve().parents[1] / "profiles" / "ieks_lorenz63_Ns65536_q3_f64_hbm_traffic.md"
    b = profiling.read_traffic(str(md), "k_filter_up<12")
    # the measured bytes per launch sit within 5% of the 44.8 MB algorithmic figure (every
    # element of the level read and written once)
    assert b is not None and abs(b / 44825940 - 1.0) < 0.05
