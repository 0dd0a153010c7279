"""The measurement helpers of bench.py on synthetic rocprofv3 CSV output
(no GPU): the kernel table of the last MD steps and the HBM bytes per MD
step from FETCH_SIZE / WRITE_SIZE.  Each case is worked by hand: two steps
of two kernels after a set-up dispatch and a calibration copy."""
import csv
import json
import os

import pytest

import bench


def _write(path, rows, fields):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def _trace(d):
    # dispatch id, kernel, start, end (ns): a set-up k_legs, then two steps
    # of k_md_pos + k_legs; the last step's k_legs takes 30 ns
    rows = [(1, "void fibk::k_legs<true, false, false>(int)", 0, 100),
            (2, "fibk::k_md_pos(double4*)", 200, 210),
            (3, "void fibk::k_legs<true, false, false>(int)", 220, 240),
            (4, "fibk::k_md_pos(double4*)", 300, 312),
            (5, "void fibk::k_legs<true, false, false>(int)", 320, 350)]
    _write(os.path.join(d, "run_kernel_trace.csv"),
           [{"Dispatch_Id": a, "Kernel_Name": b, "Start_Timestamp": c, "End_Timestamp": e}
            for a, b, c, e in rows],
           ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])


def test_last_steps_stats(tmp_path):
    _trace(tmp_path)
    rows = {bench.short_name(r["Name"]): r for r in bench._last_steps_stats(str(tmp_path), 1)}
    # only the last step: k_md_pos 12 ns, k_legs 30 ns
    assert set(rows) == {"k_md_pos", "k_legs<true, false, false>"}
    assert rows["k_md_pos"]["Calls"] == 1 and rows["k_md_pos"]["AverageNs"] == 12
    assert rows["k_legs<true, false, false>"]["AverageNs"] == 30
    assert rows["k_legs<true, false, false>"]["Percentage"] == pytest.approx(100 * 30 / 42)
    rows = {bench.short_name(r["Name"]): r for r in bench._last_steps_stats(str(tmp_path), 2)}
    assert rows["k_legs<true, false, false>"]["Calls"] == 2
    assert rows["k_legs<true, false, false>"]["AverageNs"] == 25  # (20 + 30) / 2, set-up left out
    assert rows["k_md_pos"]["MinNs"] == 10 and rows["k_md_pos"]["MaxNs"] == 12


def _pmc(d, counter, copy_kib, per_kernel_kib):
    rows = [{"Dispatch_Id": 1, "Kernel_Name": "__amd_rocclr_copyBuffer", "Counter_Name": counter,
             "Counter_Value": copy_kib}]
    k = 2
    for step in range(bench.PROFILE_STEPS + 1):  # one step more than the summary takes
        for name, kib in per_kernel_kib:
            rows.append({"Dispatch_Id": k, "Kernel_Name": name, "Counter_Name": counter,
                         "Counter_Value": kib * (1000 if step == 0 else 1)})
            k += 1
    os.makedirs(d, exist_ok=True)
    _write(os.path.join(d, "run_counter_collection.csv"), rows,
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])


def test_summarize_step_pmc(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "STEP_PMC_FILE", str(tmp_path / "step.json"))
    # the copy reads half its size in FETCH_SIZE (factor 2) and its size in
    # WRITE_SIZE (factor 1); per step k_md_pos fetches 10 KiB raw, k_legs 5
    _pmc(tmp_path / "f", "FETCH_SIZE", bench.COPY_KIB / 2,
         [("fibk::k_md_pos(double4*)", 10.0), ("fibk::k_legs<true, false, false>(int)", 5.0)])
    _pmc(tmp_path / "w", "WRITE_SIZE", bench.COPY_KIB,
         [("fibk::k_md_pos(double4*)", 4.0), ("fibk::k_legs<true, false, false>(int)", 1.0)])
    bench.summarize_step_pmc(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "out.txt"),
                             "cmd")
    res = json.load(open(tmp_path / "step.json"))
    assert res["steps"] == bench.PROFILE_STEPS
    assert res["fetch_factor"] == pytest.approx(2.0) and res["write_factor"] == pytest.approx(1.0)
    k = res["kernels"]
    # the first (outlier) step is not among the last PROFILE_STEPS
    assert k["k_md_pos"]["fetch_bytes"] == pytest.approx(2 * 10 * 1024)
    assert k["k_md_pos"]["write_bytes"] == pytest.approx(4 * 1024)
    assert k["k_legs<true, false, false>"]["bytes_per_step"] == pytest.approx((2 * 5 + 1) * 1024)
    assert res["traffic_per_step_bytes"] == pytest.approx((20 + 4 + 10 + 1) * 1024)
    assert "total" in open(tmp_path / "out.txt").read()
