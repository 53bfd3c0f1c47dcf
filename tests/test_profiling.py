"""The rocprofv3 summaries of cloverleaf3d_ref_amd.profiling, on small
hand-made CSV files in rocprofv3's column layout."""
from __future__ import annotations

import csv

import pytest

from cloverleaf3d_ref_amd import profiling as P

D = "void clover::k_fast_accel_pdv<double>(clover::Lay, double const*)"
F = "void clover::k_fast_accel_pdv<float>(clover::Lay, float const*)"
MZ = "void clover::k_fast_mom_march<3, double>(clover::Lay)"
R = "void clover::k_reflect<double>(clover::Lay)"


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f, quoting=csv.QUOTE_ALL)
        w.writerow(header)
        w.writerows(rows)


def _stats(tmp_path):
    _write(tmp_path / "run_kernel_stats.csv", ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"],
           [[D, 3, 9000, 3000, 40], [F, 2, 3000, 1500, 10], [MZ, 2, 4000, 2000, 20], [R, 4, 400, 100, 5],
            ["k_min_final", 1, 50, 50, 1]])
    trace = [  # name, start, end: trials first, then the steps' own launches
        (D, 0, 5000), (D, 10000, 12000), (MZ, 13000, 15000), (D, 20000, 22000), (MZ, 23000, 25000),
        (F, 30000, 31000), (F, 40000, 42000)]
    _write(tmp_path / "run_kernel_trace.csv", ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
           [[n, s, e] for n, s, e in trace])


def test_names_and_type_filter(tmp_path):
    _stats(tmp_path)
    rows = P.kernel_stats(str(tmp_path), "double")
    assert [r[0] for r in rows] == ["k_fast_accel_pdv", "k_fast_mom_z", "k_reflect", "k_min_final"]
    assert sum(r[4] for r in rows) == pytest.approx(100.0)
    rows = P.kernel_stats(str(tmp_path), "float")
    assert [r[0] for r in rows] == ["k_fast_accel_pdv", "k_min_final"]
    assert rows[0][1:4] == (2, 3.0, 1.5)


def test_last_launches(tmp_path):
    _stats(tmp_path)
    tail = P.last_launches(str(tmp_path), 2, "double")
    assert tail["k_fast_accel_pdv"] == (2, 2.0)  # the 5 us trial is dropped
    assert tail["k_fast_mom_z"] == (2, 2.0)
    assert P.last_launches(str(tmp_path), 5, "float")["k_fast_accel_pdv"] == (2, 1.5)


def test_traffic_from_request_counters(tmp_path):
    rd, wr = tmp_path / "r", tmp_path / "w"
    rd.mkdir()
    wr.mkdir()
    hdr = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    G = "void clover::k_fast_ideal_gas<double>(clover::Lay)"
    rows = []
    # dispatch 1: the check kernel at 4^3 FP64 (2 arrays = 1024 B read, as 8 128-byte requests)
    rows += [[1, G, "TCC_EA0_RDREQ_128B_sum", 8], [1, G, "TCC_EA0_RDREQ_64B_sum", 0], [1, G, "TCC_EA0_RDREQ_32B_sum", 0]]
    # dispatches 2, 3: a trial and a timed launch of accel_pdv; only the last counts with --last 1
    rows += [[2, D, "TCC_EA0_RDREQ_128B_sum", 100], [2, D, "TCC_EA0_RDREQ_64B_sum", 0], [2, D, "TCC_EA0_RDREQ_32B_sum", 0]]
    rows += [[3, D, "TCC_EA0_RDREQ_128B_sum", 10], [3, D, "TCC_EA0_RDREQ_64B_sum", 3], [3, D, "TCC_EA0_RDREQ_32B_sum", 2]]
    _write(rd / "run_counter_collection.csv", hdr, rows)
    rows = [[1, G, "TCC_EA0_WRREQ_sum", 16], [1, G, "TCC_EA0_WRREQ_64B_sum", 16],
            [2, D, "TCC_EA0_WRREQ_sum", 50], [2, D, "TCC_EA0_WRREQ_64B_sum", 50],
            [3, D, "TCC_EA0_WRREQ_sum", 7], [3, D, "TCC_EA0_WRREQ_64B_sum", 5]]
    _write(wr / "run_counter_collection.csv", hdr, rows)
    t = P.traffic(str(rd), str(wr), 1, 4, 8)
    k = t["kernels"]["k_fast_accel_pdv"]
    assert k["dispatches"] == 1
    assert k["read_bytes"] == 10 * 128 + 3 * 64 + 2 * 32
    assert k["write_bytes"] == 5 * 64 + 2 * 32
    assert k["hbm_bytes"] == k["read_bytes"] + k["write_bytes"]
    assert t["check"]["read_ratio"] == 1.0 and t["check"]["write_ratio"] == 1.0
    t = P.traffic(str(rd), str(wr), 2, 4, 8)
    assert t["kernels"]["k_fast_accel_pdv"]["read_bytes"] == (100 * 128 + 10 * 128 + 3 * 64 + 2 * 32) / 2
