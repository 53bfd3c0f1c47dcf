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


def test_traffic_scaled_by_the_copy(tmp_path):
    fetch, write = tmp_path / "f", tmp_path / "w"
    fetch.mkdir()
    write.mkdir()
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    # the counters read a 1 MiB copy as 640 KiB fetched and 1600 KiB written
    _write(fetch / "run_counter_collection.csv", hdr,
           [["__amd_rocclr_copyBuffer", "FETCH_SIZE", 640], [D, "FETCH_SIZE", 100], [D, "FETCH_SIZE", 300]])
    _write(write / "run_counter_collection.csv", hdr,
           [["__amd_rocclr_copyBuffer", "WRITE_SIZE", 1600], [D, "WRITE_SIZE", 400], [D, "WRITE_SIZE", 400]])
    t = P.traffic(str(fetch), str(write), 1)
    assert t["fetch_scale"] == 1.6 and t["write_scale"] == 0.64
    k = t["kernels"]["k_fast_accel_pdv"]
    assert k["dispatches"] == 2
    assert k["fetch_bytes"] == pytest.approx(200 * 1024 * 1.6)
    assert k["write_bytes"] == pytest.approx(400 * 1024 * 0.64)
    assert k["hbm_bytes"] == k["fetch_bytes"] + k["write_bytes"]
