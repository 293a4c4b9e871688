"""bench.py's profile summaries on small hand-made rocprofv3 CSVs (no GPU):
the ROCTX window filter of --trace-summary and the calibration arithmetic of
--counter-summary."""
import csv
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(header)
        w.writerows(rows)


def test_trace_summary_keeps_only_the_timed_window(tmp_path):
    d = str(tmp_path / "trace")
    _write(os.path.join(d, "k_kernel_trace.csv"), ["Kernel_Name", "Grid_Size", "Start_Timestamp", "End_Timestamp"],
           [["a", 64, 0, 5000],            # before the window: dropped
            ["a", 64, 10000, 12000], ["a", 64, 20000, 24000], ["b", 128, 30000, 31000],
            ["b", 128, 95000, 99000]])      # after the window: dropped
    _write(os.path.join(d, "k_marker_api_trace.csv"), ["Function", "Start_Timestamp", "End_Timestamp"],
           [["timed frames", 9000, 40000]])
    out = bench.trace_summary(d, frames=2)
    assert "31.0 us wall" in out and "kernel time 7.0 us" in out
    rows = {r.split("|")[1].strip(" `"): r for r in out.splitlines() if r.startswith("| `")}
    assert set(rows) == {"a", "b"}
    # a: 2 calls, 6 us total, 3 us average, 3 us per frame
    assert [c.strip() for c in rows["a"].split("|")[2:7]] == ["2", "6.0", "3.00", "85.71", "3.0"]


def test_counter_summary_scales_by_the_calibration_copy(tmp_path):
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    kib = bench.CALIBRATION_BYTES / 1024
    fd, wd = str(tmp_path / "f"), str(tmp_path / "w")
    _write(os.path.join(fd, "k_counter_collection.csv"), hdr,
           [["__amd_rocclr_copyBuffer", "FETCH_SIZE", 3.0],         # small runtime copies: ignored
            ["__amd_rocclr_copyBuffer", "FETCH_SIZE", kib / 2],      # the calibration copy: scale 2
            ["void xpbd::k_jacobi_tiled<0u, true>(x)", "FETCH_SIZE", 1000.0],
            ["void xpbd::k_jacobi_tiled<0u, true>(x)", "FETCH_SIZE", 1002.0],
            ["xpbd::k_coll_sweep(xpbd::CollArgs)", "FETCH_SIZE", 50.0]])
    _write(os.path.join(wd, "k_counter_collection.csv"), hdr,
           [["__amd_rocclr_copyBuffer", "WRITE_SIZE", kib],          # scale 1
            ["void xpbd::k_jacobi_tiled<0u, true>(x)", "WRITE_SIZE", 400.0],
            ["void xpbd::k_jacobi_tiled<0u, true>(x)", "WRITE_SIZE", 400.0],
            ["xpbd::k_coll_sweep(xpbd::CollArgs)", "WRITE_SIZE", 10.0]])
    table, rec = bench.counter_summary(fd, wd)
    assert "scale 2.00000" in table and "scale 1.00000" in table
    assert rec["read_bytes"] == 1001 * 2 * 1024 and rec["write_bytes"] == 400 * 1024
    assert rec["bytes_per_launch"] == rec["read_bytes"] + rec["write_bytes"]
    assert rec["collision_sweep_bytes_per_launch"] == (50 * 2 + 10) * 1024
    assert "__amd_rocclr" not in table.split("|---|")[1]


def test_counter_summary_refuses_a_run_without_the_calibration_copy(tmp_path):
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    fd, wd = str(tmp_path / "f"), str(tmp_path / "w")
    _write(os.path.join(fd, "k_counter_collection.csv"), hdr,
           [["__amd_rocclr_copyBuffer", "FETCH_SIZE", 3.0], ["xpbd::k_coll_sweep(a)", "FETCH_SIZE", 50.0]])
    _write(os.path.join(wd, "k_counter_collection.csv"), hdr,
           [["__amd_rocclr_copyBuffer", "WRITE_SIZE", 2.0], ["xpbd::k_coll_sweep(a)", "WRITE_SIZE", 10.0]])
    with pytest.raises(SystemExit):
        bench.counter_summary(fd, wd)
