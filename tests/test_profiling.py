"""The rocprofv3 summary helpers (mps3d_amd/profiling.py) on small
hand-made CSV files: the copy calibration scales the counters, and only the
named kernel at the named grid size is averaged."""
import csv

from mps3d_amd import profiling as Pf


def write_counters(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        for d, name, grid, kib in rows:
            # counters arrive split over several instances (XCDs); they add up
            w.writerow([d, name, grid, counter, kib / 2])
            w.writerow([d, name, grid, counter, kib / 2])


def test_traffic_scales_by_the_calibration_copy(tmp_path):
    copy_kib = Pf.COPY_BYTES / 1024
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    write_counters(f, "FETCH_SIZE", [(1, "k_sweep_fast<float>", 1000, 400.0), (2, "k_sweep_fast<float>", 1000, 600.0),
                                     (3, "k_sweep_fast<float>", 64, 9999.0), (4, "__amd_rocclr_copyBuffer", 5,
                                                                              copy_kib * 0.5)])
    write_counters(w, "WRITE_SIZE", [(1, "k_sweep_fast<float>", 1000, 100.0), (2, "k_sweep_fast<float>", 1000, 100.0),
                                     (4, "__amd_rocclr_copyBuffer", 5, copy_kib)])
    t = Pf.traffic(str(f), str(w), "k_sweep_fast", 1000)
    assert t["fetch"]["dispatches"] == 2
    assert t["fetch"]["factor"] == 2.0 and t["write"]["factor"] == 1.0
    assert t["fetch"]["bytes"] == 500.0 * 1024 * 2
    assert t["write"]["bytes"] == 100.0 * 1024
    assert t["traffic_bytes"] == t["fetch"]["bytes"] + t["write"]["bytes"]
