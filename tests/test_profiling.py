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


def test_solve_table_cuts_one_solve_out_of_a_trace(tmp_path):
    """Kernels from the second marker launch up to the first idle gap above
    300 us, grouped by name and grid, with span and in-kernel time."""
    path = tmp_path / "trace.csv"
    rows = []
    t = 0

    def launch(name, grid, dur, gap=1000):
        nonlocal t
        t += gap
        rows.append([name + "(args)", grid, 1, 1, t, t + dur])
        t += dur

    for solve in range(3):
        launch("k_dc_scale", 64, 2000, gap=10_000_000)
        for _ in range(2):
            launch("k_sweep_fast<float>", 1000, 90_000)
        launch("k_add_interior", 5000, 60_000)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z", "Start_Timestamp", "End_Timestamp"])
        w.writerows(rows)
    out = Pf.solve_table(str(path), index=1)
    # 2 + 90 + 90 + 60 us in kernels, 3 gaps of 1 us, the 10 ms gap excluded
    assert out.startswith("4 kernels, span 0.245 ms, 0.242 ms in kernels")
    assert "| `k_sweep_fast<float>` | 1000 | 2 | 180.0 | 90.0 |" in out
    assert "| `k_add_interior` | 5000 | 1 | 60.0 | 60.0 |" in out
