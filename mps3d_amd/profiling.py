"""Summaries of rocprofv3 runs of ``bench.py`` (text for ``profiles/``).

Three runs feed them, each of the same command
(``python3 bench.py --steps 2 --warmup 1 --no-cpu``):

* ``rocprofv3 --kernel-trace --stats --output-format csv`` gives the
  per-kernel time table (``*_kernel_stats.csv``) and the per-dispatch trace
  (``*_kernel_trace.csv``);
* ``rocprofv3 --pmc FETCH_SIZE --output-format csv`` and
  ``rocprofv3 --pmc WRITE_SIZE --output-format csv`` give the HBM-side bytes
  of every dispatch (``*_counter_collection.csv``, KiB per dispatch).

The counters are scaled by what they read for a plain device copy of known
size, which ``bench.py`` runs once (``hipMemcpyAsync`` of ``COPY_BYTES``
bytes, the ``__amd_rocclr_copyBuffer`` kernel): factor = known bytes /
counter bytes, separately for reads and writes.

Usage:
    python -m mps3d_amd.profiling stats <kernel_stats.csv> [<kernel_trace.csv>]
    python -m mps3d_amd.profiling solve <kernel_trace.csv> [--index I] [--marker NAME]
    python -m mps3d_amd.profiling traffic <fetch.csv> <write.csv> --kernel NAME [--grid THREADS]
    python -m mps3d_amd.profiling table <fetch.csv> <write.csv> NAME [NAME ...]
"""
from __future__ import annotations

import argparse
import csv
import json
import sys
from collections import defaultdict

COPY_BYTES = 512 * 1024 * 1024  # bytes read (and written) by the calibration copy
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _short(name, n=90):
    name = name.replace("mps::", "")
    return name if len(name) <= n else name[: n - 3] + "..."


def stats_table(stats_csv, trace_csv=None, top=25):
    """Markdown table of kernel time from ``--stats`` (and, with the trace,
    the average per distinct grid of the top kernels)."""
    rows = list(csv.DictReader(open(stats_csv)))
    out = ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append("| `%s` | %s | %.3f | %.2f | %.1f |" % (
            _short(r["Name"]), r["Calls"], float(r["TotalDurationNs"]) / 1e6, float(r["AverageNs"]) / 1e3,
            float(r["Percentage"])))
    if trace_csv:
        by = defaultdict(list)
        for r in csv.DictReader(open(trace_csv)):
            key = (r["Kernel_Name"], _grid_total(r))
            by[key].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        ranked = sorted(by.items(), key=lambda kv: -sum(kv[1]))[:top]
        out += ["", "| kernel | grid (threads) | calls | avg us | total ms |", "|---|---|---|---|---|"]
        for (name, g), d in ranked:
            out.append("| `%s` | %d | %d | %.2f | %.3f |" % (
                _short(name, 60), g, len(d), sum(d) / len(d) / 1e3, sum(d) / 1e6))
    return "\n".join(out)


def solve_table(trace_csv, index=1, marker="k_dc_scale", top=12):
    """Markdown table of one solve out of a kernel trace: the kernels from
    the index-th launch of `marker` (one per FP32+DC solve) up to the first
    idle gap above 300 us, by kernel and launch shape, with the solve's span
    (first start to last end) and the time inside kernels."""
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    marks = [i for i, r in enumerate(rows) if marker in r["Kernel_Name"]]
    if len(marks) <= index:
        raise SystemExit("%d launches of %s in %s" % (len(marks), marker, trace_csv))
    seg = rows[marks[index]:marks[index + 1] if index + 1 < len(marks) else len(rows)]
    for i in range(1, len(seg)):
        if int(seg[i]["Start_Timestamp"]) - int(seg[i - 1]["End_Timestamp"]) > 300000:
            seg = seg[:i]
            break
    span = (int(seg[-1]["End_Timestamp"]) - int(seg[0]["Start_Timestamp"])) / 1e6
    busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in seg) / 1e6
    agg = defaultdict(list)
    for r in seg:
        agg[(_short(r["Kernel_Name"].split("(")[0], 70), _grid_total(r))].append(
            int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    out = ["%d kernels, span %.3f ms, %.3f ms in kernels" % (len(seg), span, busy), "",
           "| kernel | grid (threads) | launches | total us | avg us |", "|---|---|---|---|---|"]
    for (name, g), d in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append("| `%s` | %d | %d | %.1f | %.1f |" % (name, g, len(d), sum(d) / 1e3, sum(d) / len(d) / 1e3))
    return "\n".join(out)


def _grid_total(r):
    """Total grid size in threads of a CSV row (either column layout)."""
    if "Grid_Size_X" in r:
        return int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
    return int(r["Grid_Size"])


def _dispatches(counter_csv, counter):
    """{dispatch id: (kernel name, grid threads, counter value in bytes)}"""
    vals = {}
    for r in csv.DictReader(open(counter_csv)):
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        name, g, v = vals.get(d, (r["Kernel_Name"], _grid_total(r), 0.0))
        vals[d] = (name, g, v + float(r["Counter_Value"]) * 1024.0)
    return vals


def traffic(fetch_csv, write_csv, kernel, grid=None):
    """HBM bytes per dispatch of ``kernel`` (optionally only dispatches of one
    grid size), scaled by the calibration copy."""
    out = {}
    for label, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        disp = _dispatches(path, counter)
        copies = [v for (n, g, v) in disp.values() if COPY_KERNEL in n and v > 0.1 * COPY_BYTES]
        if not copies:
            raise SystemExit("no calibration copy of %d bytes in %s" % (COPY_BYTES, path))
        copy = max(copies)
        factor = COPY_BYTES / copy
        ks = [v for (n, g, v) in disp.values() if kernel in n and (grid is None or g == grid)]
        if not ks:
            raise SystemExit("no dispatch of %s (grid %s) in %s" % (kernel, grid, path))
        avg = sum(ks) / len(ks)
        out[label] = {"counter": counter, "dispatches": len(ks), "avg_counter_bytes": avg,
                      "copy_counter_bytes": copy, "factor": factor, "bytes": avg * factor}
    out["kernel"] = kernel
    out["grid_threads"] = grid
    out["traffic_bytes"] = out["fetch"]["bytes"] + out["write"]["bytes"]
    return out


def traffic_table(fetch_csv, write_csv, kernels):
    """Markdown table of scaled HBM bytes per dispatch for each kernel name
    filter, at the launch shape that reads the most (the finest level; grid
    sizes alone do not say which level a launch is, since the z-chunking
    differs per level)."""
    disp = _dispatches(fetch_csv, "FETCH_SIZE")
    out = ["| kernel | grid (threads) | dispatches | read MB | write MB | total MB |", "|---|---|---|---|---|---|"]
    for k in kernels:
        per = defaultdict(list)
        for n, g, v in disp.values():
            if k in n:
                per[g].append(v)
        if not per:
            continue
        g = max(per, key=lambda x: sum(per[x]) / len(per[x]))
        t = traffic(fetch_csv, write_csv, k, g)
        out.append("| `%s` | %d | %d | %.1f | %.1f | %.1f |" % (
            k, g, t["fetch"]["dispatches"], t["fetch"]["bytes"] / 1e6, t["write"]["bytes"] / 1e6,
            t["traffic_bytes"] / 1e6))
    t = traffic(fetch_csv, write_csv, COPY_KERNEL)
    out.append("")
    out.append("Calibration copy (%d bytes read and written): FETCH_SIZE factor %.4f, WRITE_SIZE factor %.4f." % (
        COPY_BYTES, t["fetch"]["factor"], t["write"]["factor"]))
    return "\n".join(out)


def counter_table(csvs, kernels, grid=None):
    """Markdown table of the mean value per dispatch of every counter in the
    given --pmc CSVs, for each kernel name filter (dispatches of one grid
    size when given, else of the largest grid the kernel was launched
    with)."""
    vals = defaultdict(lambda: defaultdict(dict))  # kernel filter -> counter -> {dispatch: value}
    grids = defaultdict(dict)
    names = []
    for path in csvs:
        for r in csv.DictReader(open(path)):
            n, c = r["Kernel_Name"], r["Counter_Name"]
            for k in kernels:
                if k in n:
                    d = int(r["Dispatch_Id"])
                    vals[k][c][d] = vals[k][c].get(d, 0.0) + float(r["Counter_Value"])
                    grids[k][d] = _grid_total(r)
            if c not in names:
                names.append(c)
    out = ["| kernel | grid (threads) | dispatches | " + " | ".join(names) + " |",
           "|---|---|---|" + "---|" * len(names)]
    for k in kernels:
        if k not in grids:
            continue
        g = grid if grid is not None else max(grids[k].values())
        cells = []
        nd = 0
        for c in names:
            ds = [v for d, v in vals[k][c].items() if grids[k][d] == g]
            nd = max(nd, len(ds))
            cells.append("%.3g" % (sum(ds) / len(ds)) if ds else "-")
        out.append("| `%s` | %d | %d | %s |" % (k, g, nd, " | ".join(cells)))
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m mps3d_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("stats_csv")
    s.add_argument("trace_csv", nargs="?")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--kernel", required=True)
    t.add_argument("--grid", type=int, default=None, help="total grid size in threads (one launch shape)")
    so = sub.add_parser("solve")
    so.add_argument("trace_csv")
    so.add_argument("--index", type=int, default=1, help="which launch of the marker starts the solve")
    so.add_argument("--marker", default="k_dc_scale")
    tt = sub.add_parser("table")
    tt.add_argument("fetch_csv")
    tt.add_argument("write_csv")
    tt.add_argument("kernels", nargs="+", help="kernel name filters")
    ct = sub.add_parser("counters")
    ct.add_argument("--csv", action="append", required=True, help="a --pmc counter_collection CSV (repeatable)")
    ct.add_argument("--grid", type=int, default=None)
    ct.add_argument("kernels", nargs="+", help="kernel name filters")
    a = ap.parse_args(argv)
    if a.cmd == "counters":
        print(counter_table(a.csv, a.kernels, a.grid))
    elif a.cmd == "table":
        print(traffic_table(a.fetch_csv, a.write_csv, a.kernels))
    elif a.cmd == "stats":
        print(stats_table(a.stats_csv, a.trace_csv))
    elif a.cmd == "solve":
        print(solve_table(a.trace_csv, a.index, a.marker))
    else:
        print(json.dumps(traffic(a.fetch_csv, a.write_csv, a.kernel, a.grid), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
