"""Summarise rocprofv3 CSV output per kernel and launch shape.

    python -m partconv_amd.prof_summary trace   <kernel_trace.csv>
    python -m partconv_amd.prof_summary counter <counter_collection.csv> [--calib-grid G]

`trace` prints calls / mean / min / max duration (us) per (kernel, grid, block);
--split NAME K splits the launches of one shape of the kernels named NAME by
call index mod K (e.g. the levels of a pass that launches one kernel per level).
`counter` prints the mean counter value (KiB per dispatch for FETCH_SIZE and
WRITE_SIZE) per (kernel, grid, counter); with --calib-grid the first copy kernel
of that grid size (the 64 MiB device copy bench.py --calib-copy issues first)
gives the scale factor 65536 KiB / measured for each counter.
"""
from __future__ import annotations

import argparse
import collections
import csv


def _short(name, n=60):
    return name if len(name) <= n else name[:n - 3] + "..."


def trace(path, split=None):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    d = collections.defaultdict(list)
    seen = collections.Counter()
    for r in rows:
        key = (r["Kernel_Name"], r.get("Grid_Size_X", r.get("Grid_Size", "")),
               r.get("Workgroup_Size_X", r.get("Workgroup_Size", "")))
        if split and split[0] in key[0]:
            # launches of one shape in time order, phase = index mod K
            ph = seen[key] % split[1]
            seen[key] += 1
            key = (f"[call%{split[1]}={ph}] " + key[0], key[1], key[2])
        d[key].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    print(f"{'kernel':60s} {'grid':>9s} {'block':>5s} {'calls':>6s} {'mean_us':>9s} {'min_us':>8s} {'max_us':>8s}")
    for (n, g, b), v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
        print(f"{_short(n):60s} {g:>9s} {b:>5s} {len(v):6d} {sum(v) / len(v):9.2f} {min(v):8.2f} {max(v):8.2f}")


def counter(path, calib_grid=None):
    rows = list(csv.DictReader(open(path)))
    d = collections.defaultdict(list)
    for r in rows:
        d[(r["Kernel_Name"], r.get("Grid_Size", r.get("Grid_Size_X", "")), r["Counter_Name"])].append(
            float(r["Counter_Value"]))
    scale = {}
    if calib_grid is not None:
        # the calibration copy is the first copy of that grid in the process
        # (later copies of the same grid may move other sizes)
        first = {}
        for r in rows:
            if "copyBuffer" in r["Kernel_Name"] and r.get("Grid_Size", r.get("Grid_Size_X", "")) == str(calib_grid):
                c, i = r["Counter_Name"], int(r.get("Dispatch_Id", 0))
                if c not in first or i < first[c][0]:
                    first[c] = (i, float(r["Counter_Value"]))
        for c, (_, v) in first.items():
            scale[c] = 65536.0 / v
        for c, s in scale.items():
            print(f"calibration {c}: 64 MiB copy reads {65536.0 / s:.1f} KiB -> scale {s:.4f}")
    print(f"{'kernel':60s} {'grid':>9s} {'counter':>11s} {'calls':>6s} {'KiB/disp':>11s} {'scaled':>11s}")
    for (n, g, c), v in sorted(d.items()):
        m = sum(v) / len(v)
        print(f"{_short(n):60s} {g:>9s} {c:>11s} {len(v):6d} {m:11.1f} {m * scale.get(c, 1.0):11.1f}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=["trace", "counter"])
    ap.add_argument("path")
    ap.add_argument("--calib-grid", type=int, default=None)
    ap.add_argument("--split", nargs=2, metavar=("NAME", "K"), default=None,
                    help="trace: split the launches of kernels whose name contains NAME by call index mod K")
    a = ap.parse_args()
    if a.kind == "trace":
        trace(a.path, (a.split[0], int(a.split[1])) if a.split else None)
    else:
        counter(a.path, a.calib_grid)


if __name__ == "__main__":
    main()
