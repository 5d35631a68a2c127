"""Summarise rocprofv3 CSV output per kernel and launch shape.

    python -m partconv_amd.prof_summary trace   <kernel_trace.csv>
    python -m partconv_amd.prof_summary counter <counter_collection.csv> [--calib-grid G]
    python -m partconv_amd.prof_summary step    <FETCH_SIZE csv> <WRITE_SIZE csv> --calib-grid G [--json OUT]

`trace` prints calls / mean / min / max duration (us) per (kernel, grid, block);
--split NAME K splits the launches of one shape of the kernels named NAME by
call index mod K (e.g. the levels of a pass that launches one kernel per level).
`counter` prints the mean counter value (KiB per dispatch for FETCH_SIZE and
WRITE_SIZE) per (kernel, grid, counter); with --calib-grid the first copy kernel
of that grid size (the 64 MiB device copy bench.py --calib-copy issues first)
gives the scale factor 65536 KiB / measured for each counter.
`step` sums the scaled FETCH_SIZE + WRITE_SIZE bytes of every dispatch between
the last two 4 MiB copies of a run (the markers bench.py --workload cfg5
--profile-step puts around one step; told apart from other copies by size), per kernel of the differentiable form:
per level (from the transforms' LOGN template argument, levels ascending) the
filter, forward, gradient and inverse transforms, the product, the dL/dh
partials, reduce and inverse; every other kernel (the fused loss's, torch's) counts as the loss.
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


def _scaled_dispatches(path, calib_grid):
    """[(dispatch id, kernel, grid, scaled bytes)] of a one-counter run, and the scale."""
    rows = list(csv.DictReader(open(path)))
    grid = lambda r: r.get("Grid_Size", r.get("Grid_Size_X", ""))
    rows.sort(key=lambda r: int(r.get("Dispatch_Id", 0)))
    cal = [r for r in rows if "copyBuffer" in r["Kernel_Name"] and grid(r) == str(calib_grid)]
    scale = 65536.0 / float(cal[0]["Counter_Value"])
    return [(int(r["Dispatch_Id"]), r["Kernel_Name"], int(grid(r) or 0),
             1024.0 * float(r["Counter_Value"]) * scale) for r in rows], scale


def _step_kind(name):
    import re
    m = re.search(r"k_off_(i?rfft)_reg<(\d+)", name)
    if m:
        return m.group(1), 1 << int(m.group(2))
    for k in ("k_off_mac_w", "k_off_mac", "k_off_dh_part", "k_off_dh_reduce"):
        if name.startswith(k) or name.startswith("void " + k):
            return k, None
    return "loss", None


def step(fetch_csv, write_csv, calib_grid, out_json=None, commit=None, marker_bytes=4 << 20):
    total, by_kernel, scales = 0.0, {}, {}
    for cname, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        ds, scales[cname] = _scaled_dispatches(path, calib_grid)
        # the markers: copies of marker_bytes (the copy kernel's grid does not
        # depend on the size, so the size tells them from the step's own copies)
        copies = [i for i, (_, n, _, v) in enumerate(ds)
                  if "copyBuffer" in n and abs(v - marker_bytes) < 0.1 * marker_bytes]
        a, b = copies[-2], copies[-1]
        inner = ds[a + 1:b]
        kinds = [(_step_kind(n), g, v) for _, n, g, v in inner]
        levels = sorted({N for (k, N), _, _ in kinds if N})
        big = {}
        for (k, N), g, _ in kinds:
            if N:
                big[(k, N)] = max(big.get((k, N), 0), g)
        seen = collections.Counter()
        for (k, N), g, v in kinds:
            if k == "rfft":
                if g < big[(k, N)]:
                    key = f"filter_transform@{N}"
                else:
                    key = ("forward_transform@" if seen[(k, N)] == 0 else "gradient_transform@") + str(N)
                    seen[(k, N)] += 1
            elif k == "irfft":
                key = (f"inverse_transform@{N}" if g == big[(k, N)] else f"dh_inverse@{N}")
            elif k == "loss":
                key = "loss"
            else:
                i = seen[k]
                seen[k] += 1
                name = {"k_off_mac_w": "product", "k_off_mac": "product", "k_off_dh_part": "dh_partial",
                        "k_off_dh_reduce": "dh_reduce"}[k]
                key = f"{name}@{levels[i]}" if i < len(levels) else f"{name}@?"
            by_kernel[key] = by_kernel.get(key, 0.0) + v
            total += v
        print(f"{cname}: {len(inner)} dispatches between the markers, scale {scales[cname]:.4f}")
    print(f"{'kernel':28s} {'MB (FETCH+WRITE)':>17s}")
    for k in sorted(by_kernel, key=lambda k: -by_kernel[k]):
        print(f"{k:28s} {by_kernel[k] / 1e6:17.1f}")
    print(f"{'step total':28s} {total / 1e6:17.1f}")
    if out_json:
        import json
        with open(out_json, "w") as fh:
            json.dump({"bytes_per_step": total, "by_kernel": by_kernel,
                       "calibration": {k: round(v, 4) for k, v in scales.items()},
                       "source": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, of bench.py --workload cfg5 "
                                 "--profile-step --calib-copy", "commit": commit}, fh, indent=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=["trace", "counter", "step"])
    ap.add_argument("path")
    ap.add_argument("path2", nargs="?", help="step: the WRITE_SIZE csv")
    ap.add_argument("--calib-grid", type=int, default=None)
    ap.add_argument("--json", default=None, help="step: write the per-step sums here")
    ap.add_argument("--commit", default=None, help="step: the build's commit, recorded in the json")
    ap.add_argument("--split", nargs=2, metavar=("NAME", "K"), default=None,
                    help="trace: split the launches of kernels whose name contains NAME by call index mod K")
    a = ap.parse_args()
    if a.kind == "trace":
        trace(a.path, (a.split[0], int(a.split[1])) if a.split else None)
    elif a.kind == "step":
        step(a.path, a.path2, a.calib_grid, a.json, a.commit)
    else:
        counter(a.path, a.calib_grid)


if __name__ == "__main__":
    main()
