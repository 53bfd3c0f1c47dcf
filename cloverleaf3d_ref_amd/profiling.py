"""Summaries of rocprofv3 output for the hydro kernels.

    python -m cloverleaf3d_ref_amd.profiling stats DIR --command CMD -o OUT.txt [--type double|float] [--last N]
        per-kernel calls / total / average from a --kernel-trace --stats run
        (DIR holds *_kernel_stats.csv and *_kernel_trace.csv), optionally only
        the kernels instantiated for one element type.  With --last N, also
        each marching kernel's average over its last N launches in the trace:
        with N = the run's warmup + timed steps these are the steps' own
        launches, which follow the autotuner's trial launches (the stats
        table counts those trials too, for every candidate shape)

    python -m cloverleaf3d_ref_amd.profiling traffic FETCH_DIR WRITE_DIR --copy-mib M --command CMD -o OUT.json
        HBM bytes per dispatch of every kernel from two --pmc runs (FETCH_SIZE,
        then WRITE_SIZE; DIRs hold *_counter_collection.csv).  Both counters
        are scaled by what they read for the plain device copy of M MiB that
        the profiled command ran first (bench.py --calib-copy-mib M), which
        reads and writes exactly M MiB.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict
from typing import Dict, List, Tuple


# device symbols of template kernels -> the names the library's launch timer uses
_TEMPLATE_NAMES = {"k_fast_mom_march<2": "k_fast_mom_y", "k_fast_mom_march<3": "k_fast_mom_z",
                   "k_fast_cell_march<2": "k_fast_cell_y", "k_fast_cell_march<3": "k_fast_cell_z"}


def _short(name: str) -> str:
    name = name.split("(")[0].replace("void ", "").strip().replace("clover::", "")
    for prefix, short in _TEMPLATE_NAMES.items():
        if name.startswith(prefix):
            return short
    return name.split("<")[0]


def _find(d: str, suffix: str) -> str:
    hits = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return hits[0]


def _elem_type(name: str) -> str:
    """'float' or 'double' for a kernel template instantiated on one, else ''."""
    head = name.split("(")[0]
    if "<" not in head:
        return ""
    args = head.split("<", 1)[1]
    return "float" if "float" in args else "double" if "double" in args else ""


def _keep(name: str, elem: str) -> bool:
    return not elem or _elem_type(name) in ("", elem)


def kernel_stats(d: str, elem: str = "") -> List[Tuple[str, int, float, float, float]]:
    """(kernel, calls, total_us, avg_us, percent of the kept rows), largest
    total first; elem keeps only the kernels instantiated for that type
    (and the untyped ones)."""
    rows = []
    with open(_find(d, "kernel_stats.csv")) as f:
        for r in csv.DictReader(f):
            if _keep(r["Name"], elem):
                rows.append((_short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]) / 1e3,
                             float(r["AverageNs"]) / 1e3))
    total = sum(r[2] for r in rows) or 1.0
    rows = [(n, c, t, a, 100.0 * t / total) for n, c, t, a in rows]
    rows.sort(key=lambda r: -r[2])
    return rows


def last_launches(d: str, n: int, elem: str = "") -> Dict[str, Tuple[int, float]]:
    """kernel -> (launches, avg_us) over its last n launches in the kernel
    trace (fewer if it ran fewer times)."""
    per: Dict[str, List[Tuple[int, float]]] = defaultdict(list)
    with open(_find(d, "kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            if _keep(r["Kernel_Name"], elem):
                t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                per[_short(r["Kernel_Name"])].append((t0, (t1 - t0) / 1e3))
    out = {}
    for k, launches in per.items():
        us = [t for _, t in sorted(launches)[-n:]]
        out[k] = (len(us), sum(us) / len(us))
    return out


def counter_per_dispatch(d: str, counter: str) -> List[Tuple[str, float]]:
    """(kernel, value) for every dispatch in a --pmc run, in dispatch order."""
    out = []
    with open(_find(d, "counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                out.append((r["Kernel_Name"], float(r["Counter_Value"])))
    return out


def traffic(fetch_dir: str, write_dir: str, copy_mib: int) -> Dict:
    known = copy_mib * (1 << 20)
    fetch = counter_per_dispatch(fetch_dir, "FETCH_SIZE")
    write = counter_per_dispatch(write_dir, "WRITE_SIZE")

    def calib(rows, what):
        copies = [v for n, v in rows if "copy" in n.lower()]
        if not copies:
            raise RuntimeError(f"no copy dispatch in the {what} run")
        v = max(copies) * 1024.0
        if not 0.5 * known < v < 2.0 * known:
            raise RuntimeError(f"{what}: largest copy dispatch read {v:.3e} B, expected about {known:.3e} B")
        return known / v, v

    ff, fraw = calib(fetch, "FETCH_SIZE")
    wf, wraw = calib(write, "WRITE_SIZE")
    acc: Dict[str, List[float]] = defaultdict(lambda: [0.0, 0.0, 0])
    for n, v in fetch:
        a = acc[_short(n)]
        a[0] += v * 1024.0 * ff
        a[2] += 1
    wsum: Dict[str, float] = defaultdict(float)
    for n, v in write:
        wsum[_short(n)] += v * 1024.0 * wf
    kernels = {}
    for n, (fb, _, calls) in acc.items():
        if "copy" in n.lower() or "fill" in n.lower():
            continue
        kernels[n] = dict(dispatches=calls, fetch_bytes=fb / calls, write_bytes=wsum[n] / calls,
                          hbm_bytes=(fb + wsum[n]) / calls)
    return dict(copy_mib=copy_mib, fetch_scale=ff, write_scale=wf, copy_fetch_raw_bytes=fraw,
                copy_write_raw_bytes=wraw, kernels=kernels)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m cloverleaf3d_ref_amd.profiling")
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command", required=True)
    s.add_argument("-o", "--out", required=True)
    s.add_argument("--type", choices=("double", "float"), default="",
                   help="only the kernels instantiated for this element type")
    s.add_argument("--last", type=int, default=0, help="also average each marching kernel's last N launches")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--copy-mib", type=int, required=True)
    t.add_argument("--command", required=True)
    t.add_argument("-o", "--out", required=True)
    a = ap.parse_args(argv)
    if a.what == "stats":
        rows = kernel_stats(a.dir, a.type)
        what = f", kernels instantiated for {a.type} (untyped ones count every launch)" if a.type else ""
        with open(a.out, "w") as f:
            f.write(f"# rocprofv3 --kernel-trace --stats{what}\n# command: {a.command}\n")
            f.write("# every launch, the autotuner's trials of each candidate shape included:\n")
            f.write(f"{'kernel':<40} {'calls':>7} {'total_us':>12} {'avg_us':>10} {'pct':>6}\n")
            for name, calls, tot, avg, pct in rows:
                f.write(f"{name:<40} {calls:>7} {tot:>12.1f} {avg:>10.2f} {pct:>6.2f}\n")
            if a.last:
                tail = last_launches(a.dir, a.last, a.type)
                f.write(f"# from the trace: each marching kernel's last {a.last} launches (the steps' own):\n")
                f.write(f"{'kernel':<40} {'calls':>7} {'avg_us':>10}\n")
                for name, calls, tot, avg, pct in rows:
                    if name.startswith("k_fast_") and calls > 1:
                        f.write(f"{name:<40} {tail[name][0]:>7} {tail[name][1]:>10.2f}\n")
    else:
        res = traffic(a.fetch_dir, a.write_dir, a.copy_mib)
        res["command"] = a.command
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1, sort_keys=True)
    sys.stdout.write(open(a.out).read())
    return 0


if __name__ == "__main__":
    sys.exit(main())
