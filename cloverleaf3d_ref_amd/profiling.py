"""Summaries of rocprofv3 output for the hydro kernels.

    python -m cloverleaf3d_ref_amd.profiling stats DIR --command CMD -o OUT.txt [--type double|float] [--last N]
        per-kernel calls / total / average from a --kernel-trace --stats run
        (DIR holds *_kernel_stats.csv and *_kernel_trace.csv), optionally only
        the kernels instantiated for one element type.  With --last N, also
        each marching kernel's average over its last N launches in the trace:
        with N = the run's warmup + timed steps these are the steps' own
        launches, which follow the autotuner's trial launches (the stats
        table counts those trials too, for every candidate shape)

    python -m cloverleaf3d_ref_amd.profiling traffic READ_DIR WRITE_DIR --last N --cells C --elem-bytes B --command CMD -o OUT.json
        HBM bytes per dispatch of every kernel from two --pmc runs of the same
        command: READ_DIR with TCC_EA0_RDREQ_128B_sum, TCC_EA0_RDREQ_64B_sum
        and TCC_EA0_RDREQ_32B_sum (read requests to memory by size), WRITE_DIR
        with TCC_EA0_WRREQ_sum and TCC_EA0_WRREQ_64B_sum (64- and 32-byte
        write transactions); each kernel averaged over its last N dispatches
        (the timed steps', after the autotuner's trials).  The check: the
        equation-of-state kernel the chunk runs at set-up reads density and
        energy and writes pressure and soundspeed, 2 + 2 arrays of C^3
        elements of B bytes each, so its counters must read exactly that.
        (rocprofv3's FETCH_SIZE takes its 128-byte reads from TCC_BUBBLE,
        which on this GPU sees few of them: it reads about 0.57 of the bytes
        a plain copy moves; the request-size counters are used instead.)
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


READ_COUNTERS = ("TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_32B_sum")
WRITE_COUNTERS = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")
CHECK_KERNEL = "k_fast_ideal_gas"


def dispatch_counters(d: str) -> List[Tuple[str, Dict[str, float]]]:
    """(kernel, {counter: value}) for every dispatch of a --pmc run, in
    dispatch order."""
    per: Dict[int, Tuple[str, Dict[str, float]]] = {}
    with open(_find(d, "counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            i = int(r["Dispatch_Id"])
            if i not in per:
                per[i] = (r["Kernel_Name"], {})
            per[i][1][r["Counter_Name"]] = per[i][1].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return [per[i] for i in sorted(per)]


def _last_avg(rows: List[Tuple[str, float]], last: int) -> Dict[str, Tuple[int, float]]:
    per: Dict[str, List[float]] = defaultdict(list)
    for n, v in rows:
        per[_short(n)].append(v)
    return {k: (len(v[-last:]), sum(v[-last:]) / len(v[-last:])) for k, v in per.items()}


def traffic(read_dir: str, write_dir: str, last: int, cells: int, elem_bytes: int) -> Dict:
    reads, writes = [], []
    for n, c in dispatch_counters(read_dir):
        reads.append((n, 128.0 * c.get(READ_COUNTERS[0], 0.0) + 64.0 * c.get(READ_COUNTERS[1], 0.0) +
                      32.0 * c.get(READ_COUNTERS[2], 0.0)))
    for n, c in dispatch_counters(write_dir):
        w, w64 = c.get(WRITE_COUNTERS[0], 0.0), c.get(WRITE_COUNTERS[1], 0.0)
        writes.append((n, 64.0 * w64 + 32.0 * (w - w64)))
    r, w = _last_avg(reads, last), _last_avg(writes, last)
    kernels = {}
    for k in r:
        if "copy" in k.lower() or "fill" in k.lower():
            continue
        rb, wb = r[k][1], w.get(k, (0, 0.0))[1]
        kernels[k] = dict(dispatches=r[k][0], read_bytes=rb, write_bytes=wb, hbm_bytes=rb + wb)
    out = dict(read_counters=list(READ_COUNTERS), write_counters=list(WRITE_COUNTERS), last=last, kernels=kernels)
    if CHECK_KERNEL in kernels:
        want = 2.0 * cells ** 3 * elem_bytes  # read density, energy; write pressure, soundspeed
        got = kernels[CHECK_KERNEL]
        out["check"] = dict(kernel=CHECK_KERNEL, expected_read_bytes=want, expected_write_bytes=want,
                            read_ratio=got["read_bytes"] / want, write_ratio=got["write_bytes"] / want)
    return out


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
    t.add_argument("read_dir")
    t.add_argument("write_dir")
    t.add_argument("--last", type=int, required=True, help="average each kernel's last N dispatches")
    t.add_argument("--cells", type=int, required=True, help="edge of the profiled cube (the check kernel's size)")
    t.add_argument("--elem-bytes", type=int, default=8)
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
        res = traffic(a.read_dir, a.write_dir, a.last, a.cells, a.elem_bytes)
        res["command"] = a.command
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1, sort_keys=True)
    sys.stdout.write(open(a.out).read())
    return 0


if __name__ == "__main__":
    sys.exit(main())
