"""Benchmark of the CloverLeaf3D hydro timestep on one MI355X.

Workload: the two-state cuboid deck at 512^3 cells (domain [0,10]^3,
background rho 0.2 e 1.0, cuboid x in [0,5], y,z in [0,2] at rho 1.0 e 2.5,
gamma 1.4, reflective walls, dt0 0.04, dt rise 1.5), FP64, one tile, every
field resident in HBM, field_summary every 25 steps, as the reference's hydro
loop does it.  A "step" is one pass of hydro.f90:48-72 over the whole mesh.
Setup (mesh, states, halo prime) is outside the timed region, as it is
outside the reference's Wall clock (hydro.f90:44, 90).

Prints one JSON line:
  value        cell-steps/s (cells x timed steps / wall time of the timed steps)
  roofline     the kernel with the most device time in the timed region,
               timed live with HIP events on the chunk's stream: algorithmic
               HBM bytes per launch / average launch time, against 8 TB/s
  cpu_baseline the reference's own kernels (compiled from its sources with
               flang -O3 -fopenmp, oracle/_ref) driven through the same
               hydro loop on this host's cores, on a bounded sample of the
               512^3 workload; also the 64^3 deck on 1 thread and all cores
  fp32_store   the FP32-store form timed the same way, and its accuracy:
               every total against the FP64 form after 300 steps (untimed)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import pathlib
import subprocess
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch of each kernel at the bench workload, from two
# rocprofv3 --pmc runs (read requests and write transactions to memory by
# size), checked against the set-up equation-of-state kernel, whose traffic
# is known exactly (python -m cloverleaf3d_ref_amd.profiling traffic ...)
TRAFFIC_FILE = ROOT / "profiles" / "traffic_cuboid512_fp64.json"

# Algorithmic (kernel-boundary compulsory) traffic of each kernel per launch,
# in FP64 arrays per cell of its domain: every distinct array it reads or
# writes, touched once.  Volume kernels read V plus 1-3 volume fluxes
# depending on the sweep (advec_cell_kernel.f90:97-116, 193-210, 289-306;
# advec_mom_kernel.f90:95-163): 6, 5 and 4 arrays for sweeps 1, 2, 3, i.e. 5
# per launch averaged over a step.
ARRAYS_PER_CELL = {
    # per-call kernels (one per reference loop nest)
    "k_ideal_gas": 4, "k_viscosity": 9, "k_calc_dt": 10, "k_revert": 4, "k_accelerate": 13,
    "k_flux_calc": 12, "k_cell_vols": 5, "k_cell_flux": 6, "k_cell_update": 12, "k_mom_vols": 5,
    "k_node_flux": 2, "k_node_mass": 5, "k_mom_flux": 4, "k_mom_update": 5, "k_reset_field": 10,
    "k_field_summary": 7, "k_pdv<1>": 13, "k_pdv<0>": 16,
    # fused kernels (geometry as scalars).  Sweep-dependent ones averaged over
    # the sweeps they run in: cell_x/y/z read 3/2/1 volume fluxes (sweep
    # 1/2/3), density1 and energy1 and write the sweep's mass flux, density1
    # and energy1; mom_x/y/z read the sweep's mass flux, density1, 2/1/0
    # volume fluxes and three velocities and write three velocities.  x and
    # z run in sweeps 1 and 3, y always in sweep 2; the third sweep also
    # writes pressure (the next step's ideal_gas), one of every two x and z
    # sweeps.  visc_dt reads rho0, p, u0, v0, w0 (it forms the soundspeed
    # from rho0 and p) and writes q; pdv_predict reads rho0, e0, q, u0, v0, w0
    # (it forms the level-0 pressure from rho0 and e0) and writes p.
    # accel_pdv reads rho0, e0, p, q, u0, v0, w0 and writes u1, v1, w1, rho1, e1 and the three volume fluxes
    "k_fast_visc_dt": 6, "k_fast_pdv_predict": 7, "k_fast_accel_pdv": 15,
    "k_fast_cell_x": 7.5, "k_fast_cell_y": 7, "k_fast_cell_z": 7.5,
    "k_fast_mom_x": 9, "k_fast_mom_y": 9, "k_fast_mom_z": 9,
    "k_fast_field_summary": 6, "k_fast_ideal_gas": 4,
}


def kernel_domain_cells(name: str, n: int) -> int:
    """Cells (or nodes) one launch of `name` covers on an n^3 tile."""
    if name in ("k_cell_vols", "k_mom_vols"):
        return (n + 4) ** 3
    if name in ("k_accelerate", "k_reset_field", "k_flux_calc", "k_mom_flux", "k_mom_update", "k_node_flux",
                "k_node_mass", "k_fast_accel_pdv", "k_fast_mom_x", "k_fast_mom_y",
                "k_fast_mom_z"):
        return (n + 1) ** 3
    return n ** 3


def gpu_bench(cells: int, steps: int, warmup: int, summary_frequency: int = 25, fp32: bool = False, tiles: int = 1,
              tune: bool = True):
    from cloverleaf3d_ref_amd import _lib
    from cloverleaf3d_ref_amd.chunk import Chunk
    from cloverleaf3d_ref_amd.deck import cuboid_deck, parse_deck

    deck = parse_deck(cuboid_deck(cells, warmup + steps + 1000, summary_frequency, tiles=tiles))
    t0 = time.perf_counter()
    c = Chunk(deck, fp32=fp32)
    setup_s = time.perf_counter() - t0
    autotune = None
    if tune:  # launch shapes picked on the chunk's own arrays, before the warmup
        t0 = time.perf_counter()
        before, after = c.autotune()
        autotune = dict(seconds=time.perf_counter() - t0, pass_ms_before=before, pass_ms_after=after,
                        shapes=c.shapes())
    L = _lib.lib()
    for _ in range(warmup):
        c.step()
    _lib.check(L.clover_chunk_set_kernel_timing(c._h, 1), "kernel timing")
    c.run(max_steps=0)  # synchronise the stream
    t0 = time.perf_counter()
    rows = c.run(max_steps=steps)  # summaries every summary_frequency steps, as the hydro loop does
    wall = time.perf_counter() - t0
    buf = ctypes.create_string_buffer(1 << 16)
    _lib.check(L.clover_chunk_kernel_times(c._h, buf, len(buf)), "kernel times")
    ktimes = {}
    for line in buf.value.decode().splitlines():
        name, ms, launches = line.split()
        ktimes[name] = (float(ms), int(launches))
    final = c.field_summary()  # after the timed region
    info = dict(setup_s=setup_s, device_bytes=c.device_bytes(), rows=rows, state=c.state(), final=final,
                autotune=autotune)
    c.close()
    return wall, ktimes, info


def dominant_roofline(ktimes, n: int, bytes_per_value: int, traffic_file=None):
    """The kernel with the most device time: algorithmic bytes per launch
    over its average launch time (HIP events on the chunk's stream)."""
    name, (ms, launches) = max(ktimes.items(), key=lambda kv: kv[1][0])
    per_launch_bytes = ARRAYS_PER_CELL.get(name, 0) * bytes_per_value * kernel_domain_cells(name, n)
    avg_s = ms / launches / 1e3
    achieved = per_launch_bytes / avg_s / 1e9 if per_launch_bytes else None
    traffic = None
    if traffic_file is not None and traffic_file.exists():
        measured = json.loads(traffic_file.read_text())["kernels"].get(name)
        traffic = measured["hbm_bytes"] if measured else None
    return dict(bound="hbm", kernel=name, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=(achieved / HBM_PEAK_GBS) if achieved else None, traffic=traffic,
                avg_launch_us=avg_s * 1e6, launches=launches, bytes_per_launch=per_launch_bytes)


def ensemble_bench(ndecks: int, steps: int, fp32: bool = False, warmup: int = 2, tune: bool = True):
    """The batched form: all decks advanced together (graph replays), then
    each deck's final summary.  Setup (mesh and states) and the warmup steps
    that capture the graph of each advect_x parity are outside the timed
    region, which covers `steps` steps of every deck."""
    from cloverleaf3d_ref_amd.deck import ensemble_decks
    from cloverleaf3d_ref_amd.ensemble import Ensemble

    decks = ensemble_decks(ndecks, end_step=warmup + steps)
    t0 = time.perf_counter()
    e = Ensemble(decks, fp32=fp32)
    setup_s = time.perf_counter() - t0
    autotune = None
    if tune:
        t0 = time.perf_counter()
        before, after = e.autotune()
        autotune = dict(seconds=time.perf_counter() - t0, pass_ms_before=before, pass_ms_after=after,
                        shapes=e.shapes(), workgroups=e.workgroups())
    e.run(warmup)
    t0 = time.perf_counter()
    e.run(steps)
    sums = e.summaries()
    wall = time.perf_counter() - t0
    _, _, _, small = e.state()
    out = dict(wall_s=wall, setup_s=setup_s, device_gib=e.device_bytes() / 2 ** 30, small=sum(small), sums=sums,
               autotune=autotune)
    e.close()
    return out, decks


def ensemble_errors(sums32, sums64):
    """Relative difference of every deck's totals, FP32 vs FP64: median,
    99th percentile and maximum per total."""
    import numpy as np
    out = {}
    for k in ("mass", "ie", "ke", "press"):
        e = np.array([abs(a[k] / b[k] - 1) for a, b in zip(sums32, sums64)])
        out[k] = dict(median=float(np.median(e)), p99=float(np.percentile(e, 99)), max=float(e.max()))
    return out


def one_at_a_time(decks, steps: int) -> float:
    """The same decks, each run alone by the plain FP64 form (one deck's
    hydro loop after another, dt back to the host every step as the
    reference driver does); only the steps and final summaries are timed."""
    from cloverleaf3d_ref_amd.chunk import Chunk

    total = 0.0
    for d in decks:
        with Chunk(d) as c:
            c.run(max_steps=0)
            t0 = time.perf_counter()
            for _ in range(steps):
                c.step()
            c.field_summary()
            total += time.perf_counter() - t0
    return total


def hbm_probe(mib: int = 1024, reps: int = 5) -> dict:
    """What plain streaming kernels attain on this device (torch's
    elementwise kernels on FP64 arrays of `mib` MiB): copy (1 read, 1
    write), add (2 reads, 1 write) and fill (1 write), in GB/s.  The
    context for the stencil kernels' roofline fractions, which are
    against the published 8 TB/s."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.ones_like(a)
    c = torch.empty_like(a)
    out = {}
    for name, arrays, op in (("copy", 2, lambda: c.copy_(a)), ("add", 3, lambda: torch.add(a, b, out=c)),
                             ("fill", 1, lambda: c.fill_(1.0))):
        op()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            op()
        e1.record()
        e1.synchronize()
        out[f"{name}_gbs"] = arrays * n * 8 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b, c
    torch.cuda.empty_cache()
    out["arrays_mib"] = mib
    return out


def cpu_run(cells: int, steps: int, threads: int, summary_frequency: int = 0) -> dict:
    """The reference's compiled kernels (oracle/_ref: its kernel sources
    built with flang -O3 -fopenmp) through the hydro loop on the host, timed
    like the reference's Wall clock (hydro loop only, hydro.f90:44, 90).  A
    child process, so that its OpenMP runtime starts with `threads` threads,
    pinned one per core (OMP_PLACES=cores, OMP_PROC_BIND=close)."""
    code = f"""
import sys, json
sys.path[:0] = [{str(ROOT)!r}, {str(ROOT / 'tests')!r}]
import oracle_lib as ol
from cloverleaf3d_ref_amd.deck import cuboid_deck, parse_deck
deck = parse_deck(cuboid_deck({cells}, {steps}, {summary_frequency}))
use_ref = ol.ref_available()
ol.run(deck, use_ref=use_ref)
import ctypes
f = ol.lib().oracle_last_hydro_seconds
f.restype = ctypes.c_double
print(json.dumps(dict(seconds=f(), ref=use_ref)))
"""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="close")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        raise RuntimeError(out.stderr[-2000:])
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return dict(cells=cells, steps=steps, threads=threads, seconds=r["seconds"],
                value=cells ** 3 * steps / r["seconds"], kind="reference" if r["ref"] else "port")


def host_cpu() -> dict:
    """CPU model and core counts from lscpu, and the cores this process may use."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        sockets, cores = int(kv.get("Socket(s)", "1")), int(kv.get("Core(s) per socket", "0"))
        info["physical_cores"] = sockets * cores if cores else None
        info["logical_cpus"] = int(kv.get("CPU(s)", "0")) or None
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    info["usable_cpus"] = len(os.sched_getaffinity(0))
    return info


def cpu_baseline(cells: int, steps: int, threads: int) -> dict:
    """BASELINE.md's CPU plan on this host: the 64^3 deck (100 steps) on one
    thread and on `threads` threads, then a bounded sample of the GPU
    workload (cells^3 for `steps` steps) on `threads` threads, which is the
    reported value."""
    t1 = cpu_run(64, 100, 1)
    tn = cpu_run(64, 100, threads)
    big = cpu_run(cells, steps, threads)
    host = host_cpu()
    return dict(value=big["value"], unit="cell-steps/s", cores=threads, kind=big["kind"],
                sample=f"cuboid deck {cells}^3, {steps} steps, hydro loop only ({big['seconds']:.2f} s); "
                       f"the reference's kernels (flang -O3 -fopenmp) through the reference's hydro loop order",
                cpu_model=host.get("model"), physical_cores=host.get("physical_cores"),
                usable_cpus=host.get("usable_cpus"), omp="OMP_PLACES=cores OMP_PROC_BIND=close",
                config0_64cubed_100_steps={"threads_1": dict(seconds=t1["seconds"], value=t1["value"]),
                                           f"threads_{threads}": dict(seconds=tn["seconds"], value=tn["value"])})


def accuracy_leg(cells: int = 512, steps: int = 300) -> dict:
    """The FP32-store form against the FP64 form on the same deck after
    `steps` steps (untimed, default launch shapes): relative difference of
    every field_summary total."""
    from cloverleaf3d_ref_amd.chunk import Chunk
    from cloverleaf3d_ref_amd.deck import cuboid_deck, parse_deck

    deck = parse_deck(cuboid_deck(cells, steps, 25))
    final = {}
    for fp32 in (False, True):
        with Chunk(deck, fp32=fp32) as c:
            rows = c.run(max_steps=steps)
        final[fp32] = rows[-1]
    f64, f32 = final[False], final[True]
    return dict(workload=f"cuboid deck {cells}^3", after_steps=f32["step"], time_fp64=f64["time"],
                time_fp32=f32["time"],
                rel_err_vs_fp64={k: abs(f32[k] / f64[k] - 1) for k in ("vol", "mass", "ie", "ke", "press")},
                bound=1e-5)


def ensemble_equal_time(ndecks: int, steps: int, dt: float) -> dict:
    """FP32-store ensemble against FP64 at equal simulated time: both forms
    on dt = max_timestep every step (pinned below every deck's stability
    limit), so each deck reaches the same time in both and the totals
    differ by the FP32 storage and arithmetic alone."""
    from cloverleaf3d_ref_amd.deck import ensemble_decks
    from cloverleaf3d_ref_amd.ensemble import Ensemble

    decks = ensemble_decks(ndecks, end_step=steps)
    for d in decks:
        d.dtmax = dt
    out = {}
    for fp32 in (False, True):
        with Ensemble(decks, fp32=fp32) as e:
            e.run(steps)
            out[fp32] = (e.summaries(), e.state())
    (s64, st64), (s32, st32) = out[False], out[True]
    same_time = st64[1] == st32[1] and set(st64[2]) == {dt} and set(st32[2]) == {dt}
    return dict(pinned_dt=dt, steps=steps, same_time_every_deck=same_time,
                rel_err_vs_fp64=ensemble_errors(s32, s64))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=512)
    ap.add_argument("--cpu-cells", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=2)  # about 25 s of the reference on 16 host cores at 512^3
    ap.add_argument("--accuracy-steps", type=int, default=300,
                    help="steps of the untimed FP32-vs-FP64 accuracy leg (0 skips it)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-store run")
    ap.add_argument("--decks", type=int, default=512, help="decks of the batched run (0 skips it)")
    ap.add_argument("--tiled-cells", type=int, default=640, help="edge of the tiled run (0 skips it)")
    ap.add_argument("--tiled-steps", type=int, default=50)
    ap.add_argument("--tiles", type=int, default=8)
    ap.add_argument("--deck-steps", type=int, default=200)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the hydro path here is single-GPU by design "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    probe = hbm_probe()  # before the hydro library opens the device
    tune = not args.no_autotune
    wall, ktimes, info = gpu_bench(args.cells, args.steps, args.warmup, tune=tune)
    n = args.cells
    value = n ** 3 * args.steps / wall
    roofline = dominant_roofline(ktimes, n, 8, TRAFFIC_FILE if n == 512 else None)

    fp32 = None
    if not args.no_fp32:
        wall32, ktimes32, info32 = gpu_bench(args.cells, args.steps, args.warmup, fp32=True, tune=tune)
        value32 = n ** 3 * args.steps / wall32
        f64, f32 = info["final"], info32["final"]
        fp32 = dict(value=value32, unit="cell-steps/s", ms_per_step=wall32 / args.steps * 1e3,
                    speedup_vs_fp64=value32 / value,
                    dtype="f32 storage and stencils; f64 dt minimum, cell mass/energy update, summary sums",
                    rel_err_vs_fp64={k: abs(f32[k] / f64[k] - 1) for k in ("vol", "mass", "ie", "ke", "press")},
                    after_steps=f32["step"],
                    roofline=dominant_roofline(ktimes32, n, 4),
                    kernel_ms_per_step={k: round(v[0] / args.steps, 4) for k, v in
                                        sorted(ktimes32.items(), key=lambda kv: -kv[1][0])},
                    device_gib=info32["device_bytes"] / 2 ** 30, autotune=info32["autotune"])

    tiled = None
    if args.tiled_cells:
        nt = args.tiled_cells
        w64, k64, i64 = gpu_bench(nt, args.tiled_steps, 2, tiles=args.tiles, tune=tune)
        w32, k32, i32 = gpu_bench(nt, args.tiled_steps, 2, tiles=args.tiles, fp32=True, tune=tune)
        f64, f32 = i64["final"], i32["final"]
        tiled = dict(workload=f"cuboid deck {nt}^3 cells as {args.tiles} tiles (3d decomposition), summary every 25",
                     steps=args.tiled_steps,
                     value=nt ** 3 * args.tiled_steps / w64, unit="cell-steps/s", ms_per_step=w64 / args.tiled_steps * 1e3,
                     fp32_value=nt ** 3 * args.tiled_steps / w32, fp32_ms_per_step=w32 / args.tiled_steps * 1e3,
                     fp32_rel_err_vs_fp64={k: abs(f32[k] / f64[k] - 1) for k in ("mass", "ie", "ke", "press")},
                     tile_halo_ms_per_step=round(sum(k64.get(n, (0.0, 0))[0] for n in ("k_tile_copy", "kb_tile_copy"))
                                                 / args.tiled_steps, 4),
                     reflect_ms_per_step=round(sum(k64.get(n, (0.0, 0))[0] for n in ("k_reflect", "kb_tile_reflect"))
                                               / args.tiled_steps, 4),
                     device_gib=i64["device_bytes"] / 2 ** 30, setup_s=i64["setup_s"],
                     autotune=i64["autotune"], fp32_autotune=i32["autotune"])

    batched = None
    if args.decks:
        b64, decks = ensemble_bench(args.decks, args.deck_steps, tune=tune)
        b32, _ = ensemble_bench(args.decks, args.deck_steps, fp32=True, tune=tune)
        alone = one_at_a_time(decks, args.deck_steps)
        ncell = sum(d.x_cells * d.y_cells * d.z_cells for d in decks)
        batched = dict(decks=args.decks, steps_per_deck=args.deck_steps, cells=ncell,
                       workload="seeded ensemble: edges {16..96 step 8}^3, background rho U[0.1,0.5] e U[0.5,1.5], "
                                "one cuboid rho U[0.5,2] e U[1.5,5] over 10-40% of each axis",
                       value=args.decks / b64["wall_s"], unit="decks/s", wall_s=b64["wall_s"],
                       cell_steps_per_s=ncell * args.deck_steps / b64["wall_s"],
                       one_at_a_time=dict(value=args.decks / alone, unit="decks/s", wall_s=alone),
                       speedup_vs_one_at_a_time=alone / b64["wall_s"],
                       fp32_store=dict(value=args.decks / b32["wall_s"], unit="decks/s", wall_s=b32["wall_s"],
                                       equal_time=ensemble_equal_time(args.decks, args.deck_steps, 0.004)),
                       small_dt_decks=b64["small"], setup_s=b64["setup_s"], device_gib=b64["device_gib"],
                       autotune=b64["autotune"], fp32_autotune=b32["autotune"])

    accuracy = None
    if not args.no_fp32 and args.accuracy_steps:
        accuracy = accuracy_leg(args.cells, args.accuracy_steps)
        fp32["accuracy"] = accuracy

    cpu = None
    if not args.no_cpu_baseline:
        # every core this process may use: the GPU's share of the host where
        # the environment sets one (OMP_NUM_THREADS), else the affinity mask
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        cpu = cpu_baseline(args.cpu_cells, args.cpu_steps, threads)

    total_ms = sum(v[0] for v in ktimes.values())
    if roofline.get("achieved"):
        probe["dominant_frac_of_add"] = roofline["achieved"] / probe["add_gbs"]
    out = {
        "metric": "cell-steps/s at 512^3 FP64 vs FP32-store; HBM GB/s; batched decks/s",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"cuboid deck {n}^3 cells, FP64, 1 tile, summary every 25 steps",
                   "cells": n ** 3, "form": "fp64"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp32_store": fp32,
        "batched": batched,
        "tiled": tiled,
        "step_hbm_gbs_kbc": 1192 * value / 1e9,
        "hbm_attainable": probe,
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in
                               sorted(ktimes.items(), key=lambda kv: -kv[1][0])},
        "device_time_fraction": total_ms / (wall * 1e3),
        "setup_s": info["setup_s"],
        "autotune": info["autotune"],
        "device_gib": info["device_bytes"] / 2 ** 30,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
