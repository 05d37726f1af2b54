"""Benchmark: cell-updates/s of the Armon Lagrange + remap step on one MI355X.

Workload: Sedov blast, 8192 x 8192 real cells, 4 ghost layers, GAD + minmod fluxes,
second-order remap, Sequential X/Y splitting, CFL 0.7, perfect gas gamma = 1.4, reference
initial state (blast in the central 2x2 cells of [-1, 1]^2). One step = one full cycle (the
X sweep, the Y sweep and the CFL time step), run by the fused engine with the state resident
in HBM and cycles replayed from a captured hipGraph.

Printed on one JSON line: the FP64 rate as `value`, the FP32-mixed rate beside it, the HBM
roofline of the dominant kernel (timed with HIP events on the solver stream), and the CPU
baseline (the oracle's fast-math build on this host's cores, on a bounded sample).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK = 8.0e12           # MI355X HBM3E, bytes/s
WORKLOAD = dict(scheme=1, projection=2, limiter=1, axis_splitting=0, nb_ghosts=4, cfl=0.7)
ACCURACY_FIELDS = ("rho", "umat", "vmat", "pmat", "Emat")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of a FP64 sweep at 8192^2, from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for a copy of known size)."""
    path = os.path.join(ROOT, "profiles", "sweeps_sedov8192_hbm_traffic_valu.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                vals = dict(kv.split("=") for kv in line.split()[1:])
                return float(vals[kernel])
    except (OSError, KeyError, ValueError):
        pass
    return None


def measure_gpu(armon, nx, ny, steps, warmup, precision, profile_cycles, graph_chunk):
    p = armon.make_params("Sedov", nx=nx, ny=ny, max_cycles=10**9, max_time=1e9, verbose=5, engine=1,
                          precision=precision, graph_chunk=graph_chunk, **WORKLOAD)
    s = armon.Solver(p)
    s.init_test()
    lib = armon.load()
    stream = lib.armon_solver_stream(s.h)

    def sync():
        lib.armon_set_stream(stream)
        lib.armon_synchronize()
        lib.armon_set_stream(None)

    s.begin()
    s.advance(warmup)
    sync()
    # timed region: exactly `steps` cycles, synchronised on both sides
    t0 = time.perf_counter()
    s.advance(steps)
    sync()
    elapsed = time.perf_counter() - t0
    # per-kernel durations: `profile_cycles` more cycles launched one kernel at a time between
    # HIP events on the solver stream, straight after the timed region (an idle device would
    # first lower its clocks and the durations would not be those of the timed cycles)
    kt = s.kernel_times(profile_cycles)
    r = armon.ArmonResult()
    st = lib.armon_solver_finish(s.h, C.byref(r))
    if st != 0 or r.cycles != warmup + steps + profile_cycles:
        raise RuntimeError(f"run failed: status {st}, cycles {r.cycles}, {armon.errors()}")
    g, rl = p.nb_ghosts, p.row_length
    state = {f: s.read(f).reshape(-1, rl)[g:g + ny, g:g + nx].copy() for f in ACCURACY_FIELDS}   # real cells
    s.close()
    return elapsed, kt, state


def measure_batch(armon, n_domains, seed, steps, warmup, precision):
    """The seeded ensemble of uneven domains, advanced together (one launch per sweep)."""
    specs = armon.random_specs(seed, n_domains)
    for sp in specs:
        sp.max_time = 1e9   # every domain stays active through the timed cycles
    common = armon.batch_params(max_cycles=10**9, graph_chunk=16, precision=precision, verbose=5, **WORKLOAD)
    b = armon.Batch(common, specs)
    b.init()
    tuning = b.autotune()
    b.begin()
    lib = armon.load()
    stream = b.stream()

    def sync():
        lib.armon_set_stream(stream)
        lib.armon_synchronize()
        lib.armon_set_stream(None)

    b.advance(warmup)
    sync()
    t0 = time.perf_counter()
    b.advance(steps)
    sync()
    elapsed = time.perf_counter() - t0
    kt = b.kernel_times(4)
    if b.finish() != 0:
        raise RuntimeError(f"batch run failed: {armon.errors()}")
    cells = b.cells()
    tests = [int(sp.test) for sp in specs]
    b.close()
    return elapsed, cells, kt, tuning, tests


def cpu_baseline(nx, ny, cycles, repeats):
    """Best of `repeats` runs of `cycles` cycles of the oracle's throughput build, threads bound
    to cores. Returns (rate, threads, best loop seconds, library file)."""
    # read by the OpenMP runtime when the oracle library loads it (nothing loads it earlier)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    lib = po.load(strict=False)
    threads = lib.o_num_threads()
    best = None
    for _ in range(repeats):
        p = po.make_params(lib, "Sedov", nx=nx, ny=ny, max_cycles=cycles, verbose=5, **WORKLOAD)
        st, d, ncyc, _, loop_time = po.run(lib, p)
        del d
        if st != 0 or ncyc != cycles:
            raise RuntimeError("CPU baseline run failed")
        best = loop_time if best is None else min(best, loop_time)
    return nx * ny * cycles / best, threads, best, os.path.basename(po.lib_path(False))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=8192)
    ap.add_argument("--cpu-cycles", type=int, default=15)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--graph-chunk", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-domains", type=int, default=256)
    ap.add_argument("--batch-seed", type=int, default=2024)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    sys.path.insert(0, ROOT)
    import __graft_entry__ as g

    armon = g.import_package()
    armon.load()
    n = args.cells
    cells = n * n

    copy_gbps = armon.load().armon_copy_bandwidth(1 << 31, 20)
    log(f"[bench] device copy bandwidth (2 GiB buffers): {copy_gbps:.0f} GB/s")
    log(f"[bench] FP64 fused, Sedov {n}x{n}, warmup {args.warmup}, steps {args.steps}")
    el64, kt64, st64 = measure_gpu(armon, n, n, args.steps, args.warmup, 0, 10, args.graph_chunk)
    rate64 = cells * args.steps / el64
    log(f"[bench] FP64: {rate64 / 1e9:.3f} G cell-updates/s, {el64 / args.steps * 1e3:.3f} ms/cycle")

    # dominant kernel and its roofline (bytes per launch: read 4 + write 4 FP64 state words per cell)
    names = ["sweep_x", "sweep_y", "time_step"]
    dom = max(range(2), key=lambda k: kt64.total_ms[k])
    avg_s = kt64.total_ms[dom] / kt64.launches[dom] / 1e3
    bytes_per_launch = 64 * cells
    achieved = bytes_per_launch / avg_s
    sweep_avg = {names[k]: kt64.total_ms[k] / max(kt64.launches[k], 1) for k in range(3)}

    fp32 = None
    if not args.no_fp32:
        log("[bench] FP32-mixed fused")
        el32, kt32, st32 = measure_gpu(armon, n, n, args.steps, args.warmup, 1, 10, args.graph_chunk)
        rate32 = cells * args.steps / el32
        # relative L1 and absolute Linf error of each field against the FP64 run, same cycles
        errors = {}
        for f in ACCURACY_FIELDS:
            d = np.abs(st32[f] - st64[f])
            errors[f] = {"rel_L1": float(np.sum(d) / np.sum(np.abs(st64[f]))), "Linf": float(np.max(d))}
        del st32, st64
        l1, linf = errors["rho"]["rel_L1"], errors["rho"]["Linf"]
        fp32 = {"value": rate32, "ms_per_step": el32 / args.steps * 1e3, "speedup_vs_fp64": rate32 / rate64,
                "rho_rel_L1_vs_fp64": l1, "rho_Linf_vs_fp64": linf, "errors_vs_fp64": errors,
                "cycles_compared": args.warmup + args.steps + 10,
                "sweep_ms": {names[k]: kt32.total_ms[k] / max(kt32.launches[k], 1) for k in range(3)}}
        log(f"[bench] FP32-mixed: {rate32 / 1e9:.3f} G cell-updates/s, speedup {rate32 / rate64:.2f}, "
            f"rho L1 {l1:.2e}")
        log("[bench] FP32-mixed vs FP64 (rel L1 / Linf): " +
            ", ".join(f"{f} {e['rel_L1']:.2e} / {e['Linf']:.2e}" for f, e in errors.items()))

    batched = None
    if not args.no_batch:
        batched = {}
        for prec, key in ((0, "fp64"), (1, "fp32_mixed")):
            log(f"[bench] batched ensemble {key}: {args.batch_domains} domains (seed {args.batch_seed})")
            el, bcells, bkt, tun, tests = measure_batch(armon, args.batch_domains, args.batch_seed, args.steps,
                                                        args.warmup, prec)
            rate = bcells * args.steps / el
            avg = {names[k]: bkt.total_ms[k] / max(bkt.launches[k], 1) for k in range(3)}
            word = 8 if prec == 0 else 4
            batched[key] = {
                "value": rate, "unit": "cell-updates/s", "ms_per_step": el / args.steps * 1e3,
                "sweep_ms": avg,
                "sweep_GBps": {k: 8 * word * bcells / (avg[k] * 1e-3) / 1e9 for k in ("sweep_x", "sweep_y")},
                "tuning": {"y_run_rows": tun.y_run_rows, "grid_x": tun.grid_x, "grid_y": tun.grid_y,
                           "items_x": tun.items_x, "items_y": tun.items_y},
            }
            log(f"[bench] batched {key}: {rate / 1e9:.3f} G cell-updates/s over {bcells} cells, "
                f"{el / args.steps * 1e3:.3f} ms/cycle, tuning L={tun.y_run_rows} gx={tun.grid_x} gy={tun.grid_y}")
        batched["domains"] = args.batch_domains
        batched["cells"] = bcells
        batched["tests"] = {"Sod": tests.count(0), "Sod_circ": tests.count(2), "Sedov": tests.count(4)}
        batched["workload"] = (f"{args.batch_domains} domains from splitmix64 seed {args.batch_seed}: area "
                               f"log-uniform 64^2..1024^2, aspect 1..4, Sod/Sod_circ/Sedov, states x U(0.8, 1.2), "
                               f"same scheme and CFL as the single domain")

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: oracle fast-math build, Sedov {n}x{n}, best of {args.cpu_repeats} x "
            f"{args.cpu_cycles} cycles")
        rate_cpu, threads, secs, libname = cpu_baseline(n, n, args.cpu_cycles, args.cpu_repeats)
        isa = "x86-64-v4" if "_v4" in libname else "x86-64-v3"
        cpu = {"value": rate_cpu, "unit": "cell-updates/s", "cores": threads, "kind": "port",
               "sample": f"Sedov {n}x{n}, best of {args.cpu_repeats} runs of {args.cpu_cycles} cycles (15 per-stage "
                         f"loops per cycle, -O3 -ffast-math -march={isa} OpenMP build of the oracle, row-by-row 2D "
                         f"iteration, threads bound to cores), best run {secs:.1f} s"}
        log(f"[bench] CPU: {rate_cpu / 1e6:.1f} M cell-updates/s on {threads} threads")

    out = {
        "metric": "cell-updates/s at 8192^2 Sedov (FP64 fused; FP32-mixed beside it)",
        "value": rate64,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el64 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (reference Sedov initial state)",
        "config": {"workload": f"Sedov {n}x{n}, 4 ghosts, GAD+minmod, Euler 2nd order remap, Sequential "
                               f"splitting, CFL 0.7, fused engine, hipGraph of {args.graph_chunk} cycles",
                   "cells": cells, "engine": "fused"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(names[dom]) if n == 8192 else None, "kernel": names[dom],
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_s * 1e3,
                     "kernel_avg_ms": sweep_avg, "copy_GBps": copy_gbps},
        "fp32_mixed": fp32,
        "batched": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
