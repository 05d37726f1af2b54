"""Benchmark: cell-updates/s of the Armon Lagrange + remap step on one MI355X.

Workload: Sedov blast, 8192 x 8192 real cells, 4 ghost layers, GAD + minmod fluxes,
second-order remap, Sequential X/Y splitting, CFL 0.7, perfect gas gamma = 1.4, reference
initial state (blast in the central 2x2 cells of [-1, 1]^2). One step = one full cycle (the
X sweep, the Y sweep and the CFL time step), run by the fused engine with the state resident
in HBM and cycles replayed from a captured hipGraph.

Printed on one JSON line: the FP64 rate as `value`, the FP32-mixed rate beside it, the HBM
roofline of the dominant kernel (timed with HIP events on the solver stream), and the CPU
baseline (the oracle's fast-math build on this host's cores, on a bounded sample).

Beside the headline:
- `developed_flow`: the circular Sod of radius 0.2 at 8192^2, set in a vortex of speed 0.5 so
  that every cell moves, in every direction (the headline Sedov moves only the cells its blast
  has reached), run 200 cycles in FP64 and in FP32-mixed from the same initial state; the error
  of every FP32-mixed field against the FP64 state after those cycles (the FP64 state is the
  oracle's, bit for bit, by the golden digests of the test suite), then the rate of both forms
  timed from that developed state;
- `batched`: the 256-domain ensemble, FP64 and FP32-mixed.
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


def launch_mode(steps, chunk):
    """How armon_solver_advance enqueues `steps` cycles: steps // chunk graph launches of `chunk`
    cycles (one captured graph per parity of the start cycle, captured at begin), then the rest
    one cycle at a time."""
    graphs, eager = (steps // chunk, steps % chunk) if chunk > 0 else (0, steps)
    parts = []
    if graphs:
        parts.append(f"{graphs} hipGraph launch{'es' if graphs > 1 else ''} of {chunk} cycles")
    if eager:
        parts.append(f"{eager} cycle{'s' if eager > 1 else ''} launched kernel by kernel")
    return " + ".join(parts)


def sod_circ_fields(p, radius):
    """Initial rho, umat, vmat, Emat of the circular Sod test (test_cases.h:73-76,
    init_test.cpp:12-22) with radius `radius` about (0.5, 0.5), over every cell of params `p`
    (ghost cells by the same test on their centres, as init_test does)."""
    dx, dy = p.domain_size[0] / p.nx, p.domain_size[1] / p.ny
    g = p.nb_ghosts
    x = p.domain_origin[0] + (np.arange(p.row_length) - g + 0.5) * dx
    y = p.domain_origin[1] + (np.arange(p.col_length) - g + 0.5) * dy
    inside = ((x[None, :] - 0.5) ** 2 + (y[:, None] - 0.5) ** 2 <= radius * radius).reshape(-1)[:p.nb_cells]
    n = p.nb_cells
    return {"rho": np.where(inside, 1.0, 0.125), "umat": np.zeros(n), "vmat": np.zeros(n),
            "Emat": np.where(inside, 2.5, 2.0)}


def swirl_fields(p, radius, speed):
    """The circular Sod state of sod_circ_fields set in a vortex of speed `speed`:
    u = speed sin(pi x) cos(pi y), v = -speed cos(pi x) sin(pi y) over the unit square (zero
    normal to its reflective walls), E raised by the kinetic energy so that the pressure is the
    test's. Every cell moves from the first cycle, in every direction."""
    f = sod_circ_fields(p, radius)
    dx, dy = p.domain_size[0] / p.nx, p.domain_size[1] / p.ny
    g = p.nb_ghosts
    x = p.domain_origin[0] + (np.arange(p.row_length) - g + 0.5) * dx
    y = p.domain_origin[1] + (np.arange(p.col_length) - g + 0.5) * dy
    n = p.nb_cells
    u = (speed * np.sin(np.pi * x)[None, :] * np.cos(np.pi * y)[:, None]).reshape(-1)[:n]
    v = (-speed * np.cos(np.pi * x)[None, :] * np.sin(np.pi * y)[:, None]).reshape(-1)[:n]
    f["umat"], f["vmat"] = u, v
    f["Emat"] = f["Emat"] + 0.5 * (u * u + v * v)
    return f


def field_errors(a, b):
    """Relative L1 (absolute L1 per cell when the reference field is zero everywhere) and
    absolute Linf error of `a` against `b`."""
    d = np.abs(a - b)
    ref = float(np.sum(np.abs(b)))
    out = {"Linf": float(np.max(d))}
    if ref > 0:
        out["rel_L1"] = float(np.sum(d) / ref)
    else:
        out["abs_L1"] = float(np.mean(d))
    return out


def measure_gpu(armon, nx, ny, steps, warmup, precision, profile_cycles, graph_chunk, test="Sedov",
                init_fields=None, develop=0):
    """`warmup` cycles, then exactly `steps` timed cycles (synchronised on both sides), then
    `profile_cycles` cycles timed kernel by kernel. With `develop` > 0 the state is first run
    `develop` cycles (and read back), and the time loop starts again from that state.
    Returns (seconds, kernel times, real cells of ACCURACY_FIELDS: after the `develop` cycles, or
    at the end)."""
    p = armon.make_params(test, nx=nx, ny=ny, max_cycles=10**9, max_time=1e9, verbose=5, engine=1,
                          precision=precision, graph_chunk=graph_chunk, **WORKLOAD)
    s = armon.Solver(p)
    s.init_test()
    for k, v in (init_fields(p) if init_fields else {}).items():
        s.write(k, v)
    lib = armon.load()
    stream = lib.armon_solver_stream(s.h)

    def sync():
        lib.armon_set_stream(stream)
        lib.armon_synchronize()
        lib.armon_set_stream(None)

    def finish(cycles):
        r = armon.ArmonResult()
        st = lib.armon_solver_finish(s.h, C.byref(r))
        if st != 0 or r.cycles != cycles:
            raise RuntimeError(f"run failed: status {st}, cycles {r.cycles} (expected {cycles}), {armon.errors()}")

    state = None
    if develop:
        s.begin()
        s.advance(develop)
        finish(develop)
        state = real_state(s, p)
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
    finish(warmup + steps + profile_cycles)
    if state is None:
        state = real_state(s, p)
    s.close()
    return elapsed, kt, state


def real_state(s, p):
    """The real cells of ACCURACY_FIELDS after armon_solver_finish."""
    g, rl = p.nb_ghosts, p.row_length
    return {f: s.read(f).reshape(-1, rl)[g:g + p.ny, g:g + p.nx].copy() for f in ACCURACY_FIELDS}


def measure_batch(armon, n_domains, seed, steps, warmup, precision, chunk):
    """The seeded ensemble of uneven domains, advanced together (one launch per sweep)."""
    specs = armon.random_specs(seed, n_domains)
    for sp in specs:
        sp.max_time = 1e9   # every domain stays active through the timed cycles
    common = armon.batch_params(max_cycles=10**9, graph_chunk=chunk, precision=precision, verbose=5, **WORKLOAD)
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


def cpu_baseline(nx, ny, cycles, repeats, iteration="2d"):
    """Best of `repeats` runs of `cycles` cycles of the oracle's throughput build iterating as the
    reference's USE_2D_ITER ("2d") or USE_MD_ITER + BALANCE_MD_ITER ("md") build, threads bound
    to cores. Returns (rate, threads, best loop seconds, library file)."""
    # read by the OpenMP runtime when the oracle library loads it (nothing loads it earlier)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    lib = po.load(strict=False, iteration=iteration)
    threads = lib.o_num_threads()
    best = None
    for _ in range(repeats):
        p = po.make_params(lib, "Sedov", nx=nx, ny=ny, max_cycles=cycles, verbose=5, **WORKLOAD)
        st, d, ncyc, _, loop_time = po.run(lib, p)
        del d
        if st != 0 or ncyc != cycles:
            raise RuntimeError("CPU baseline run failed")
        best = loop_time if best is None else min(best, loop_time)
    return nx * ny * cycles / best, threads, best, os.path.basename(po.lib_path(False, iteration))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=8192)
    ap.add_argument("--cpu-cycles", type=int, default=15)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--graph-chunk", type=int, default=-1,
                    help="cycles per graph launch; default: the timed steps (one launch) up to 256, else 16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-developed", action="store_true")
    ap.add_argument("--develop-cycles", type=int, default=200)
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

    # the timed cycles as one graph launch when they are few enough to capture in one graph
    chunk = args.graph_chunk if args.graph_chunk >= 0 else (args.steps if args.steps <= 256 else 16)
    copy_gbps = armon.load().armon_copy_bandwidth(1 << 31, 20)
    log(f"[bench] device copy bandwidth (2 GiB buffers): {copy_gbps:.0f} GB/s")
    log(f"[bench] FP64 fused, Sedov {n}x{n}, warmup {args.warmup}, steps {args.steps} "
        f"({launch_mode(args.steps, chunk)})")
    el64, kt64, _ = measure_gpu(armon, n, n, args.steps, args.warmup, 0, 10, chunk)
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
        log("[bench] FP32-mixed fused, Sedov")
        el32, kt32, _ = measure_gpu(armon, n, n, args.steps, args.warmup, 1, 10, chunk)
        rate32 = cells * args.steps / el32
        fp32 = {"value": rate32, "ms_per_step": el32 / args.steps * 1e3, "speedup_vs_fp64": rate32 / rate64,
                "sweep_ms": {names[k]: kt32.total_ms[k] / max(kt32.launches[k], 1) for k in range(3)}}
        log(f"[bench] FP32-mixed: {rate32 / 1e9:.3f} G cell-updates/s, speedup {rate32 / rate64:.2f}")

    developed = None
    if not args.no_developed:
        # circular Sod of radius 0.2 in a vortex: 200 cycles in both forms from the same state, the
        # FP32-mixed error against the FP64 state, then both rates timed from the developed state
        dev = {}
        states = {}
        for prec, key in ((0, "fp64"), (1, "fp32_mixed")):
            log(f"[bench] developed flow {key}: circular Sod r=0.2 in a vortex {n}x{n}, {args.develop_cycles} "
                f"cycles, then {args.steps} timed")
            el, kt, st = measure_gpu(armon, n, n, args.steps, args.warmup, prec, 10, chunk, test="Sod_circ",
                                     init_fields=lambda p: swirl_fields(p, 0.2, 0.5), develop=args.develop_cycles)
            states[key] = st
            dev[key] = {"value": cells * args.steps / el, "ms_per_step": el / args.steps * 1e3,
                        "sweep_ms": {names[k]: kt.total_ms[k] / max(kt.launches[k], 1) for k in range(3)}}
            log(f"[bench] developed {key}: {dev[key]['value'] / 1e9:.3f} G cell-updates/s")
        errors = {f: field_errors(states["fp32_mixed"][f], states["fp64"][f]) for f in ACCURACY_FIELDS}
        moving = float(np.mean(np.abs(states["fp64"]["umat"]) + np.abs(states["fp64"]["vmat"]) > 1e-6))
        del states
        dev["fp32_mixed"]["speedup_vs_fp64"] = dev["fp32_mixed"]["value"] / dev["fp64"]["value"]
        dev["fp32_mixed"]["errors_vs_fp64"] = errors
        dev["fp64_vs_headline"] = dev["fp64"]["value"] / rate64
        dev["moving_cell_fraction"] = moving
        dev["workload"] = (f"circular Sod, radius 0.2 about (0.5, 0.5), in the vortex u = 0.5 sin(pi x) cos(pi y), "
                           f"v = -0.5 cos(pi x) sin(pi y), {n}x{n}, same scheme and CFL as the "
                           f"headline; FP64 and FP32-mixed each run {args.develop_cycles} cycles from the same "
                           f"initial state, errors taken there; then {args.warmup} warm-up and {args.steps} timed "
                           f"cycles from that state")
        developed = dev
        if fp32 is not None:
            fp32["rho_rel_L1_vs_fp64"] = errors["rho"]["rel_L1"]
            fp32["rho_Linf_vs_fp64"] = errors["rho"]["Linf"]
            fp32["errors_vs_fp64"] = errors
            fp32["errors_from"] = f"developed_flow ({args.develop_cycles} cycles of circular Sod r=0.2 in a vortex)"
        log("[bench] FP32-mixed vs FP64 after the developed run: " +
            ", ".join(f"{f} " + " / ".join(f"{k} {v:.2e}" for k, v in e.items()) for f, e in errors.items()))
        log(f"[bench] developed flow: {moving:.0%} of cells moving; FP64 {dev['fp64_vs_headline']:.3f}x the "
            f"headline rate")

    batched = None
    if not args.no_batch:
        batched = {}
        for prec, key in ((0, "fp64"), (1, "fp32_mixed")):
            log(f"[bench] batched ensemble {key}: {args.batch_domains} domains (seed {args.batch_seed})")
            el, bcells, bkt, tun, tests = measure_batch(armon, args.batch_domains, args.batch_seed, args.steps,
                                                        args.warmup, prec, chunk)
            rate = bcells * args.steps / el
            avg = {names[k]: bkt.total_ms[k] / max(bkt.launches[k], 1) for k in range(3)}
            word = 8 if prec == 0 else 4
            batched[key] = {
                "value": rate, "unit": "cell-updates/s", "ms_per_step": el / args.steps * 1e3,
                "sweep_ms": avg,
                "sweep_GBps": {k: 8 * word * bcells / (avg[k] * 1e-3) / 1e9 for k in ("sweep_x", "sweep_y")},
                "tuning": {"tile_x": tun.tile_x, "x_cells_per_thread": tun.x_cells_per_thread,
                           "y_run_rows": tun.y_run_rows, "grid_x": tun.grid_x, "grid_y": tun.grid_y,
                           "items_x": tun.items_x, "items_y": tun.items_y},
            }
            log(f"[bench] batched {key}: {rate / 1e9:.3f} G cell-updates/s over {bcells} cells, "
                f"{el / args.steps * 1e3:.3f} ms/cycle, tuning tile {tun.tile_x}x{tun.x_cells_per_thread} "
                f"L={tun.y_run_rows} gx={tun.grid_x} "
                f"gy={tun.grid_y}")
        batched["domains"] = args.batch_domains
        batched["cells"] = bcells
        batched["tests"] = {"Sod": tests.count(0), "Sod_circ": tests.count(2), "Sedov": tests.count(4)}
        batched["workload"] = (f"{args.batch_domains} domains from splitmix64 seed {args.batch_seed}: area "
                               f"log-uniform 64^2..1024^2, aspect 1..4, Sod/Sod_circ/Sedov, states x U(0.8, 1.2), "
                               f"same scheme and CFL as the single domain")

    cpu = None
    if not args.no_cpu:
        variants = {}
        for it, what in (("2d", "USE_2D_ITER: rows shared among the threads, a vector loop per row"),
                         ("md", "USE_MD_ITER + BALANCE_MD_ITER: row tiles balanced over the threads")):
            log(f"[bench] CPU baseline: oracle fast-math build ({it} iteration), Sedov {n}x{n}, best of "
                f"{args.cpu_repeats} x {args.cpu_cycles} cycles")
            rate_cpu, threads, secs, libname = cpu_baseline(n, n, args.cpu_cycles, args.cpu_repeats, it)
            variants[it] = {"value": rate_cpu, "best_run_s": secs, "library": libname, "iteration": what}
            log(f"[bench] CPU ({it}): {rate_cpu / 1e6:.1f} M cell-updates/s on {threads} threads")
        best = max(variants, key=lambda k: variants[k]["value"])
        isa = "x86-64-v4" if "_v4" in variants[best]["library"] else "x86-64-v3"
        cpu = {"value": variants[best]["value"], "unit": "cell-updates/s", "cores": threads, "kind": "port",
               "sample": f"Sedov {n}x{n}, best of {args.cpu_repeats} runs of {args.cpu_cycles} cycles (15 per-stage "
                         f"loops per cycle, -O3 -ffast-math -march={isa} OpenMP build of the oracle, threads bound to "
                         f"cores), the faster of the reference's two CPU iteration modes ({best}: "
                         f"{variants[best]['iteration']}), best run {variants[best]['best_run_s']:.1f} s",
               "variants": variants}

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
                               f"splitting, CFL 0.7, fused engine",
                   "cells": cells, "engine": "fused", "timed_launches": launch_mode(args.steps, chunk)},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(names[dom]) if n == 8192 else None, "kernel": names[dom],
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_s * 1e3,
                     "kernel_avg_ms": sweep_avg, "copy_GBps": copy_gbps},
        "fp32_mixed": fp32,
        "developed_flow": developed,
        "batched": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
