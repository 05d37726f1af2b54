"""Writes full_runs.json: whole runs of the oracle at the benchmark's parity sizes, too long
for the oracle to run inside a GPU test, reduced to per-field digests of the final state.

Cases (GAD, Euler 2nd-order remap, Sequential splitting, 4 ghost layers):
- Sedov 1024x1024, minmod, CFL 0.7, run to its end time t = 1 (thousands of cycles);
- Bizarrium 4096x4096, superbee, CFL 0.6, 500 steps;
- circular Sod 8192x8192, minmod, CFL 0.7, 200 steps (the FP64 side of the FP32-mixed run);
- the corner Sedov blast (outside the reference's tests, through the parameter overrides and
  initial fields of tests/conftest.py: blast in the 2x2 cells at the origin corner of [0, 1]^2,
  reflective walls, background internal energy 1e-5) at 1024x1024, minmod, CFL 0.7, run to
  t = 1;
- circular Sod of radius 0.2 (the benchmark list's; the reference's is 0.3) at 8192x8192,
  minmod, CFL 0.7, 200 steps, through the initial fields of tests/conftest.py;
- the same circular Sod set in a vortex of speed 0.5 (tests/conftest.py sod_circ_swirl_fields:
  every cell moving from the first cycle; the benchmark's developed flow) at 8192x8192, minmod,
  CFL 0.7, 200 steps;
- the Bizarrium tube with its left state moving at u = 250 (the benchmark list's; the reference
  moves the right one) at 4096x4096, superbee, CFL 0.6, 500 steps;
- the batched ensemble of the benchmark: 256 domains drawn from seed 2024 by the library's
  domain generator (armon_batch_random_specs, host code), each run alone by the oracle for 40
  cycles with the benchmark's scheme (minmod, CFL 0.7, end times pushed out so that every domain
  runs all 40 cycles). The specs are stored with the digests, so the GPU test also checks that
  the generator still draws the same domains.

For each: the cycle count, the last dt (exact, as float.hex) and, per field, the SHA-256 of
the real cells' FP64 values (row by row, little-endian) with their sum and sum of magnitudes
for diagnostics. Run from the repository root after building the oracle:

    python tests/golden/make_full_runs.py          # all cases, about 65 min on 8 cores
    python tests/golden/make_full_runs.py sedov_1024_to_end_time

The GPU test (tests/test_gpu_golden_runs.py) runs the same parameters through the library and
compares its digests with these."""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "full_runs.json")
FIELDS = ("rho", "umat", "vmat", "Emat", "pmat", "cmat")

_COMMON = dict(scheme=1, projection=2, axis_splitting=0, nb_ghosts=4, verbose=5)
CASES = {
    "sedov_1024_to_end_time": dict(test="Sedov", params=dict(_COMMON, limiter=1, cfl=0.7, nx=1024, ny=1024,
                                                             max_cycles=10**6)),
    "bizarrium_4096_superbee_500_steps": dict(test="Bizarrium", params=dict(_COMMON, limiter=2, cfl=0.6, max_time=1.0,
                                                                            nx=4096, ny=4096, max_cycles=500)),
    "sod_circ_8192_200_steps": dict(test="Sod_circ", params=dict(_COMMON, limiter=1, cfl=0.7, max_time=1e9,
                                                                 nx=8192, ny=8192, max_cycles=200)),
    "sod_circ_r02_8192_200_steps": dict(test="Sod_circ", init="sod_circ_r02",
                                        params=dict(_COMMON, limiter=1, cfl=0.7, max_time=1e9, nx=8192, ny=8192,
                                                    max_cycles=200)),
    "sod_circ_r02_swirl_8192_200_steps": dict(test="Sod_circ", init="sod_circ_r02_swirl",
                                              params=dict(_COMMON, limiter=1, cfl=0.7, max_time=1e9, nx=8192,
                                                          ny=8192, max_cycles=200)),
    "bizarrium_left_moving_4096_superbee_500_steps": dict(
        test="Bizarrium", init="bizarrium_left_moving",
        params=dict(_COMMON, limiter=2, cfl=0.6, max_time=1.0, nx=4096, ny=4096, max_cycles=500)),
    "batch_256_seed2024_40_cycles": dict(kind="batch", seed=2024, domains=256,
                                         params=dict(_COMMON, limiter=1, cfl=0.7, max_cycles=40)),
    "sedov_corner_1024_to_end_time": dict(
        test="Sedov", init="sedov_corner",
        params=dict(_COMMON, limiter=1, cfl=0.7, nx=1024, ny=1024, max_cycles=10**6, max_time=1.0,
                    domain_origin=[0.0, 0.0], domain_size=[1.0, 1.0],
                    bc_factors=[[-1.0, 1.0], [-1.0, 1.0], [1.0, -1.0], [1.0, -1.0]])),
}


def initial_fields(case: dict, p):
    """The fields a case writes after init_test (None: the test's own initial state)."""
    init = case.get("init")
    if init is None:
        return None
    sys.path.insert(0, os.path.join(HERE, ".."))
    from conftest import bizarrium_fields, sedov_corner_fields, sod_circ_radius_fields, sod_circ_swirl_fields

    if init == "sedov_corner":
        return sedov_corner_fields(p)
    if init == "sod_circ_r02":
        return sod_circ_radius_fields(p, 0.2)
    if init == "sod_circ_r02_swirl":
        return sod_circ_swirl_fields(p, 0.2, 0.5)
    if init == "bizarrium_left_moving":
        return bizarrium_fields(p, left_moving=True)
    raise ValueError(init)


STATE_FIELDS = ("high_rho", "low_rho", "high_E", "low_E", "high_u", "low_u", "high_v", "low_v")
BATCH_END_TIME = 1e9   # every domain runs max_cycles cycles


def batch_specs(seed: int, n: int):
    """The ensemble's domain specs from the library's generator (host code of libarmon.so) as
    plain data: test, nx, ny and the high/low states (as float.hex)."""
    sys.path.insert(0, os.path.join(HERE, "..", ".."))
    import __graft_entry__ as g

    armon = g.import_package()
    return [{"test": int(sp.test), "nx": int(sp.nx), "ny": int(sp.ny),
             "state": {f: float(getattr(sp.state, f)).hex() for f in STATE_FIELDS}}
            for sp in armon.random_specs(seed, n)]


def batch_domain_params(po, lib, spec: dict, params: dict):
    """The oracle's parameters for one domain of the ensemble run alone, and its states."""
    p = po.make_params(lib, spec["test"], nx=spec["nx"], ny=spec["ny"], max_time=BATCH_END_TIME, **params)
    st = po.OTestState(**{f: float.fromhex(spec["state"][f]) for f in STATE_FIELDS})
    return p, st


def domain_digest(d, p) -> str:
    """One SHA-256 over the real cells of every field of FIELDS, in that order."""
    h = hashlib.sha256()
    for f in FIELDS:
        h.update(real_cells(d.array(f), p.nx, p.ny, p.row_length, p.index_start).tobytes())
    return h.hexdigest()


def run_batch_case(po, lib, name: str) -> dict:
    case = CASES[name]
    t = time.time()
    specs = batch_specs(case["seed"], case["domains"])
    domains = []
    for k, spec in enumerate(specs):
        p, st = batch_domain_params(po, lib, spec, case["params"])
        status, d, cycles, final_dt, _ = po.run(lib, p, states=st)
        if status != 0:
            raise RuntimeError(f"{name}: domain {k}: oracle status {status}")
        domains.append(dict(spec, cycles=cycles, final_dt=float(final_dt).hex(), sha256=domain_digest(d, p),
                            rho_sum=float(np.sum(real_cells(d.array("rho"), p.nx, p.ny, p.row_length,
                                                            p.index_start)))))
    print(f"{name}: {len(domains)} domains in {time.time() - t:.0f} s", flush=True)
    return {"kind": "batch", "seed": case["seed"], "params": case["params"], "end_time": BATCH_END_TIME,
            "domains": domains}


def real_cells(arr: np.ndarray, nx: int, ny: int, row_length: int, index_start: int) -> np.ndarray:
    """The nx*ny real cells of a field in the reference layout, row by row
    (index(i, j) = index_start + j*row_length + i; ranges.h:44-47)."""
    rows = arr[index_start:index_start + ny * row_length].reshape(ny, row_length)
    return np.ascontiguousarray(rows[:, :nx], dtype="<f8")


def digest(cells: np.ndarray) -> dict:
    return {"sha256": hashlib.sha256(cells.tobytes()).hexdigest(),
            "sum": float(np.sum(cells)), "abs_sum": float(np.sum(np.abs(cells)))}


def run_case(po, lib, name: str) -> dict:
    case = CASES[name]
    if case.get("kind") == "batch":
        return run_batch_case(po, lib, name)
    p = po.make_params(lib, case["test"], **case["params"])
    t = time.time()
    st, d, cycles, final_dt, _ = po.run(lib, p, init_fields=initial_fields(case, p))
    if st != 0:
        raise RuntimeError(f"{name}: oracle status {st}")
    out = {"test": case["test"], "params": case["params"], "cycles": cycles, "final_dt": float(final_dt).hex(),
           "init": case.get("init"),
           "fields": {f: digest(real_cells(d.array(f), p.nx, p.ny, p.row_length, p.index_start)) for f in FIELDS}}
    print(f"{name}: {cycles} cycles in {time.time() - t:.0f} s", flush=True)
    return out


def main(names):
    sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
    import pyoracle as po

    lib = po.load(strict=True)
    data = {"source": "oracle/armon_oracle.c (IEEE build), tests/golden/make_full_runs.py", "cases": {}}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for name in names or list(CASES):
        data["cases"][name] = run_case(po, lib, name)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1:])
