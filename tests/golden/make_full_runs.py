"""Writes full_runs.json: whole runs of the oracle at the benchmark's parity sizes, too long
for the oracle to run inside a GPU test, reduced to per-field digests of the final state.

Cases (GAD, Euler 2nd-order remap, Sequential splitting, 4 ghost layers):
- Sedov 1024x1024, minmod, CFL 0.7, run to its end time t = 1 (thousands of cycles);
- Bizarrium 4096x4096, superbee, CFL 0.6, 500 steps;
- circular Sod 8192x8192, minmod, CFL 0.7, 200 steps (the FP64 side of the FP32-mixed run);
- the corner Sedov blast (outside the reference's tests, through the parameter overrides and
  initial fields of tests/conftest.py: blast in the 2x2 cells at the origin corner of [0, 1]^2,
  reflective walls, background internal energy 1e-5) at 1024x1024, minmod, CFL 0.7, run to
  t = 1.

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
    "sedov_corner_1024_to_end_time": dict(
        test="Sedov", init="sedov_corner",
        params=dict(_COMMON, limiter=1, cfl=0.7, nx=1024, ny=1024, max_cycles=10**6, max_time=1.0,
                    domain_origin=[0.0, 0.0], domain_size=[1.0, 1.0],
                    bc_factors=[[-1.0, 1.0], [-1.0, 1.0], [1.0, -1.0], [1.0, -1.0]])),
}


def initial_fields(case: dict, p):
    """The fields a case writes after init_test (None: the test's own initial state)."""
    if case.get("init") != "sedov_corner":
        return None
    sys.path.insert(0, os.path.join(HERE, ".."))
    from conftest import sedov_corner_fields

    return sedov_corner_fields(p)


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
