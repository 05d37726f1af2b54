"""Whole runs at the benchmark's parity sizes against oracle digests (tests/golden/full_runs.json,
written by tests/golden/make_full_runs.py; the oracle needs minutes for these, so its final
states are kept as per-field SHA-256 digests of the real cells):

- Sedov 1024x1024 to its end time t = 1, fused and staged engines;
- Bizarrium 4096x4096, superbee, 500 steps;
- circular Sod 8192x8192, 200 steps, and beside it the FP32-mixed form, whose error is taken
  against that FP64 state (shown identical to the oracle's by its digests);
- the corner Sedov blast with reflective walls (parameter overrides and initial fields),
  1024x1024 to t = 1, fused and staged engines.

Identical cycle count, last dt and digests mean every real cell of every field is bit-identical
to the oracle's."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_full_runs import FIELDS, digest, initial_fields, real_cells  # noqa: E402

pytestmark = pytest.mark.gpu

with open(os.path.join(HERE, "golden", "full_runs.json")) as _f:
    GOLDEN = json.load(_f)["cases"]

RUNS = [("sedov_1024_to_end_time", 1), ("sedov_1024_to_end_time", 0),
        ("bizarrium_4096_superbee_500_steps", 1), ("sod_circ_8192_200_steps", 1),
        ("sedov_corner_1024_to_end_time", 1), ("sedov_corner_1024_to_end_time", 0)]


def state_cells(s, f):
    p = s.params()
    return real_cells(s.read(f), p.nx, p.ny, p.row_length, p.index_start)


@pytest.mark.parametrize("name,engine", RUNS)
def test_whole_run_matches_oracle_digests(armon, name, engine):
    g = GOLDEN[name]
    print(f"case {name} engine {engine}: {g['test']} {g['params']}", flush=True)
    q = armon.make_params(g["test"], engine=engine, **g["params"])
    r, s = armon.run(q, initial_fields(g, q))
    try:
        assert r.status == 0, armon.errors()
        assert (r.cycles, float(r.final_dt).hex()) == (g["cycles"], g["final_dt"])
        bad = []
        for f in FIELDS:
            got, want = digest(state_cells(s, f)), g["fields"][f]
            if got["sha256"] != want["sha256"]:
                bad.append(f"{f}: sum {got['sum']!r} vs {want['sum']!r}, |sum| {got['abs_sum']!r} vs {want['abs_sum']!r}")
        assert not bad, "; ".join(bad)
        if name.startswith("sod_circ"):
            fp32_mixed_error(armon, g, s)
    finally:
        s.close()


def fp32_mixed_error(armon, g, s64):
    """The FP32-mixed form of the same run: relative L1 and Linf error of each field against
    the FP64 state; the density's relative L1 error must stay within 1e-5."""
    r32, s32 = armon.run(armon.make_params(g["test"], engine=1, precision=1, **g["params"]))
    try:
        assert r32.status == 0 and r32.cycles == g["cycles"], armon.errors()
        errs = {}
        for f in ("rho", "umat", "vmat", "Emat", "pmat"):
            a, b = state_cells(s32, f), state_cells(s64, f)
            errs[f] = (float(np.sum(np.abs(a - b)) / np.sum(np.abs(b))), float(np.max(np.abs(a - b))))
        print("FP32-mixed vs FP64 oracle state (rel L1, Linf):", errs, flush=True)
        assert errs["rho"][0] <= 1e-5
    finally:
        s32.close()
