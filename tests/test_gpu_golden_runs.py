"""Whole runs at the benchmark's parity sizes against oracle digests (tests/golden/full_runs.json,
written by tests/golden/make_full_runs.py; the oracle needs minutes for these, so its final
states are kept as per-field SHA-256 digests of the real cells):

- Sedov 1024x1024 to its end time t = 1, fused and staged engines;
- Bizarrium 4096x4096, superbee, 500 steps;
- circular Sod 8192x8192, 200 steps, and beside it the FP32-mixed form, whose error is taken
  against that FP64 state (shown identical to the oracle's by its digests);
- the corner Sedov blast with reflective walls (parameter overrides and initial fields),
  1024x1024 to t = 1, fused and staged engines;
- circular Sod of radius 0.2 at 8192x8192, 200 steps, with the FP32-mixed form beside it;
- the Bizarrium tube with its left state moving, 4096x4096, superbee, 500 steps;
- the benchmark's ensemble of 256 domains (seed 2024), 40 cycles in one batch, every domain
  against the digest of the oracle running it alone, and the FP32-mixed batch beside it.

Identical cycle count, last dt and digests mean every real cell of every field is bit-identical
to the oracle's."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_full_runs import (BATCH_END_TIME, FIELDS, STATE_FIELDS, digest, initial_fields,  # noqa: E402
                            real_cells)

pytestmark = pytest.mark.gpu

with open(os.path.join(HERE, "golden", "full_runs.json")) as _f:
    GOLDEN = json.load(_f)["cases"]

RUNS = [("sedov_1024_to_end_time", 1), ("sedov_1024_to_end_time", 0),
        ("bizarrium_4096_superbee_500_steps", 1), ("sod_circ_8192_200_steps", 1),
        ("sedov_corner_1024_to_end_time", 1), ("sedov_corner_1024_to_end_time", 0),
        ("sod_circ_r02_8192_200_steps", 1), ("sod_circ_r02_swirl_8192_200_steps", 1),
        ("bizarrium_left_moving_4096_superbee_500_steps", 1)]


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
    q = armon.make_params(g["test"], engine=1, precision=1, **g["params"])
    r32, s32 = armon.run(q, initial_fields(g, q))
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


def batch_digest(b, k, p):
    import hashlib

    h = hashlib.sha256()
    for f in FIELDS:
        h.update(real_cells(b.read(k, f), p.nx, p.ny, p.row_length, p.index_start).tobytes())
    return h.hexdigest()


def test_batch_256_domains_match_oracle_digests(armon):
    """The benchmark's ensemble (256 domains from seed 2024, 56 M cells) in one FP64 batch, with
    the autotuner's launch configuration, 40 cycles: the generator draws the stored domains, and
    every domain ends with the cycle count, last dt and state digest of the oracle running it
    alone. Then the FP32-mixed batch: the density's relative L1 error against the FP64 batch is
    within 1e-5 in every domain."""
    g = GOLDEN["batch_256_seed2024_40_cycles"]
    doms = g["domains"]
    specs = armon.random_specs(g["seed"], len(doms))
    for k, (sp, want) in enumerate(zip(specs, doms)):
        assert (int(sp.test), sp.nx, sp.ny) == (want["test"], want["nx"], want["ny"]), k
        assert all(float(getattr(sp.state, f)).hex() == want["state"][f] for f in STATE_FIELDS), k
        sp.max_time = g["end_time"]
    assert g["end_time"] == BATCH_END_TIME
    print(f"case batch 256 domains seed {g['seed']} 40 cycles fp64", flush=True)
    batches = {}
    try:
        for prec in (0, 1):
            common = armon.batch_params(precision=prec, graph_chunk=16, **g["params"])
            b = armon.Batch(common, specs)
            batches[prec] = b
            b.init()
            b.autotune()
            b.begin()
            b.run()
            assert b.finish() == 0, armon.errors()
            if prec == 0:
                bad = []
                for k, want in enumerate(doms):
                    r, p = b.result(k), b.params(k)
                    if (r.cycles, float(r.final_dt).hex()) != (want["cycles"], want["final_dt"]):
                        bad.append(f"domain {k}: cycles {r.cycles} vs {want['cycles']}, dt {r.final_dt!r}")
                    elif batch_digest(b, k, p) != want["sha256"]:
                        rho = float(np.sum(real_cells(b.read(k, "rho"), p.nx, p.ny, p.row_length, p.index_start)))
                        bad.append(f"domain {k} ({p.nx}x{p.ny}): digest differs, rho sum {rho!r} vs "
                                   f"{want['rho_sum']!r}")
                assert not bad, "; ".join(bad[:8])
                print("case batch 256 domains fp32-mixed", flush=True)
        worst = 0.0
        for k in range(len(doms)):
            p = batches[0].params(k)
            a = real_cells(batches[1].read(k, "rho"), p.nx, p.ny, p.row_length, p.index_start)
            ref = real_cells(batches[0].read(k, "rho"), p.nx, p.ny, p.row_length, p.index_start)
            l1 = float(np.sum(np.abs(a - ref)) / np.sum(np.abs(ref)))
            assert np.isfinite(l1) and l1 <= 1e-5, (k, l1)
            worst = max(worst, l1)
        print(f"FP32-mixed batch: worst domain density relative L1 {worst:.2e}", flush=True)
    finally:
        for b in batches.values():
            b.close()
