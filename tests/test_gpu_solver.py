"""Whole-run parity of the solver (staged drop-in engine and fused engine) against the oracle.

Both engines compute in FP64 without FMA contraction, in the reference's operation order, so
the final state, the cycle count and the final time step must match the oracle bit for bit.
The reference's own whole-run known answers (cycle counts of tests/armon_2D_tests.cpp:238-242,
conservation to 1e-12) are checked on the GPU as well.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import reference_cycle_counts

pytestmark = pytest.mark.gpu

STATE = ("rho", "umat", "vmat", "Emat", "pmat", "cmat")
ENGINES = {"staged": 0, "fused": 1}


def oracle_run(po, oracle, test, **cfg):
    p = po.make_params(oracle, test, **cfg)
    st, d, cycles, final_dt, _ = po.run(oracle, p)
    return st, d, cycles, final_dt


def product_run(armon, test, engine, init=None, **cfg):
    p = armon.make_params(test, engine=engine, **cfg)
    r, s = armon.run(p, init)
    return r, s


def report(name, a, b):
    d = np.abs(a - b)
    rel = d / np.maximum(np.abs(b), 1e-300)
    return f"{name}: {int(np.count_nonzero(a != b))} differ, max abs {d.max():.3e}, max rel {rel.max():.3e}"


def assert_same_state(s, od, fields=STATE):
    mask = od.array("domain_mask") == 1
    msgs = []
    for f in fields:
        a, b = s.read(f)[mask], od.array(f)[mask]
        if not np.array_equal(a.view(np.uint64), b.view(np.uint64)):
            msgs.append(report(f, a, b))
    assert not msgs, "; ".join(msgs)


def check(armon, po, oracle, test, engine, fields=STATE, **cfg):
    st, od, cycles, final_dt = oracle_run(po, oracle, test, **cfg)
    assert st == 0
    print(f"case {test} engine={engine} {cfg}", flush=True)
    r, s = product_run(armon, test, engine, **cfg)
    assert r.status == 0, armon.errors()
    assert r.cycles == cycles
    assert r.final_dt == final_dt, (r.final_dt, final_dt)
    assert_same_state(s, od, fields)
    s.close()
    return cycles


REF = dict(scheme=1, projection=2, limiter=1, axis_splitting=0, nb_ghosts=5, nx=100, ny=100, max_cycles=1000,
           verbose=5)


@pytest.mark.parametrize("engine", ["fused", "staged"])
@pytest.mark.parametrize("test,cycles", reference_cycle_counts())
def test_reference_cycle_counts(armon, po, oracle, engine, test, cycles):
    """tests/armon_2D_tests.cpp:238-242 with the reference parameters (reference.cpp:33-64)."""
    assert check(armon, po, oracle, test, ENGINES[engine], **REF) == cycles


@pytest.mark.parametrize("engine", ["fused", "staged"])
@pytest.mark.parametrize("test,nx,ny", [("Sedov", 4, 300), ("Sedov", 300, 4), ("Sod_circ", 257, 5),
                                        ("Sod_y", 5, 257), ("Bizarrium", 1023, 4)])
def test_thin_and_ragged_domains(armon, po, oracle, engine, test, nx, ny):
    """Domains as thin as the stencil's read extent (4 cells for GAD + 2nd-order remap) and sizes
    that fill no tile or column group evenly: X tiles spanning several rows of the padded-row
    stream, Y runs over partial groups of 64 columns."""
    cfg = dict(scheme=1, projection=2, limiter=1, axis_splitting=0, nb_ghosts=4, nx=nx, ny=ny, max_cycles=60,
               verbose=5)
    assert check(armon, po, oracle, test, ENGINES[engine], **cfg) > 0


@pytest.mark.parametrize("test,nx,ny", [("Sod", 1, 40), ("Sod_y", 40, 1), ("Sod", 1, 1)])
def test_single_cell_wide_domains_staged(armon, po, oracle, test, nx, ny):
    """One-cell-wide domains, below the fused engine's minimum width, through the staged engine:
    every stage's range degenerates to one row or column."""
    cfg = dict(scheme=0, projection=1, axis_splitting=0, nb_ghosts=2, nx=nx, ny=ny, max_cycles=30, verbose=5)
    assert check(armon, po, oracle, test, ENGINES["staged"], **cfg) > 0


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_sod_first_order(armon, po, oracle, engine):
    """Sod 100x100, Godunov + first-order remap, CFL 0.6, to t = 0.2, default 3 ghosts and a
    stencil width of 1: the second ghost layer is read but never refilled (stale values)."""
    cfg = dict(scheme=0, projection=1, cfl=0.6, max_time=0.2, nx=100, ny=100, max_cycles=100000, verbose=5)
    check(armon, po, oracle, "Sod", ENGINES[engine], **cfg)


@pytest.mark.parametrize("scheme,proj,limiter", [(0, 1, 1), (0, 2, 1), (1, 1, 1), (1, 2, 0), (1, 2, 2)])
def test_fused_schemes(armon, po, oracle, scheme, proj, limiter):
    cfg = dict(scheme=scheme, projection=proj, limiter=limiter, nb_ghosts=4, nx=64, ny=48, max_cycles=40,
               verbose=5)
    check(armon, po, oracle, "Sedov", 1, **cfg)


@pytest.mark.parametrize("split", [1, 2, 3, 4])
@pytest.mark.parametrize("test", ["Sedov", "Sod_circ"])
def test_fused_splittings(armon, po, oracle, split, test):
    cfg = dict(scheme=1, projection=2, limiter=1, axis_splitting=split, nb_ghosts=4, nx=50, ny=70, max_cycles=30,
               verbose=5)
    check(armon, po, oracle, test, 1, **cfg)


Y_SWEEP_CASES = [
    ("Sedov", dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=64, ny=150)),
    ("Sedov", dict(scheme=0, projection=1, limiter=1, nb_ghosts=4, nx=70, ny=64)),
    ("Sedov", dict(scheme=0, projection=2, nb_ghosts=3, nx=40, ny=36)),          # stale 3rd ghost layer
    ("Bizarrium", dict(scheme=1, projection=2, limiter=2, nb_ghosts=4, nx=96, ny=64)),
    ("Sod_circ", dict(scheme=1, projection=2, limiter=1, axis_splitting=1, nb_ghosts=4, nx=50, ny=70)),
    ("Sod_y", dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=5, ny=257)),
    ("Sedov", dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=300, ny=4)),
]


@pytest.mark.parametrize("y_run_rows", [1, 3, 1000])
@pytest.mark.parametrize("test,cfg", Y_SWEEP_CASES)
def test_fused_y_run_lengths(armon, po, oracle, y_run_rows, test, cfg):
    """The fused Y sweep with marching runs of any length (ArmonParams.y_run_rows) gives the
    oracle's bits: the run length changes how the work is split, never the arithmetic."""
    cfg = dict(cfg, max_cycles=30, verbose=5)
    st, od, cycles, final_dt = oracle_run(po, oracle, test, **cfg)
    assert st == 0
    print(f"case {test} y_run_rows={y_run_rows} {cfg}", flush=True)
    r, s = product_run(armon, test, 1, y_run_rows=y_run_rows, **cfg)
    assert r.status == 0, armon.errors()
    assert (r.cycles, r.final_dt) == (cycles, final_dt)
    assert_same_state(s, od)
    s.close()


@pytest.mark.parametrize("limiter", [1, 2])
def test_fused_bizarrium(armon, po, oracle, limiter):
    cfg = dict(scheme=1, projection=2, limiter=limiter, nb_ghosts=4, nx=96, ny=64, max_cycles=60, verbose=5)
    check(armon, po, oracle, "Bizarrium", 1, **cfg)


def test_fused_stencil_and_ghosts(armon, po, oracle):
    """More ghost layers than the stencil reads, and a stencil wider than the minimum."""
    cfg = dict(scheme=1, projection=1, nb_ghosts=6, stencil_width=5, nx=40, ny=36, max_cycles=30, verbose=5)
    check(armon, po, oracle, "Sod_circ", 1, **cfg)
    cfg = dict(scheme=0, projection=2, nb_ghosts=3, nx=40, ny=36, max_cycles=30, verbose=5)   # stale 3rd layer
    check(armon, po, oracle, "Sedov", 1, **cfg)


@pytest.mark.parametrize("chunk", [0, 2, 6])
def test_fused_graph_chunks(armon, po, oracle, chunk):
    cfg = dict(scheme=1, projection=2, nb_ghosts=4, nx=64, ny=64, max_cycles=23, verbose=5)
    st, od, cycles, final_dt = oracle_run(po, oracle, "Sedov", **cfg)
    r, s = product_run(armon, "Sedov", 1, graph_chunk=chunk, **cfg)
    assert r.status == 0 and r.cycles == cycles == 23 and r.final_dt == final_dt
    assert_same_state(s, od)


def test_fused_time_bound(armon, po, oracle):
    """The loop stops on max_time with the same cycle count (device-side condition)."""
    cfg = dict(scheme=1, projection=2, nb_ghosts=4, nx=64, ny=64, max_cycles=100000, max_time=0.05, verbose=5)
    check(armon, po, oracle, "Sedov", 1, **cfg)


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_constant_dt(armon, po, oracle, engine):
    cfg = dict(scheme=1, projection=2, nb_ghosts=4, nx=48, ny=48, max_cycles=20, cst_dt=1, Dt=1e-3, verbose=5)
    check(armon, po, oracle, "Sod_circ", ENGINES[engine], **cfg)


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_dt_on_even_cycles(armon, po, oracle, engine):
    cfg = dict(scheme=1, projection=2, nb_ghosts=4, nx=48, ny=48, max_cycles=20, dt_on_even_cycles=1, verbose=5)
    check(armon, po, oracle, "Sedov", ENGINES[engine], **cfg)


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_invalid_time_step(armon, po, oracle, engine):
    """Zero sound speed and velocity everywhere: dt_CFL returns DBL_MAX and CFL 2 overflows it
    to +inf, an invalid time step (armon_2D.cpp:238-242) at cycle 0."""
    cfg = dict(scheme=1, projection=2, nb_ghosts=4, nx=32, ny=32, max_cycles=5, cfl=2.0, verbose=5)
    p = po.make_params(oracle, "Sod", **cfg)
    d = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(d.c), False)
    for f in ("Emat", "umat", "vmat"):
        d.array(f)[:] = 0.0
    lt, fdt, cyc = C.c_double(), C.c_double(), C.c_int()
    assert oracle.o_time_loop(C.byref(p), C.byref(d.c), C.byref(lt), C.byref(fdt), C.byref(cyc)) == 1
    q = armon.make_params("Sod", engine=ENGINES[engine], **cfg)
    s = armon.Solver(q)
    s.init_test()
    z = np.zeros(q.nb_cells)
    for f in ("Emat", "umat", "vmat"):
        s.write(f, z)
    r = s.time_loop()
    errs = armon.errors()
    assert r.status == 1
    assert any("Invalid dt at cycle 0" in m for _, m in errs), errs


def test_fused_conservation(armon):
    """armon_2D_tests.cpp:246-275 on the fused engine."""
    for test in ("Sod", "Sod_y", "Sod_circ", "Sedov"):
        p = armon.reference_params(test, max_time=10000.0, max_cycles=400, engine=1)
        s = armon.Solver(p)
        s.init_test()
        m0, e0 = C.c_double(), C.c_double()
        lib = armon.load()
        s.begin()
        lib.armon_solver_conservation(s.h, C.byref(m0), C.byref(e0))
        s.advance(400)
        r = s.finish()
        m1, e1 = C.c_double(), C.c_double()
        lib.armon_solver_conservation(s.h, C.byref(m1), C.byref(e1))
        assert r.status == 0 and r.cycles == 400
        assert abs(m1.value - m0.value) <= 1e-12 * abs(m0.value), test
        assert abs(e1.value - e0.value) <= 1e-12 * abs(e0.value), test
        s.close()


def test_sedov_1024_fused(armon, po, oracle):
    """Sedov 1024x1024 (GAD + minmod, second-order remap, CFL 0.7) for 60 cycles, bit for bit."""
    cfg = dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=1024, ny=1024, max_cycles=60, verbose=5)
    check(armon, po, oracle, "Sedov", 1, **cfg)


@pytest.mark.parametrize("test", ["Sedov", "Sod_circ", "Bizarrium"])
def test_fp32_mixed_accuracy(armon, test):
    """FP32 storage and flux arithmetic against the FP64 engine after 100 cycles: relative L1
    error of the density <= 1e-5 (and the other fields reported). For Bizarrium the EOS runs in
    FP64 on the FP32 state, whose `e - eps_k0` and `p - pk0` terms cancel to a few digits."""
    cfg = dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=256, ny=256, max_cycles=100, max_time=1.0,
               verbose=5)
    r64, s64 = product_run(armon, test, 1, **cfg)
    r32, s32 = product_run(armon, test, 1, precision=1, **cfg)
    assert r64.status == 0 and r32.status == 0 and r32.cycles == r64.cycles == 100
    p = s64.params()
    mask = s64.read("domain_mask") == 1
    out = {}
    for f in ("rho", "umat", "vmat", "Emat", "pmat"):
        a, b = s32.read(f)[mask], s64.read(f)[mask]
        ref = np.sum(np.abs(b))
        # relative L1, or the absolute L1 per cell for a field that is zero everywhere (v of the
        # 1D tubes)
        l1 = np.sum(np.abs(a - b)) / ref if ref > 0 else np.mean(np.abs(a - b))
        out[f] = (float(l1), float(np.max(np.abs(a - b))))
    print(test, {k: f"L1={v[0]:.2e} Linf={v[1]:.2e}" for k, v in out.items()}, file=sys.stderr)
    assert all(np.isfinite(v).all() for v in out.values()), out
    assert out["rho"][0] <= 1e-5
    assert p.nx == 256


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_sedov_corner_overrides(armon, po, oracle, engine):
    """A configuration outside the reference's tests, given through the explicit overrides (origin,
    size, per-side boundary factors, initial fields): the Sedov blast in the 2x2 cells at the
    origin corner of [0, 1]^2, reflective walls, background internal energy 1e-5 (GAD + minmod,
    second-order remap, CFL 0.7), 256x256 for 300 cycles, bit for bit against the oracle given the
    same overrides."""
    from conftest import SEDOV_CORNER, sedov_corner_fields

    cfg = dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=256, ny=256, cfl=0.7, max_cycles=300,
               max_time=1.0, verbose=5, **SEDOV_CORNER)
    p = po.make_params(oracle, "Sedov", **cfg)
    st, od, cycles, final_dt, _ = po.run(oracle, p, init_fields=sedov_corner_fields(p))
    assert st == 0 and cycles == 300
    print(f"case Sedov corner 256^2 engine={engine}", flush=True)
    q = armon.make_params("Sedov", engine=ENGINES[engine], **cfg)
    assert q.bc_override == 1 and q.nb_cells == p.nb_cells
    r, s = armon.run(q, sedov_corner_fields(q))
    assert r.status == 0, armon.errors()
    assert (r.cycles, r.final_dt) == (cycles, final_dt)
    assert_same_state(s, od)
    s.close()


# Boundary factors other than +-1: a no-slip left wall (u = 0), partial factors elsewhere.
NON_UNIT_FACTORS = ((0.0, 1.0), (0.5, 0.75), (1.0, 0.0), (-0.25, -1.0))


@pytest.mark.parametrize("engine", ["fused", "staged"])
@pytest.mark.parametrize("scheme,proj", [(1, 2), (0, 1)])
def test_non_unit_boundary_factors(armon, po, oracle, engine, scheme, proj):
    """Per-side factors other than +-1: a mirror ghost takes its real cell's p and c
    (halo_exchange.cpp:23-30), not the EOS of its scaled velocities. Circular Sod 96x80 until the
    waves have crossed the walls (100 cycles), both engines bit for bit against the oracle; with
    Godunov + first order the stencil also reads a stale layer beyond the mirrored one."""
    check(armon, po, oracle, "Sod_circ", ENGINES[engine], scheme=scheme, projection=proj, limiter=1, nb_ghosts=4,
          nx=96, ny=80, max_cycles=100, max_time=0.6, verbose=5, bc_factors=NON_UNIT_FACTORS)


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_sod_circ_radius_override(armon, po, oracle, engine):
    """The circular Sod of radius 0.2 (the reference's is 0.3), through initial fields written
    after init_test: 256x256, GAD + minmod, second-order remap, CFL 0.7, 200 cycles, bit for bit
    against the oracle given the same fields."""
    from conftest import sod_circ_radius_fields

    cfg = dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=256, ny=256, cfl=0.7, max_cycles=200,
               max_time=1.0, verbose=5)
    p = po.make_params(oracle, "Sod_circ", **cfg)
    st, od, cycles, final_dt, _ = po.run(oracle, p, init_fields=sod_circ_radius_fields(p, 0.2))
    assert st == 0 and cycles == 200
    print(f"case Sod_circ r=0.2 256^2 engine={engine}", flush=True)
    q = armon.make_params("Sod_circ", engine=ENGINES[engine], **cfg)
    r, s = armon.run(q, sod_circ_radius_fields(q, 0.2))
    assert r.status == 0, armon.errors()
    assert (r.cycles, r.final_dt) == (cycles, final_dt)
    assert_same_state(s, od)
    s.close()


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_bizarrium_left_moving_override(armon, po, oracle, engine):
    """The Bizarrium tube with its left state moving at u = 250 and the right one at rest (the
    reference moves the right one), through initial fields written after init_test: 256x256,
    GAD + superbee, second-order remap, CFL 0.6, 200 cycles, bit for bit against the oracle given
    the same fields."""
    from conftest import bizarrium_fields

    cfg = dict(scheme=1, projection=2, limiter=2, nb_ghosts=4, nx=256, ny=256, cfl=0.6, max_cycles=200,
               max_time=1.0, verbose=5)
    p = po.make_params(oracle, "Bizarrium", **cfg)
    st, od, cycles, final_dt, _ = po.run(oracle, p, init_fields=bizarrium_fields(p, left_moving=True))
    assert st == 0 and cycles == 200
    print(f"case Bizarrium left-moving 256^2 engine={engine}", flush=True)
    q = armon.make_params("Bizarrium", engine=ENGINES[engine], **cfg)
    r, s = armon.run(q, bizarrium_fields(q, left_moving=True))
    assert r.status == 0, armon.errors()
    assert (r.cycles, r.final_dt) == (cycles, final_dt)
    assert_same_state(s, od)
    s.close()


@pytest.mark.parametrize("engine", ["staged", "fused"])
def test_stage_time_table(armon, engine):
    """With verbose < 3 the solver times each stage with HIP events under the reference's labels
    (armon_2D.cpp:306-322, 401): every stage of every cycle is counted, and the times are
    positive."""
    cfg = dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=128, ny=96, max_cycles=20, verbose=2,
               graph_chunk=8)
    print(f"case stage table {engine}", flush=True)
    r, s = product_run(armon, "Sedov", ENGINES[engine], **cfg)
    assert r.status == 0 and r.cycles == 20
    t = s.stage_times()
    s.close()
    if engine == "staged":
        want = {"init_test": 1, "dt_CFL": 20, "EOS": 40, "boundary_conditions": 40, "numerical_fluxes": 40,
                "cell_update": 40, "euler_proj": 40}
    else:   # the time loop launches graphs of 8 cycles until the device reports the end: 3
        want = {"init_test": 1, "fused cycles (graph)": 3}
    assert {k: v[1] for k, v in t.items()} == want, t
    assert all(ms > 0 for ms, _ in t.values()), t
