"""Parity at the benchmark's sizes and over whole runs.

- Sedov 8192x8192 (the benchmark configuration) for 20 cycles: every field bit-identical to
  the oracle.
- Sedov 256x256 run to its end time t = 1 (thousands of cycles): identical cycle count, final
  dt and state.
- Bizarrium 1024x1024 for 200 cycles (minmod) and 500 steps (superbee): identical.
- The 1D tubes at full size (Bizarrium 4096x4096 for 500 steps, Sod 8192x8192 for 200):
  every row stays equal to the first.
- A size-independent property at 8192x8192: mass and energy conserved to 1e-12 over 100
  cycles (FP64 fused).

Whole runs whose oracle side takes minutes (Sedov 1024^2 to t = 1, Bizarrium 4096^2, circular
Sod 8192^2 with its FP32-mixed form) are in test_gpu_golden_runs.py."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("rho", "umat", "vmat", "Emat", "pmat", "cmat")
BENCH = dict(scheme=1, projection=2, limiter=1, axis_splitting=0, nb_ghosts=4, cfl=0.7, verbose=5)


def compare_run(armon, po, oracle, test, **cfg):
    print(f"case {test} {cfg}: oracle", flush=True)
    p = po.make_params(oracle, test, **cfg)
    st, od, cycles, final_dt, _ = po.run(oracle, p)
    assert st == 0
    print(f"case {test}: GPU", flush=True)
    q = armon.make_params(test, engine=1, **cfg)
    r, s = armon.run(q)
    assert r.status == 0, armon.errors()
    assert (r.cycles, r.final_dt) == (cycles, final_dt)
    mask = od.array("domain_mask") == 1
    bad = []
    for f in STATE:
        a, b = s.read(f)[mask], od.array(f)[mask]
        if not np.array_equal(a.view(np.uint64), b.view(np.uint64)):
            bad.append(f"{f}: {int(np.count_nonzero(a != b))} cells differ, max {np.max(np.abs(a - b)):.3e}")
    s.close()
    del od
    assert not bad, "; ".join(bad)
    return cycles


def test_sedov_8192_20_cycles(armon, po, oracle):
    assert compare_run(armon, po, oracle, "Sedov", nx=8192, ny=8192, max_cycles=20, **BENCH) == 20


def test_sedov_256_to_end_time(armon, po, oracle):
    cycles = compare_run(armon, po, oracle, "Sedov", nx=256, ny=256, max_cycles=10**6, **BENCH)
    print(f"Sedov 256^2 reached t = 1 in {cycles} cycles", flush=True)
    assert cycles > 1000


def test_bizarrium_1024_200_cycles(armon, po, oracle):
    cfg = dict(BENCH, cfl=0.6, max_time=1.0)
    assert compare_run(armon, po, oracle, "Bizarrium", nx=1024, ny=1024, max_cycles=200, **cfg) == 200


def test_bizarrium_superbee_1024_500_steps(armon, po, oracle):
    """The Bizarrium shock tube of the benchmark's parity set (GAD + superbee, 500 steps), at
    1024^2 so that the oracle finishes in seconds."""
    cfg = dict(BENCH, limiter=2, cfl=0.6, max_time=1.0)
    assert compare_run(armon, po, oracle, "Bizarrium", nx=1024, ny=1024, max_cycles=500, **cfg) == 500


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("test,n", [("Bizarrium", 4096), ("Sod", 8192)])
def test_one_dimensional_tubes_stay_row_invariant(armon, test, n, precision):
    """At full size, FP64 and FP32-mixed: the Bizarrium and Sod tubes vary only along x and the y
    walls reflect, so every row must stay equal to the first one (the zero y velocity may change
    sign). Every cell of a column is computed by the same arithmetic whatever its position in
    the tiles, runs and unrolled march, so this holds exactly in both precisions."""
    steps = 500 if test == "Bizarrium" else 200
    cfg = dict(BENCH, limiter=2 if test == "Bizarrium" else 1, cfl=0.6 if test == "Bizarrium" else 0.95,
               max_time=1.0, nx=n, ny=n, max_cycles=steps)
    print(f"case {test} {n}^2 {steps} steps precision {precision}: row invariance", flush=True)
    r, s = armon.run(armon.make_params(test, engine=1, precision=precision, **cfg))
    assert r.status == 0 and r.cycles == steps, armon.errors()
    p = s.params()
    g, rl = p.nb_ghosts, p.row_length
    for f in ("rho", "umat", "vmat", "Emat", "pmat"):
        a = s.read(f).reshape(-1, rl)[g:g + n, g:g + n]
        assert np.all(np.isfinite(a)), f
        assert np.array_equal(a, np.broadcast_to(a[0], a.shape)), f   # == treats -0 and +0 alike
    rho = s.read("rho").reshape(-1, rl)[g, g:g + n]
    assert rho.max() > rho.min()   # the tube has evolved, not just stayed put
    s.close()


def test_conservation_8192(armon):
    p = armon.make_params("Sedov", nx=8192, ny=8192, max_cycles=100, engine=1, **BENCH)
    s = armon.Solver(p)
    s.init_test()
    lib = armon.load()
    s.begin()
    m0, e0 = C.c_double(), C.c_double()
    assert lib.armon_solver_conservation(s.h, C.byref(m0), C.byref(e0)) == 0
    s.advance(100)
    r = s.finish()
    m1, e1 = C.c_double(), C.c_double()
    assert lib.armon_solver_conservation(s.h, C.byref(m1), C.byref(e1)) == 0
    assert r.status == 0 and r.cycles == 100
    assert abs(m1.value - m0.value) <= 1e-12 * abs(m0.value)
    assert abs(e1.value - e0.value) <= 1e-12 * abs(e0.value)
    s.close()


def test_conservation_32768(armon):
    """The largest square domain below the 2^31-cell limit of the Params indexing that is a
    power of two: Sedov 32768^2 (1.07 G cells, 137 GB for the reference's 16 FP64 arrays in
    HBM), 20 cycles of the fused engine, mass and energy conserved to 1e-12."""
    print("case Sedov 32768^2 fused, 20 cycles", flush=True)
    p = armon.make_params("Sedov", nx=32768, ny=32768, max_cycles=20, engine=1, **BENCH)
    s = armon.Solver(p)
    try:
        s.init_test()
        lib = armon.load()
        s.begin()
        m0, e0 = C.c_double(), C.c_double()
        assert lib.armon_solver_conservation(s.h, C.byref(m0), C.byref(e0)) == 0
        s.advance(20)
        r = s.finish()
        m1, e1 = C.c_double(), C.c_double()
        assert lib.armon_solver_conservation(s.h, C.byref(m1), C.byref(e1)) == 0
        assert r.status == 0 and r.cycles == 20, armon.errors()
        assert abs(m1.value - m0.value) <= 1e-12 * abs(m0.value)
        assert abs(e1.value - e0.value) <= 1e-12 * abs(e0.value)
    finally:
        s.close()
