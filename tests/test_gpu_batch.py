"""Batched ensemble on the GPU: every domain of a batch must end bit for bit where the oracle
ends when it runs that domain alone (same test, size, perturbed states, parameters), whatever
the launch configuration, and the FP32-mixed batch must stay within the FP32 error bound."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("rho", "umat", "vmat", "Emat", "pmat", "cmat")
NAMES = {0: "Sod", 1: "Sod_y", 2: "Sod_circ", 3: "Bizarrium", 4: "Sedov"}


def small_specs(armon, seed, n, min_side=24, max_side=96):
    return armon.random_specs(seed, n, min_area=min_side * min_side, max_area=max_side * max_side, max_aspect=4.0)


def oracle_domain(po, oracle, bp, spec, **cfg):
    """The oracle run of one domain with the batch's parameters for it."""
    extra = {}
    if bp.bc_override:
        extra["bc_factors"] = [(bp.bc_factors[side][0], bp.bc_factors[side][1]) for side in range(4)]
    p = po.make_params(oracle, int(bp.test), nx=bp.nx, ny=bp.ny, nb_ghosts=bp.nb_ghosts, scheme=bp.scheme,
                       projection=bp.projection, limiter=bp.limiter, axis_splitting=bp.axis_splitting,
                       stencil_width=bp.stencil_width, cfl=bp.cfl, max_time=bp.max_time,
                       max_cycles=bp.max_cycles, cst_dt=bp.cst_dt, Dt=bp.Dt,
                       dt_on_even_cycles=bp.dt_on_even_cycles, verbose=5, **extra)
    st = po.OTestState(**{f: getattr(spec.state, f) for f, _ in po.OTestState._fields_})
    status, d, cycles, final_dt, _ = po.run(oracle, p, states=st)
    return status, d, cycles, final_dt


def compare_batch(armon, po, oracle, b, specs, fields=STATE):
    msgs = []
    for k in range(b.n):
        bp = b.params(k)
        r = b.result(k)
        status, od, cycles, final_dt = oracle_domain(po, oracle, bp, specs[k])
        assert status == 0
        tag = f"domain {k} ({NAMES[int(specs[k].test)]} {bp.nx}x{bp.ny})"
        if r.cycles != cycles or r.final_dt != final_dt:
            msgs.append(f"{tag}: cycles {r.cycles} vs {cycles}, final dt {r.final_dt} vs {final_dt}")
            continue
        mask = od.array("domain_mask") == 1
        for f in fields:
            a, o = b.read(k, f)[mask], od.array(f)[mask]
            if not np.array_equal(a.view(np.uint64), o.view(np.uint64)):
                msgs.append(f"{tag} {f}: {int(np.count_nonzero(a != o))} cells differ, "
                            f"max {np.max(np.abs(a - o)):.3e}")
    assert not msgs, "; ".join(msgs[:8])


def run_batch(armon, specs, tuning=None, autotune=False, advance=None, bc_factors=None, **cfg):
    common = armon.batch_params(**cfg)
    if bc_factors is not None:
        common.bc_override = 1
        for side, (uf, vf) in enumerate(bc_factors):
            common.bc_factors[side][0], common.bc_factors[side][1] = uf, vf
    b = armon.Batch(common, specs)
    b.init()
    if autotune:
        b.autotune()
    if tuning:
        b.set_tuning(*tuning)
    b.begin()
    if advance is None:
        b.run()
    else:
        b.advance(advance)
    assert b.finish() == 0, armon.errors()
    return b


BASE = dict(scheme=1, projection=2, limiter=1, axis_splitting=0, nb_ghosts=4, max_cycles=40, verbose=5)


def test_batch_matches_oracle_per_domain(armon, po, oracle):
    """12 uneven domains (Sod, circular Sod, Sedov with perturbed states), 40 cycles each."""
    specs = small_specs(armon, 11, 12)
    print("case batch seed 11 n 12 default tuning", flush=True)
    b = run_batch(armon, specs, **BASE)
    cycles = [b.result(k).cycles for k in range(b.n)]
    assert max(cycles) == 40 and min(cycles) > 0, cycles   # small Sod domains reach t = 0.2 earlier
    compare_batch(armon, po, oracle, b, specs)
    b.close()


@pytest.mark.parametrize("tuning", [(16, 0, 0, 128), (64, 7, 5, 512), (1000, 1, 1, 256), (33, 300, 0, 128),
                                    (8, 3, 0, 512), (48, 0, 2, 256, 2), (128, 0, 0, 512, 2)])
def test_batch_launch_configurations(armon, po, oracle, tuning):
    """Y run lengths, X tile widths, cells per thread and persistent grids change the work split,
    never the result (tuning = (Y run rows, X grid, Y grid, X tile cells[, X cells per thread]))."""
    specs = small_specs(armon, 5, 6)
    print(f"case batch seed 5 tuning {tuning}", flush=True)
    b = run_batch(armon, specs, tuning=tuning, **dict(BASE, max_cycles=25))
    compare_batch(armon, po, oracle, b, specs)
    b.close()


def test_batch_autotune(armon, po, oracle):
    specs = small_specs(armon, 3, 8)
    print("case batch autotune", flush=True)
    b = run_batch(armon, specs, autotune=True, **dict(BASE, max_cycles=20))
    t = b.tuning()
    assert t.y_run_rows in (16, 32, 64, 128, 256, 512, 1024) and t.ms_x > 0 and t.ms_y > 0
    assert t.tile_x in (128, 256, 512) and t.x_cells_per_thread in (1, 2)
    assert t.x_cells_per_thread == 1 or (t.tile_x in (256, 512) and t.grid_x == 0)
    assert t.items_x > 0 and t.items_y > 0
    compare_batch(armon, po, oracle, b, specs)
    b.close()


def test_batch_end_times_differ(armon, po, oracle):
    """Domains reach their own end times at different cycles; finished domains stop changing
    while the others continue (per-domain dt, time and cycle on the device)."""
    specs = small_specs(armon, 21, 6, 16, 48)
    for k, s in enumerate(specs):
        s.max_time = 0.004 * (k + 1)
    print("case batch end times", flush=True)
    b = run_batch(armon, specs, **dict(BASE, max_cycles=100000))
    cycles = [b.result(k).cycles for k in range(b.n)]
    assert len(set(cycles)) > 1, cycles
    compare_batch(armon, po, oracle, b, specs)
    b.close()


@pytest.mark.parametrize("scheme,proj,limiter,split", [(0, 1, 1, 0), (1, 2, 2, 1), (1, 1, 0, 2), (0, 2, 1, 3)])
def test_batch_schemes_and_splittings(armon, po, oracle, scheme, proj, limiter, split):
    specs = small_specs(armon, 8, 5)
    cfg = dict(BASE, scheme=scheme, projection=proj, limiter=limiter, axis_splitting=split, max_cycles=20)
    print(f"case batch scheme {scheme} proj {proj} limiter {limiter} split {split}", flush=True)
    b = run_batch(armon, specs, **cfg)
    compare_batch(armon, po, oracle, b, specs)
    b.close()


@pytest.mark.parametrize("chunk", [0, 4])
def test_batch_graph_chunks(armon, po, oracle, chunk):
    specs = small_specs(armon, 4, 4)
    print(f"case batch graph chunk {chunk}", flush=True)
    b = run_batch(armon, specs, advance=23, **dict(BASE, max_cycles=23, graph_chunk=chunk))
    compare_batch(armon, po, oracle, b, specs)
    b.close()


def test_batch_bizarrium(armon, po, oracle):
    """A batch of Bizarrium domains (the EOS is shared by all domains of a batch)."""
    lib = armon.load()
    specs = (armon.ArmonDomainSpec * 3)()
    for k, (nx, ny) in enumerate([(40, 24), (32, 56), (64, 40)]):
        q = armon.make_params("Bizarrium", nx=nx, ny=ny)
        specs[k].test, specs[k].nx, specs[k].ny, specs[k].max_time = 3, nx, ny, 0.0
        specs[k].state = armon.reference_state("Bizarrium", lib.armon_test_option(C.byref(q)))
    specs[1].state.high_rho *= 1.1
    print("case batch bizarrium", flush=True)
    b = run_batch(armon, specs, **dict(BASE, max_cycles=30))
    compare_batch(armon, po, oracle, b, specs)
    b.close()


def test_batch_fp32_mixed(armon):
    """FP32-mixed batch against the FP64 batch after 60 cycles: density L1 <= 1e-5 per domain."""
    specs = small_specs(armon, 17, 8, 48, 160)
    for s in specs:
        s.max_time = 1e9   # stop on the cycle count in both precisions
    print("case batch fp32 mixed", flush=True)
    b64 = run_batch(armon, specs, **dict(BASE, max_cycles=60))
    b32 = run_batch(armon, specs, **dict(BASE, max_cycles=60, precision=1))
    for k in range(b64.n):
        assert b32.result(k).cycles == b64.result(k).cycles == 60
        mask = b64.read(k, "domain_mask") == 1
        a, o = b32.read(k, "rho")[mask], b64.read(k, "rho")[mask]
        l1 = np.sum(np.abs(a - o)) / np.sum(np.abs(o))
        assert l1 <= 1e-5, (k, l1)
    b64.close()
    b32.close()


def test_batch_fp32_mixed_independent_of_launch_configuration(armon):
    """FP32-mixed results are not bit-comparable with the oracle, but they must not depend on how
    the work is split: every run length and grid gives the same bits (each cell is computed by the
    same expressions, contracted the same way, wherever its tile or run falls)."""
    specs = small_specs(armon, 23, 6)
    ref = None
    for tuning in [(16, 0, 0, 128), (64, 7, 5, 512), (1000, 1, 1, 256), (33, 300, 0, 128),
                   (8, 3, 0, 512)]:
        print(f"case batch fp32 mixed seed 23 tuning {tuning}", flush=True)
        b = run_batch(armon, specs, tuning=tuning, **dict(BASE, max_cycles=25, precision=1))
        out = [b.read(k, f) for k in range(b.n) for f in STATE]
        b.close()
        if ref is None:
            ref = out
            continue
        bad = [i for i, (a, o) in enumerate(zip(out, ref)) if not np.array_equal(a.view(np.uint64), o.view(np.uint64))]
        assert not bad, (tuning, [(i // len(STATE), STATE[i % len(STATE)]) for i in bad[:8]])



def test_batch_boundary_factor_override(armon, po, oracle):
    """Boundary factors given in the batch's common parameters replace every domain's test
    factors (here reflective walls on all four sides, which no reference test has): each domain
    ends bit for bit where the oracle ends with the same override."""
    specs = small_specs(armon, 31, 6)
    print("case batch seed 31 n 6 reflective walls", flush=True)
    walls = ((-1.0, 1.0), (-1.0, 1.0), (1.0, -1.0), (1.0, -1.0))
    b = run_batch(armon, specs, bc_factors=walls, **BASE)
    for k in range(b.n):
        bp = b.params(k)
        assert bp.bc_override == 1 and [tuple(bp.bc_factors[s]) for s in range(4)] == list(walls)
    compare_batch(armon, po, oracle, b, specs)
    b.close()


def test_batch_non_unit_boundary_factors(armon, po, oracle):
    """Boundary factors other than +-1 in the batch's common parameters (a no-slip wall and
    partial factors): every domain bit for bit against the oracle with the same override, whose
    mirror ghosts copy the real cells' p and c."""
    specs = small_specs(armon, 37, 6)
    print("case batch seed 37 n 6 non-unit factors", flush=True)
    factors = ((0.0, 1.0), (0.5, 0.75), (1.0, 0.0), (-0.25, -1.0))
    b = run_batch(armon, specs, bc_factors=factors, **BASE)
    compare_batch(armon, po, oracle, b, specs)
    b.close()


def test_batch_set_tuning_rejects_invalid_cells_per_thread(armon):
    """Two cells per thread exist for X tiles of 256 and 512 cells with one block per tile only;
    other values are refused with the tuning left as it was."""
    specs = small_specs(armon, 5, 3)
    b = armon.Batch(armon.batch_params(**dict(BASE, max_cycles=5)), specs)
    try:
        b.init()
        before = b.tuning()
        for bad in ((16, 0, 0, 128, 2), (16, 4, 0, 256, 2), (16, 0, 0, 256, 3), (16, 0, 0, 256, -1)):
            with pytest.raises(armon.ArmonError, match="cells per thread|x_cells_per_thread"):
                b.set_tuning(*bad)
            armon.errors()
            t = b.tuning()
            assert (t.tile_x, t.x_cells_per_thread, t.grid_x, t.y_run_rows) == \
                (before.tile_x, before.x_cells_per_thread, before.grid_x, before.y_run_rows)
        b.set_tuning(16, 0, 0, 512, 2)
        t = b.tuning()
        assert (t.tile_x, t.x_cells_per_thread, t.y_run_rows) == (512, 2, 16)
    finally:
        b.close()
