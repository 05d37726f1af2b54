"""Per-entry-point parity of the drop-in C-ABI (include/armon/kernels.h) against the CPU oracle.

Each test fills the arrays of a small domain with seeded random states, runs the HIP entry
point over the same Range/InnerRange the reference driver would use, runs the oracle on host
copies, and compares every element of every array the kernel may write, bit for bit (FP64,
both built without FMA contraction, same operation order).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20241202


def bits_equal(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def assert_bits(name, got, want):
    if not bits_equal(got, want):
        bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
        i = bad[0]
        raise AssertionError(f"{name}: {bad.size} differing elements, first at {i}: {got[i]!r} vs {want[i]!r}")


class Case:
    """Random arrays of one domain on both sides (device views for the product, host arrays for the oracle)."""

    RANGES = {"rho": (0.5, 2.0), "umat": (-1.0, 1.0), "vmat": (-1.0, 1.0), "u": (-1.0, 1.0),
              "Emat": (2.0, 5.0), "pmat": (0.1, 2.0), "cmat": (0.5, 2.0), "gmat": (0.5, 2.0),
              "ustar": (-1.0, 1.0), "pstar": (0.1, 2.0), "x": (0, 1), "y": (0, 1), "domain_mask": (0, 1)}

    def __init__(self, armon, n, names, seed=SEED, ranges=None):
        self.armon = armon
        rng = np.random.default_rng(seed)
        self.host, self.dev = {}, {}
        for k in names:
            lo, hi = (ranges or {}).get(k, self.RANGES.get(k, (-0.1, 0.1)))
            self.host[k] = rng.uniform(lo, hi, n)
            self.dev[k] = armon.DeviceArray(n, f"Armon_{k}", self.host[k])

    def v(self, k):
        return C.byref(self.dev[k].view)

    def h(self, k):
        return self.host[k].ctypes.data_as(C.POINTER(C.c_double))

    def compare(self, names):
        self.armon.load().armon_synchronize()
        assert not self.armon.errors()
        for k in names:
            assert_bits(k, self.dev[k].download(), self.host[k])


def domain(armon, axis=0, scheme=1, proj=2, nx=40, ny=33, g=5, test="Sod"):
    p = armon.make_params(test, scheme=scheme, projection=proj, nx=nx, ny=ny, nb_ghosts=g,
                          engine=armon.ENGINE_STAGED)
    armon.load().armon_update_axis(C.byref(p), axis)
    return p


def it2d(armon, po, d):
    lib = armon.load()
    r, ir = armon.Range(), armon.InnerRange2D()
    lib.armon_domain_iter2D(C.byref(d), C.byref(r), C.byref(ir))
    return r, ir, po.ORange(r.start, r.end), po.OInnerRange2D(ir.main_range_start, ir.main_range_step,
                                                               ir.row_range_start, ir.row_range_length)


@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0])
def test_perfect_gas_EOS(armon, po, oracle, gamma):
    lib = armon.load()
    p = domain(armon)
    names = ["rho", "Emat", "umat", "vmat", "pmat", "cmat", "gmat"]
    c = Case(armon, p.nb_cells, names)
    r, ir, orr, oir = it2d(armon, po, lib.armon_real_domain(C.byref(p)))
    lib.perfect_gas_EOS(C.byref(r), C.byref(ir), gamma, *[c.v(k) for k in names])
    oracle.o_perfect_gas_EOS(C.byref(orr), C.byref(oir), gamma, *[c.h(k) for k in names])
    c.compare(["pmat", "cmat", "gmat"])


def test_bizarrium_EOS(armon, po, oracle):
    lib = armon.load()
    p = domain(armon, test="Bizarrium")
    names = ["rho", "umat", "vmat", "Emat", "pmat", "cmat", "gmat"]
    c = Case(armon, p.nb_cells, names, ranges={"rho": (0.9e4, 1.45e4), "Emat": (3e4, 5e6),
                                               "umat": (-300, 300), "vmat": (-300, 300)})
    r, ir, orr, oir = it2d(armon, po, lib.armon_real_domain(C.byref(p)))
    lib.bizarrium_EOS(C.byref(r), C.byref(ir), *[c.v(k) for k in names])
    oracle.o_bizarrium_EOS(C.byref(orr), C.byref(oir), *[c.h(k) for k in names])
    c.compare(["pmat", "cmat", "gmat"])


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("proj", [1, 2])
def test_acoustic(armon, po, oracle, axis, proj):
    lib = armon.load()
    p = domain(armon, axis=axis, scheme=0, proj=proj)
    names = ["ustar", "pstar", "rho", "u", "pmat", "cmat"]
    c = Case(armon, p.nb_cells, names)
    r, ir, orr, oir = it2d(armon, po, lib.armon_domain_fluxes(C.byref(p)))
    lib.acoustic(C.byref(r), C.byref(ir), p.s, *[c.v(k) for k in names])
    oracle.o_acoustic(C.byref(orr), C.byref(oir), p.s, *[c.h(k) for k in names])
    c.compare(["ustar", "pstar"])


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("limiter", [0, 1, 2])
def test_acoustic_GAD(armon, po, oracle, axis, limiter):
    lib = armon.load()
    p = domain(armon, axis=axis)
    names = ["ustar", "pstar", "rho", "u", "pmat", "cmat"]
    c = Case(armon, p.nb_cells, names, seed=SEED + limiter)
    r, ir, orr, oir = it2d(armon, po, lib.armon_domain_fluxes(C.byref(p)))
    dt, dx = 0.0123, p.dx
    lib.acoustic_GAD(C.byref(r), C.byref(ir), p.s, dt, dx, *[c.v(k) for k in names], limiter)
    assert oracle.o_acoustic_GAD(C.byref(orr), C.byref(oir), p.s, dt, dx, *[c.h(k) for k in names], limiter) == 0
    c.compare(["ustar", "pstar"])


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("cut", ["full", "partial"])
def test_acoustic_GAD_shared_faces(armon, po, oracle, axis, cut):
    """The GAD flux kernels that solve each face once (LDS tiles of 254 faces along x, runs of
    rows along y): a domain of several tiles and runs, over the whole flux range or a range cut in
    the middle of its first and last rows, bit for bit against the oracle."""
    lib = armon.load()
    p = domain(armon, axis=axis, nx=600, ny=300)
    names = ["ustar", "pstar", "rho", "u", "pmat", "cmat"]
    c = Case(armon, p.nb_cells, names, seed=SEED + 11 + axis)
    r, ir, orr, oir = it2d(armon, po, lib.armon_domain_fluxes(C.byref(p)))
    if cut == "partial":
        n = ir.row_range_length
        r.start, r.end = n // 3 + 5 * n, r.end - n // 2 - 3 * n - 7
        orr.start, orr.end = r.start, r.end
    print(f"case acoustic_GAD axis {axis} range {r.start}..{r.end} of rows of {ir.row_range_length}", flush=True)
    dt, dx = 0.0123, p.dx
    lib.acoustic_GAD(C.byref(r), C.byref(ir), p.s, dt, dx, *[c.v(k) for k in names], 2)
    assert oracle.o_acoustic_GAD(C.byref(orr), C.byref(oir), p.s, dt, dx, *[c.h(k) for k in names], 2) == 0
    c.compare(["ustar", "pstar"])


def test_acoustic_GAD_other_stride(armon, po, oracle):
    """A stride that is neither 1 nor the row step goes through the generic kernel (three solves
    per face, as acoustic.cpp:65-86): s = 2 over the x flux range, bit for bit."""
    lib = armon.load()
    p = domain(armon, axis=0, nx=60, ny=20)
    names = ["ustar", "pstar", "rho", "u", "pmat", "cmat"]
    c = Case(armon, p.nb_cells, names, seed=SEED + 21)
    r, ir, orr, oir = it2d(armon, po, lib.armon_domain_fluxes(C.byref(p)))
    print("case acoustic_GAD s=2", flush=True)
    lib.acoustic_GAD(C.byref(r), C.byref(ir), 2, 0.0123, p.dx, *[c.v(k) for k in names], 1)
    assert oracle.o_acoustic_GAD(C.byref(orr), C.byref(oir), 2, 0.0123, p.dx, *[c.h(k) for k in names], 1) == 0
    c.compare(["ustar", "pstar"])


@pytest.mark.parametrize("axis", [0, 1])
def test_cell_update(armon, po, oracle, axis):
    lib = armon.load()
    p = domain(armon, axis=axis)
    names = ["ustar", "pstar", "rho", "u", "Emat"]
    c = Case(armon, p.nb_cells, names)
    r, ir, orr, oir = it2d(armon, po, lib.armon_domain_cell_update(C.byref(p)))
    dt = 0.0071
    lib.cell_update(C.byref(r), C.byref(ir), p.s, p.dx, dt, *[c.v(k) for k in names])
    oracle.o_cell_update(C.byref(orr), C.byref(oir), p.s, p.dx, dt, *[c.h(k) for k in names])
    c.compare(["rho", "u", "Emat"])


ADV = ["ustar", "rho", "umat", "vmat", "Emat", "work_array_1", "work_array_2", "work_array_3", "work_array_4"]


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("cut", ["full", "partial"])
def test_advection_second_order_tiles_and_runs(armon, po, oracle, axis, cut):
    """The second-order advection kernels that load each element once (LDS tiles of 252 faces
    along x, runs of rows along y): a domain of several tiles and runs, over the whole advection
    range or a range cut in the middle of its first and last rows, with velocities of both signs
    (both donor sides), bit for bit against the oracle."""
    lib = armon.load()
    p = domain(armon, axis=axis, proj=2, nx=600, ny=300)
    c = Case(armon, p.nb_cells, ADV, seed=SEED + 31 + axis)
    r, ir, orr, oir = it2d(armon, po, lib.armon_domain_advection(C.byref(p)))
    if cut == "partial":
        n = ir.row_range_length
        r.start, r.end = n // 3 + 5 * n, r.end - n // 2 - 3 * n - 7
        orr.start, orr.end = r.start, r.end
    print(f"case advection_second_order axis {axis} range {r.start}..{r.end} of rows of {ir.row_range_length}",
          flush=True)
    dt = 0.0071
    lib.advection_second_order(C.byref(r), C.byref(ir), p.s, p.dx, dt, *[c.v(k) for k in ADV])
    oracle.o_advection_second_order(C.byref(orr), C.byref(oir), p.s, p.dx, dt, *[c.h(k) for k in ADV])
    c.compare(ADV[5:])


def test_advection_second_order_other_stride(armon, po, oracle):
    """A stride that is neither 1 nor the row step goes through the generic per-face kernel:
    s = 2 over the x advection range, bit for bit."""
    lib = armon.load()
    p = domain(armon, axis=0, proj=2, nx=60, ny=20)
    c = Case(armon, p.nb_cells, ADV, seed=SEED + 41)
    r, ir, orr, oir = it2d(armon, po, lib.armon_domain_advection(C.byref(p)))
    print("case advection_second_order s=2", flush=True)
    lib.advection_second_order(C.byref(r), C.byref(ir), 2, p.dx, 0.0071, *[c.v(k) for k in ADV])
    oracle.o_advection_second_order(C.byref(orr), C.byref(oir), 2, p.dx, 0.0071, *[c.h(k) for k in ADV])
    c.compare(ADV[5:])


@pytest.mark.parametrize("axis", [0, 1])
def test_advection_first_order(armon, po, oracle, axis):
    lib = armon.load()
    p = domain(armon, axis=axis, proj=1)
    c = Case(armon, p.nb_cells, ADV)
    r, ir, orr, oir = it2d(armon, po, lib.armon_domain_advection(C.byref(p)))
    dt = 0.0071
    lib.advection_first_order(C.byref(r), C.byref(ir), p.s, dt, *[c.v(k) for k in ADV])
    oracle.o_advection_first_order(C.byref(orr), C.byref(oir), p.s, dt, *[c.h(k) for k in ADV])
    c.compare(ADV[5:])


@pytest.mark.parametrize("axis", [0, 1])
def test_advection_second_order(armon, po, oracle, axis):
    lib = armon.load()
    p = domain(armon, axis=axis, proj=2)
    c = Case(armon, p.nb_cells, ADV)
    r, ir, orr, oir = it2d(armon, po, lib.armon_domain_advection(C.byref(p)))
    dt = 0.0071
    lib.advection_second_order(C.byref(r), C.byref(ir), p.s, p.dx, dt, *[c.v(k) for k in ADV])
    oracle.o_advection_second_order(C.byref(orr), C.byref(oir), p.s, p.dx, dt, *[c.h(k) for k in ADV])
    c.compare(ADV[5:])


@pytest.mark.parametrize("axis", [0, 1])
def test_euler_projection(armon, po, oracle, axis):
    lib = armon.load()
    p = domain(armon, axis=axis)
    names = ["ustar", "rho", "umat", "vmat", "Emat", "work_array_1", "work_array_2", "work_array_3", "work_array_4"]
    c = Case(armon, p.nb_cells, names)
    r, ir, orr, oir = it2d(armon, po, lib.armon_real_domain(C.byref(p)))
    dt = 0.0071
    lib.euler_projection(C.byref(r), C.byref(ir), p.s, p.dx, dt, *[c.v(k) for k in names])
    oracle.o_euler_projection(C.byref(orr), C.byref(oir), p.s, p.dx, dt, *[c.h(k) for k in names])
    c.compare(["rho", "umat", "vmat", "Emat"])


@pytest.mark.parametrize("side", [0, 1, 2, 3])
@pytest.mark.parametrize("width", [1, 3, 5])
def test_boundary_conditions(armon, po, oracle, side, width):
    lib = armon.load()
    p = domain(armon)
    names = ["rho", "umat", "vmat", "pmat", "cmat", "gmat", "Emat"]
    c = Case(armon, p.nb_cells, names, seed=SEED + side)
    disp = C.c_int()
    d = lib.armon_boundary_conditions_domain(C.byref(p), side, C.byref(disp))
    r, ir = armon.Range(), armon.InnerRange1D()
    lib.armon_domain_direct_iter1D(C.byref(d), C.byref(r), C.byref(ir))
    uf, vf = (-1.0, 1.0) if side < 2 else (1.0, -1.0)
    lib.boundary_conditions(C.byref(r), C.byref(ir), width, disp.value, uf, vf, *[c.v(k) for k in names])
    oracle.o_boundary_conditions(C.byref(po.ORange(r.start, r.end)), C.byref(po.OInnerRange1D(ir.start, ir.step)),
                                 width, disp.value, uf, vf, *[c.h(k) for k in names])
    c.compare(names)


FIELDS7 = ["rho", "umat", "vmat", "pmat", "cmat", "gmat", "Emat"]


def _border(armon, po, p, g):
    """Left border block: g ghost-width columns of the ny real rows (the Julia halo hook)."""
    d = armon.DomainRange(armon.load().armon_index_1D(C.byref(p), 0, 0), p.row_length,
                          armon.load().armon_index_1D(C.byref(p), 0, p.ny - 1), 0, 1, g - 1)
    return it2d(armon, po, d)


def test_read_border_array(armon, po, oracle):
    p = domain(armon)
    g = p.nb_ghosts
    c = Case(armon, p.nb_cells, FIELDS7)
    size = 7 * g * p.ny
    vals = Case(armon, size, ["value_array"], seed=SEED + 9)
    r, ir, orr, oir = _border(armon, po, p, g)
    armon.load().read_border_array(C.byref(r), C.byref(ir), p.ny, g, *[c.v(k) for k in FIELDS7], vals.v("value_array"))
    oracle.o_read_border_array(C.byref(orr), C.byref(oir), p.ny, g, *[c.h(k) for k in FIELDS7],
                               vals.h("value_array"))
    vals.compare(["value_array"])


def test_write_border_array(armon, po, oracle):
    p = domain(armon)
    g = p.nb_ghosts
    c = Case(armon, p.nb_cells, FIELDS7)
    vals = Case(armon, 7 * g * p.ny, ["value_array"], seed=SEED + 9)
    r, ir, orr, oir = _border(armon, po, p, g)
    armon.load().write_border_array(C.byref(r), C.byref(ir), p.ny, g, *[c.v(k) for k in FIELDS7],
                                    vals.v("value_array"))
    oracle.o_write_border_array(C.byref(orr), C.byref(oir), p.ny, g, *[c.h(k) for k in FIELDS7],
                                vals.h("value_array"))
    c.compare(FIELDS7)


INIT = ["x", "y", "rho", "Emat", "umat", "vmat", "domain_mask", "pmat", "cmat", "gmat", "ustar", "pstar"]


@pytest.mark.parametrize("test", ["Sod", "Sod_y", "Sod_circ", "Bizarrium", "Sedov"])
@pytest.mark.parametrize("debug", [False, True])
def test_init_test(armon, po, oracle, test, debug):
    lib = armon.load()
    p = armon.make_params(test, nx=41, ny=29, nb_ghosts=4, engine=armon.ENGINE_STAGED)
    c = Case(armon, p.nb_cells, INIT)
    r, ir, orr, oir = it2d(armon, po, lib.armon_complete_domain(C.byref(p)))
    opt = lib.armon_test_option(C.byref(p))
    args = (p.row_length, p.nb_ghosts, p.nx, p.ny, p.domain_size[0], p.domain_size[1], p.domain_origin[0],
            p.domain_origin[1], 0, 0, p.nx, p.ny)
    lib.init_test(C.byref(r), C.byref(ir), *args, *[c.v(k) for k in INIT], p.test, debug, opt)
    assert oracle.o_init_test(C.byref(orr), C.byref(oir), *args, *[c.h(k) for k in INIT], p.test, debug, opt) == 0
    c.compare(INIT)


@pytest.mark.parametrize("nx,ny", [(40, 33), (257, 3)])
def test_dt_CFL(armon, po, oracle, nx, ny):
    lib = armon.load()
    p = domain(armon, nx=nx, ny=ny)
    names = ["umat", "vmat", "cmat", "domain_mask"]
    c = Case(armon, p.nb_cells, names, ranges={"umat": (-3, 3), "vmat": (-3, 3)})
    c.host["domain_mask"][:] = 0.0
    d = lib.armon_real_domain(C.byref(p))
    for j in range(p.ny):
        s = p.index_start + j * p.row_length
        c.host["domain_mask"][s:s + p.nx] = 1.0
    c.dev["domain_mask"].upload(c.host["domain_mask"])
    r, ir = armon.Range(), armon.InnerRange1D()
    lib.armon_domain_iter1D(C.byref(d), C.byref(r), C.byref(ir))
    dx, dy = 1.0 / nx, 1.0 / ny
    got = lib.dt_CFL(C.byref(r), C.byref(ir), dx, dy, *[c.v(k) for k in names])
    want = oracle.o_dt_CFL(C.byref(po.ORange(r.start, r.end)), C.byref(po.OInnerRange1D(ir.start, ir.step)),
                           dx, dy, *[c.h(k) for k in names])
    assert not armon.errors()
    assert got == want
    # device-resident variant
    out = armon.DeviceArray(1)
    lib.dt_CFL_device(C.byref(r), C.byref(ir), dx, dy, *[c.v(k) for k in names], out.ptr)
    assert out.download()[0] == want


def test_dt_CFL_all_infinite_is_dbl_max(armon, po, oracle):
    """Zero wave speed everywhere: every bound is +inf and the reduction returns Kokkos'
    Min identity, the largest finite double."""
    lib = armon.load()
    p = domain(armon, nx=16, ny=8)
    names = ["umat", "vmat", "cmat", "domain_mask"]
    c = Case(armon, p.nb_cells, names)
    for k in ("umat", "vmat", "cmat"):
        c.host[k][:] = 0.0
        c.dev[k].upload(c.host[k])
    c.host["domain_mask"][:] = 1.0
    c.dev["domain_mask"].upload(c.host["domain_mask"])
    d = lib.armon_real_domain(C.byref(p))
    r, ir = armon.Range(), armon.InnerRange1D()
    lib.armon_domain_iter1D(C.byref(d), C.byref(r), C.byref(ir))
    got = lib.dt_CFL(C.byref(r), C.byref(ir), 0.1, 0.1, *[c.v(k) for k in names])
    want = oracle.o_dt_CFL(C.byref(po.ORange(r.start, r.end)), C.byref(po.OInnerRange1D(ir.start, ir.step)),
                           0.1, 0.1, *[c.h(k) for k in names])
    assert got == want == np.finfo(np.float64).max


def test_conservation_vars(armon, po, oracle):
    lib = armon.load()
    p = domain(armon, nx=300, ny=200)
    names = ["rho", "Emat", "domain_mask"]
    c = Case(armon, p.nb_cells, names)
    d = lib.armon_real_domain(C.byref(p))
    r, ir = armon.Range(), armon.InnerRange1D()
    lib.armon_domain_iter1D(C.byref(d), C.byref(r), C.byref(ir))
    m, e = C.c_double(), C.c_double()
    lib.conservation_vars(C.byref(r), C.byref(ir), p.dx, *[c.v(k) for k in names], C.byref(m), C.byref(e))
    om, oe = C.c_double(), C.c_double()
    oracle.o_conservation_vars(C.byref(po.ORange(r.start, r.end)), C.byref(po.OInnerRange1D(ir.start, ir.step)),
                               p.dx, *[c.h(k) for k in names], C.byref(om), C.byref(oe))
    # both compensated; the summation orders differ
    assert abs(m.value - om.value) <= 1e-15 * abs(om.value)
    assert abs(e.value - oe.value) <= 1e-15 * abs(oe.value)
    # and against an exact sum
    from math import fsum
    idx = [ir.start + k * ir.step for k in range(r.start, r.end)]
    cm = [c.host["rho"][i] * c.host["domain_mask"][i] * (p.dx * p.dx) for i in idx]
    ce = [cm[j] * c.host["Emat"][i] for j, i in enumerate(idx)]
    assert abs(m.value - fsum(cm)) <= 2e-16 * abs(fsum(cm))
    assert abs(e.value - fsum(ce)) <= 2e-16 * abs(fsum(ce))


def test_empty_ranges_change_nothing(armon):
    """An empty iteration range (start == end) does nothing for every entry point, as a Kokkos
    parallel_for over an empty RangePolicy, and the reductions return their identities: dt_CFL
    the largest finite double (Kokkos' Min identity), conservation_vars zeros."""
    lib = armon.load()
    p = domain(armon)
    names = sorted(set(ADV) | {"pstar", "u", "pmat", "cmat", "gmat", "domain_mask"})
    c = Case(armon, p.nb_cells, names)
    r, ir = armon.Range(), armon.InnerRange2D()
    lib.armon_domain_iter2D(C.byref(lib.armon_real_domain(C.byref(p))), C.byref(r), C.byref(ir))
    r.end = r.start
    R, IR = C.byref(r), C.byref(ir)
    v = lambda *ks: [c.v(k) for k in ks]
    dt = 0.0071
    lib.acoustic(R, IR, p.s, *v("ustar", "pstar", "rho", "u", "pmat", "cmat"))
    lib.acoustic_GAD(R, IR, p.s, dt, p.dx, *v("ustar", "pstar", "rho", "u", "pmat", "cmat"), 1)
    lib.perfect_gas_EOS(R, IR, 1.4, *v("rho", "Emat", "umat", "vmat", "pmat", "cmat", "gmat"))
    lib.bizarrium_EOS(R, IR, *v("rho", "umat", "vmat", "Emat", "pmat", "cmat", "gmat"))
    lib.cell_update(R, IR, p.s, p.dx, dt, *v("ustar", "pstar", "rho", "u", "Emat"))
    lib.advection_first_order(R, IR, p.s, dt, *v(*ADV))
    lib.advection_second_order(R, IR, p.s, p.dx, dt, *v(*ADV))
    lib.euler_projection(R, IR, p.s, p.dx, dt, *v(*ADV))
    r1, ir1 = armon.Range(), armon.InnerRange1D()
    lib.armon_domain_iter1D(C.byref(lib.armon_real_domain(C.byref(p))), C.byref(r1), C.byref(ir1))
    r1.end = r1.start
    assert lib.dt_CFL(C.byref(r1), C.byref(ir1), 0.1, 0.1, *v("umat", "vmat", "cmat", "domain_mask")) == \
        np.finfo(np.float64).max
    m, e = C.c_double(1.0), C.c_double(1.0)
    lib.conservation_vars(C.byref(r1), C.byref(ir1), p.dx, *v("rho", "Emat", "domain_mask"), C.byref(m), C.byref(e))
    assert (m.value, e.value) == (0.0, 0.0)
    c.compare(names)
