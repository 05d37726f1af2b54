"""The float build of the drop-in C-ABI (lib/libarmon_f32.so: flt_t = float, the reference's
USE_SINGLE_PRECISION, common.h:8-12) against the FP64 oracle.

The inputs are float32 values; the oracle runs in FP64 on the same values (exactly representable),
so its outputs are the FP64 results the float kernels approximate. Tolerances, per kernel:
- copies and sign flips (boundary_conditions, read/write_border_array) and the region states
  of init_test: equal to the oracle's values rounded to float;
- Bizarrium EOS: computed in FP64 from the float state, as the FP64 library does, so equal to the
  oracle's FP64 results rounded to float;
- arithmetic kernels (EOS, fluxes, cell update, advection, projection): largest difference at
  most RTOL times the largest magnitude of that output (normwise), RTOL = 1e-5, about 170 float
  ulps: float rounding of a few tens of operations, with the cancellations of the GAD limiter
  ratios and the minmod slopes;
- dt_CFL and the conservation sums: relative 1e-6 and 1e-5.
The reference's float build itself differs between its CPU and GPU backends (its powf calls),
so the float library is held to the FP64 algorithm within these bounds rather than bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20241203
RTOL = 1e-5


def f32(x):
    """A Python float rounded to float32 and back: the value both sides receive."""
    return float(np.float32(x))


class Case32:
    """Random float32 arrays on the device (float library) and the same values in FP64 for the oracle."""

    RANGES = {"rho": (0.5, 2.0), "umat": (-1.0, 1.0), "vmat": (-1.0, 1.0), "u": (-1.0, 1.0),
              "Emat": (2.0, 5.0), "pmat": (0.1, 2.0), "cmat": (0.5, 2.0), "gmat": (0.5, 2.0),
              "ustar": (-1.0, 1.0), "pstar": (0.1, 2.0), "x": (0, 1), "y": (0, 1), "domain_mask": (0, 1)}

    def __init__(self, armon, n, names, seed=SEED, ranges=None):
        rng = np.random.default_rng(seed)
        self.armon = armon
        self.host, self.dev = {}, {}
        for k in names:
            lo, hi = (ranges or {}).get(k, self.RANGES.get(k, (-0.1, 0.1)))
            vals = rng.uniform(lo, hi, n).astype(np.float32)
            self.host[k] = vals.astype(np.float64)
            self.dev[k] = armon.DeviceArray(n, f"Armon_{k}", vals, single=True)

    def v(self, k):
        return C.byref(self.dev[k].view)

    def h(self, k):
        return self.host[k].ctypes.data_as(C.POINTER(C.c_double))

    def got(self, k):
        self.armon.load_f32().armon_synchronize()
        assert not self.armon.errors()
        return self.dev[k].download()

    def exact(self, names):
        for k in names:
            got, want = self.got(k), self.host[k].astype(np.float32)
            bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
            assert bad.size == 0, f"{k}: {bad.size} differ, first at {bad[0]}: {got[bad[0]]!r} vs {want[bad[0]]!r}"

    def close(self, names, rtol=RTOL):
        for k in names:
            got, want = self.got(k).astype(np.float64), self.host[k]
            err = np.max(np.abs(got - want))
            scale = np.max(np.abs(want))
            print(f"{k}: max |float - fp64| = {err:.3e}, max |fp64| = {scale:.3e}", flush=True)
            assert np.all(np.isfinite(got)) and err <= rtol * scale, k


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


def test_float_library_reports_float(armon):
    lib = armon.load_f32()
    assert lib.flt_size() == 4 and lib.idx_size() == 8


@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0])
def test_perfect_gas_EOS_f32(armon, po, oracle, gamma):
    lib, p = armon.load_f32(), domain(armon)
    names = ["rho", "Emat", "umat", "vmat", "pmat", "cmat", "gmat"]
    c = Case32(armon, p.nb_cells, names)
    r, ir, orr, oir = it2d(armon, po, armon.load().armon_real_domain(C.byref(p)))
    print(f"case float perfect_gas_EOS gamma {gamma}", flush=True)
    lib.perfect_gas_EOS(C.byref(r), C.byref(ir), gamma, *[c.v(k) for k in names])
    oracle.o_perfect_gas_EOS(C.byref(orr), C.byref(oir), f32(gamma), *[c.h(k) for k in names])
    c.close(["pmat", "cmat", "gmat"])


def test_bizarrium_EOS_f32(armon, po, oracle):
    lib, p = armon.load_f32(), domain(armon, test="Bizarrium")
    names = ["rho", "umat", "vmat", "Emat", "pmat", "cmat", "gmat"]
    c = Case32(armon, p.nb_cells, names, ranges={"rho": (0.9e4, 1.45e4), "Emat": (3e4, 5e6),
                                                 "umat": (-300, 300), "vmat": (-300, 300)})
    r, ir, orr, oir = it2d(armon, po, armon.load().armon_real_domain(C.byref(p)))
    print("case float bizarrium_EOS", flush=True)
    lib.bizarrium_EOS(C.byref(r), C.byref(ir), *[c.v(k) for k in names])
    oracle.o_bizarrium_EOS(C.byref(orr), C.byref(oir), *[c.h(k) for k in names])
    c.exact(["pmat", "cmat", "gmat"])


@pytest.mark.parametrize("axis", [0, 1])
def test_acoustic_f32(armon, po, oracle, axis):
    lib, p = armon.load_f32(), domain(armon, axis=axis, scheme=0, proj=1)
    names = ["ustar", "pstar", "rho", "u", "pmat", "cmat"]
    c = Case32(armon, p.nb_cells, names)
    r, ir, orr, oir = it2d(armon, po, armon.load().armon_domain_fluxes(C.byref(p)))
    print(f"case float acoustic axis {axis}", flush=True)
    lib.acoustic(C.byref(r), C.byref(ir), p.s, *[c.v(k) for k in names])
    oracle.o_acoustic(C.byref(orr), C.byref(oir), p.s, *[c.h(k) for k in names])
    c.close(["ustar", "pstar"])


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("limiter", [0, 1, 2])
def test_acoustic_GAD_f32(armon, po, oracle, axis, limiter):
    """Over a domain of several X tiles and Y runs (the shared-face kernels)."""
    lib, p = armon.load_f32(), domain(armon, axis=axis, nx=600, ny=300)
    names = ["ustar", "pstar", "rho", "u", "pmat", "cmat"]
    c = Case32(armon, p.nb_cells, names, seed=SEED + limiter)
    r, ir, orr, oir = it2d(armon, po, armon.load().armon_domain_fluxes(C.byref(p)))
    dt, dx = f32(0.0123), f32(p.dx)
    print(f"case float acoustic_GAD axis {axis} limiter {limiter}", flush=True)
    lib.acoustic_GAD(C.byref(r), C.byref(ir), p.s, dt, dx, *[c.v(k) for k in names], limiter)
    assert oracle.o_acoustic_GAD(C.byref(orr), C.byref(oir), p.s, dt, dx, *[c.h(k) for k in names], limiter) == 0
    c.close(["ustar", "pstar"])


@pytest.mark.parametrize("axis", [0, 1])
def test_cell_update_f32(armon, po, oracle, axis):
    lib, p = armon.load_f32(), domain(armon, axis=axis)
    names = ["ustar", "pstar", "rho", "u", "Emat"]
    c = Case32(armon, p.nb_cells, names)
    r, ir, orr, oir = it2d(armon, po, armon.load().armon_domain_cell_update(C.byref(p)))
    dt, dx = f32(0.0071), f32(p.dx)
    print(f"case float cell_update axis {axis}", flush=True)
    lib.cell_update(C.byref(r), C.byref(ir), p.s, dx, dt, *[c.v(k) for k in names])
    oracle.o_cell_update(C.byref(orr), C.byref(oir), p.s, dx, dt, *[c.h(k) for k in names])
    c.close(["rho", "u", "Emat"])


ADV = ["ustar", "rho", "umat", "vmat", "Emat", "work_array_1", "work_array_2", "work_array_3", "work_array_4"]


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("order", [1, 2])
def test_advection_f32(armon, po, oracle, axis, order):
    """Both orders, over a domain of several X tiles and Y runs, velocities of both signs. dt
    keeps dt |u*| within 6% of dx, as the CFL condition does: the Lagrangian lengths stay near dx
    and the slope ratios well conditioned."""
    lib, p = armon.load_f32(), domain(armon, axis=axis, proj=order, nx=600, ny=300)
    c = Case32(armon, p.nb_cells, ADV, seed=SEED + 31 + axis)
    r, ir, orr, oir = it2d(armon, po, armon.load().armon_domain_advection(C.byref(p)))
    dt, dx = f32(2e-4), f32(p.dx)
    print(f"case float advection order {order} axis {axis}", flush=True)
    if order == 1:
        lib.advection_first_order(C.byref(r), C.byref(ir), p.s, dt, *[c.v(k) for k in ADV])
        oracle.o_advection_first_order(C.byref(orr), C.byref(oir), p.s, dt, *[c.h(k) for k in ADV])
    else:
        lib.advection_second_order(C.byref(r), C.byref(ir), p.s, dx, dt, *[c.v(k) for k in ADV])
        oracle.o_advection_second_order(C.byref(orr), C.byref(oir), p.s, dx, dt, *[c.h(k) for k in ADV])
    c.close(ADV[5:])


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("nx", [40, 64])
def test_euler_projection_f32(armon, po, oracle, axis, nx):
    """dx = 1/40 (divided) and 1/64 (a power of two: the product with its reciprocal). Fluxes of
    the size dt u rho E that dt = 2e-4 gives, so the remapped density stays near rho."""
    lib, p = armon.load_f32(), domain(armon, axis=axis, nx=nx, ny=nx)
    names = ["ustar", "rho", "umat", "vmat", "Emat", "work_array_1", "work_array_2", "work_array_3", "work_array_4"]
    c = Case32(armon, p.nb_cells, names, ranges={f"work_array_{q}": (-1e-3, 1e-3) for q in range(1, 5)})
    r, ir, orr, oir = it2d(armon, po, armon.load().armon_real_domain(C.byref(p)))
    dt, dx = f32(2e-4), f32(p.dx)
    print(f"case float euler_projection axis {axis} dx {dx}", flush=True)
    lib.euler_projection(C.byref(r), C.byref(ir), p.s, dx, dt, *[c.v(k) for k in names])
    oracle.o_euler_projection(C.byref(orr), C.byref(oir), p.s, dx, dt, *[c.h(k) for k in names])
    c.close(["rho", "umat", "vmat", "Emat"])


FIELDS7 = ["rho", "umat", "vmat", "pmat", "cmat", "gmat", "Emat"]


@pytest.mark.parametrize("side", [0, 1, 2, 3])
def test_boundary_conditions_f32(armon, po, oracle, side):
    lib, p = armon.load_f32(), domain(armon)
    c = Case32(armon, p.nb_cells, FIELDS7, seed=SEED + side)
    disp = C.c_int()
    d = armon.load().armon_boundary_conditions_domain(C.byref(p), side, C.byref(disp))
    r, ir = armon.Range(), armon.InnerRange1D()
    armon.load().armon_domain_direct_iter1D(C.byref(d), C.byref(r), C.byref(ir))
    uf, vf = (-1.0, 1.0) if side < 2 else (1.0, -1.0)
    print(f"case float boundary_conditions side {side}", flush=True)
    lib.boundary_conditions(C.byref(r), C.byref(ir), 3, disp.value, uf, vf, *[c.v(k) for k in FIELDS7])
    oracle.o_boundary_conditions(C.byref(po.ORange(r.start, r.end)), C.byref(po.OInnerRange1D(ir.start, ir.step)),
                                 3, disp.value, uf, vf, *[c.h(k) for k in FIELDS7])
    c.exact(FIELDS7)


def test_border_arrays_f32(armon, po, oracle):
    """read_border_array then write_border_array over the left border block (the halo hook)."""
    lib, p = armon.load_f32(), domain(armon)
    g = p.nb_ghosts
    c = Case32(armon, p.nb_cells, FIELDS7)
    vals = Case32(armon, 7 * g * p.ny, ["value_array"], seed=SEED + 9)
    d = armon.DomainRange(armon.load().armon_index_1D(C.byref(p), 0, 0), p.row_length,
                          armon.load().armon_index_1D(C.byref(p), 0, p.ny - 1), 0, 1, g - 1)
    r, ir, orr, oir = it2d(armon, po, d)
    print("case float read/write_border_array", flush=True)
    lib.read_border_array(C.byref(r), C.byref(ir), p.ny, g, *[c.v(k) for k in FIELDS7], vals.v("value_array"))
    oracle.o_read_border_array(C.byref(orr), C.byref(oir), p.ny, g, *[c.h(k) for k in FIELDS7],
                               vals.h("value_array"))
    vals.exact(["value_array"])
    # new float values, written back into the same cells on both sides
    vals.host["value_array"] = (vals.host["value_array"] + 1.0).astype(np.float32).astype(np.float64)
    vals.dev["value_array"].upload(vals.host["value_array"])
    lib.write_border_array(C.byref(r), C.byref(ir), p.ny, g, *[c.v(k) for k in FIELDS7], vals.v("value_array"))
    oracle.o_write_border_array(C.byref(orr), C.byref(oir), p.ny, g, *[c.h(k) for k in FIELDS7],
                                vals.h("value_array"))
    c.exact(FIELDS7)


INIT = ["x", "y", "rho", "Emat", "umat", "vmat", "domain_mask", "pmat", "cmat", "gmat", "ustar", "pstar"]


@pytest.mark.parametrize("test", ["Sod", "Sod_y", "Sod_circ", "Bizarrium", "Sedov"])
def test_init_test_f32(armon, po, oracle, test):
    """Cell-centre coordinates within float rounding, and every cell in the same region with the
    test's states rounded to float (no centre lies within rounding of a region boundary here)."""
    lib, p = armon.load_f32(), armon.make_params(test, nx=41, ny=29, nb_ghosts=4, engine=armon.ENGINE_STAGED)
    c = Case32(armon, p.nb_cells, INIT)
    r, ir, orr, oir = it2d(armon, po, armon.load().armon_complete_domain(C.byref(p)))
    opt = f32(armon.load().armon_test_option(C.byref(p)))
    sizes = [f32(v) for v in (p.domain_size[0], p.domain_size[1], p.domain_origin[0], p.domain_origin[1])]
    args = (p.row_length, p.nb_ghosts, p.nx, p.ny, *sizes, 0, 0, p.nx, p.ny)
    print(f"case float init_test {test}", flush=True)
    lib.init_test(C.byref(r), C.byref(ir), *args, *[c.v(k) for k in INIT], p.test, False, opt)
    assert oracle.o_init_test(C.byref(orr), C.byref(oir), *args, *[c.h(k) for k in INIT], p.test, False, opt) == 0
    c.close(["x", "y"], rtol=1e-6)
    if test == "Sedov":
        # the blast energy 0.851072 / (pi r^2) is a float quotient of float operands
        c.close(["Emat"], rtol=1e-6)
        c.exact([k for k in INIT[2:] if k != "Emat"])
    else:
        c.exact(INIT[2:])


def test_dt_CFL_f32(armon, po, oracle):
    lib, p = armon.load_f32(), domain(armon)
    names = ["umat", "vmat", "cmat", "domain_mask"]
    c = Case32(armon, p.nb_cells, names, ranges={"umat": (-3, 3), "vmat": (-3, 3)})
    c.host["domain_mask"][:] = 0.0
    for j in range(p.ny):
        s = p.index_start + j * p.row_length
        c.host["domain_mask"][s:s + p.nx] = 1.0
    c.dev["domain_mask"].upload(c.host["domain_mask"])
    d = armon.load().armon_real_domain(C.byref(p))
    r, ir = armon.Range(), armon.InnerRange1D()
    armon.load().armon_domain_iter1D(C.byref(d), C.byref(r), C.byref(ir))
    dx, dy = f32(p.dx), f32(p.domain_size[1] / p.ny)
    print("case float dt_CFL", flush=True)
    got = lib.dt_CFL(C.byref(r), C.byref(ir), dx, dy, *[c.v(k) for k in names])
    want = oracle.o_dt_CFL(C.byref(po.ORange(r.start, r.end)), C.byref(po.OInnerRange1D(ir.start, ir.step)),
                           dx, dy, *[c.h(k) for k in names])
    assert not armon.errors()
    assert abs(got - want) <= 1e-6 * abs(want), (got, want)


def test_conservation_vars_f32(armon, po, oracle):
    lib, p = armon.load_f32(), domain(armon)
    names = ["rho", "Emat", "domain_mask"]
    c = Case32(armon, p.nb_cells, names)
    d = armon.load().armon_real_domain(C.byref(p))
    r, ir = armon.Range(), armon.InnerRange1D()
    armon.load().armon_domain_iter1D(C.byref(d), C.byref(r), C.byref(ir))
    dx = f32(p.dx)
    m, e = C.c_float(), C.c_float()
    print("case float conservation_vars", flush=True)
    lib.conservation_vars(C.byref(r), C.byref(ir), dx, *[c.v(k) for k in names], C.byref(m), C.byref(e))
    om, oe = C.c_double(), C.c_double()
    oracle.o_conservation_vars(C.byref(po.ORange(r.start, r.end)), C.byref(po.OInnerRange1D(ir.start, ir.step)),
                               dx, *[c.h(k) for k in names], C.byref(om), C.byref(oe))
    assert not armon.errors()
    assert abs(m.value - om.value) <= 1e-5 * abs(om.value) and abs(e.value - oe.value) <= 1e-5 * abs(oe.value)


def test_invalid_limiter_reports_f32(armon):
    lib, p = armon.load_f32(), domain(armon)
    names = ["ustar", "pstar", "rho", "u", "pmat", "cmat"]
    c = Case32(armon, p.nb_cells, names)
    r, ir = armon.Range(), armon.InnerRange2D()
    armon.load().armon_domain_iter2D(C.byref(armon.load().armon_domain_fluxes(C.byref(p))), C.byref(r), C.byref(ir))
    armon.errors()
    lib.acoustic_GAD(C.byref(r), C.byref(ir), p.s, 0.01, f32(p.dx), *[c.v(k) for k in names], 7)
    errs = armon.errors()
    assert errs and errs[0][0] == "acoustic_GAD"
