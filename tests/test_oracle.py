"""Pins the CPU oracle (oracle/armon_oracle.c) before anything is compared with it.

Each check is chosen so that a plausible transcription mistake (dropped term, wrong sign,
swapped operand, off-by-one index) fails it:
  * the reference's own known-answer tests (tests/armon_2D_tests.cpp): indexing KATs, the
    cycle counts of the NaN-check suite, conservation to 1e-12 over 400 cycles;
  * closed forms computed independently (exact rational arithmetic, textbook values);
  * invariants (mirror symmetry at reflective walls, brute-force minima, transposition).
No test re-types an oracle formula.
"""
import ctypes as C
import math
import os
import random
from fractions import Fraction

import numpy as np
import pytest

from conftest import reference_cycle_counts


def default_params(po, oracle):
    # armon_2D_tests.cpp:61-80
    return po.make_params(oracle, "Sod", scheme=po.SCHEME_GAD, limiter=po.LIM_MINMOD,
                          projection=po.PROJ_EULER_2ND, axis_splitting=po.SPLIT_SEQUENTIAL,
                          max_cycles=100, verbose=1, nx=100, ny=5, nb_ghosts=5)


# ---- indexing KATs: armon_2D_tests.cpp:83-189 --------------------------------------------

def test_indexing_domains(po, oracle, kats):
    p = default_params(po, oracle)
    assert oracle.o_params_check(C.byref(p)) == 1
    P = C.byref(p)
    doms = {
        "real": oracle.o_real_domain(P), "fluxes": oracle.o_domain_fluxes(P),
        "advection": oracle.o_domain_advection(P), "complete": oracle.o_complete_domain(P),
    }
    for k, exp in kats["indexing"]["domains"].items():
        d = C.byref(doms[k])
        assert doms[k].as_tuple() == tuple(exp["range"]), k
        assert oracle.o_dr_length(d) == exp["length"], k
        assert oracle.o_dr_row_length(d) == exp["row_length"], k
        assert oracle.o_dr_col_length(d) == exp["col_length"], k
        assert oracle.o_dr_begin(d) == exp["begin"], k
        assert oracle.o_dr_end(d) == exp["end"], k


def test_indexing_ranges(po, oracle, kats):
    k1, k2 = kats["indexing"]["iter1D_real"], kats["indexing"]["iter2D_real"]
    p = default_params(po, oracle)
    d = oracle.o_real_domain(C.byref(p))
    r, i1 = po.ORange(), po.OInnerRange1D()
    oracle.o_dr_iter1d(C.byref(d), C.byref(r), C.byref(i1))
    assert [r.start, r.end] == k1["range"]
    assert i1.start == k1["start"]
    assert oracle.o_scale_index_1d(C.byref(i1), r.start) == oracle.o_dr_begin(C.byref(d))
    assert oracle.o_scale_index_1d(C.byref(i1), r.end - 1) == oracle.o_dr_end(C.byref(d))

    i2 = po.OInnerRange2D()
    oracle.o_dr_iter2d(C.byref(d), C.byref(r), C.byref(i2))
    assert [r.start, r.end] == k2["range"]
    assert [i2.main_range_start, i2.main_range_step, i2.row_range_start, i2.row_range_length] == k2["inner"]
    s = lambda k: oracle.o_scale_index_2d(C.byref(i2), k)  # noqa: E731
    assert s(r.start) == oracle.o_dr_begin(C.byref(d))
    assert s(r.end - 1) == oracle.o_dr_end(C.byref(d))
    assert s(r.start + 100) == oracle.o_dr_begin(C.byref(d)) + p.row_length
    assert s(r.start + 200) == oracle.o_dr_begin(C.byref(d)) + 2 * p.row_length


def test_indexing_boundary_conditions(po, oracle, kats):
    p = default_params(po, oracle)
    disp = C.c_int()
    sides = {"Bottom": po.SIDE_BOTTOM, "Left": po.SIDE_LEFT, "Right": po.SIDE_RIGHT, "Top": po.SIDE_TOP}
    for name, exp in kats["indexing"]["boundary_conditions"].items():
        d = oracle.o_boundary_conditions_domain(C.byref(p), sides[name], C.byref(disp))
        got = (oracle.o_dr_begin(C.byref(d)), oracle.o_dr_end(C.byref(d)), oracle.o_dr_length(C.byref(d)), disp.value)
        assert got == (exp["begin"], exp["end"], exp["length"], exp["disp"]), name
    for i, j, idx in kats["indexing"]["index_1D"]:
        assert oracle.o_index_1D(C.byref(p), i, j) == idx


def test_memory_alignment(po, oracle, kats):
    # armon_2D_tests.cpp:192-202
    p = default_params(po, oracle)
    d = po.Data(oracle, p.nb_cells)
    for f, _ in po.OData._fields_:
        assert C.cast(getattr(d.c, f), C.c_void_p).value % kats["indexing"]["alignment_bytes"] == 0, f


# ---- whole-run KATs: armon_2D_tests.cpp:205-275 --------------------------------------------

_libm = C.CDLL("libm.so.6")
FE_INVALID = 0x01
FE_ALL_EXCEPT = 0x3D


@pytest.mark.parametrize("test,cycles", reference_cycle_counts())
def test_nan_check_cycle_counts(po, oracle, test, cycles):
    """Reference params (reference.cpp:33-64) to the test's end time: the reference's cycle
    counts, and no invalid floating-point operation (the reference traps FE_INVALID)."""
    threads = oracle.o_num_threads()
    oracle.o_set_num_threads(1)   # FP flags are per thread: run on this one
    try:
        p = po.reference_params(oracle, test)
        _libm.feclearexcept(FE_ALL_EXCEPT)
        st, d, ncycles, dt, _ = po.run(oracle, p)
        invalid = _libm.fetestexcept(FE_INVALID)
    finally:
        oracle.o_set_num_threads(threads)
    assert st == 0
    assert ncycles == cycles
    assert invalid == 0
    assert math.isfinite(dt) and dt > 0


@pytest.mark.parametrize("test", ["Sod", "Sod_y", "Sod_circ", "Sedov"])
def test_conservation(po, oracle, test):
    """armon_2D_tests.cpp:246-275: mass and energy conserved to 1e-12 over 400 cycles."""
    p = po.reference_params(oracle, test, max_time=10000.0, max_cycles=400)
    d = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(d.c), False)
    m0, e0 = C.c_double(), C.c_double()
    oracle.o_conservation_vars_params(C.byref(p), C.byref(d.c), C.byref(m0), C.byref(e0))
    lt, fdt, cyc = C.c_double(), C.c_double(), C.c_int()
    assert oracle.o_time_loop(C.byref(p), C.byref(d.c), C.byref(lt), C.byref(fdt), C.byref(cyc)) == 0
    m1, e1 = C.c_double(), C.c_double()
    oracle.o_conservation_vars_params(C.byref(p), C.byref(d.c), C.byref(m1), C.byref(e1))
    assert cyc.value == 400
    assert abs(m1.value - m0.value) <= 1e-12 * max(abs(m0.value), abs(m1.value))
    assert abs(e1.value - e0.value) <= 1e-12 * max(abs(e0.value), abs(e1.value))


# ---- closed forms -------------------------------------------------------------------------

def _eos_point(po, oracle, fn, rho, E, u, v, gamma=None):
    r, ir = po.ORange(0, 1), po.OInnerRange2D(0, 1, 0, 1)
    a = {k: np.array([val], dtype=np.float64) for k, val in
         dict(rho=rho, E=E, u=u, v=v, p=0.0, c=0.0, g=0.0).items()}
    P = po.ptr
    if fn == "perfect":
        oracle.o_perfect_gas_EOS(C.byref(r), C.byref(ir), gamma, P(a["rho"]), P(a["E"]), P(a["u"]), P(a["v"]),
                                 P(a["p"]), P(a["c"]), P(a["g"]))
    else:
        oracle.o_bizarrium_EOS(C.byref(r), C.byref(ir), P(a["rho"]), P(a["u"]), P(a["v"]), P(a["E"]),
                               P(a["p"]), P(a["c"]), P(a["g"]))
    return a["p"][0], a["c"][0], a["g"][0]


def ulps(a, b):
    return abs(a - b) / math.ulp(max(abs(a), abs(b)))


@pytest.mark.parametrize("rho,E,u,v,p_exact", [
    (1.0, 2.5, 0.0, 0.0, Fraction(1)),               # Sod left state: p = 1
    (0.125, 2.0, 0.0, 0.0, Fraction(1, 10)),         # Sod right state: p = 0.1
    (2.0, 5.0, 1.0, 2.0, Fraction(2)),               # e = 5 - (1 + 4)/2 = 2.5, p = 0.4*2*2.5
    (1.0, 3.0, -2.0, 0.0, Fraction(2, 5)),           # e = 1, p = 0.4
])
def test_perfect_gas_closed_form(po, oracle, rho, E, u, v, p_exact):
    p, c, g = _eos_point(po, oracle, "perfect", rho, E, u, v, 1.4)
    assert ulps(p, float(p_exact)) <= 4
    c_exact = math.sqrt(1.4 * float(p_exact) / rho)
    assert ulps(c, c_exact) <= 4
    assert g == 1.2                                   # (1 + gamma) / 2


def _bizarrium_exact(rho, E, u, v):
    """fluid_eos.cpp:26-62 in exact rational arithmetic (the reference's constants)."""
    F = Fraction
    rho0, K0, Cv0, T0, eps0 = F(10**4), F(10**11), F(1000), F(300), F(0)
    G0, s = F(3, 2), F(3, 2)
    q, r = F(-42080895, 14941154), F(727668333, 149411540)
    rho, E, u, v = F(rho), F(E), F(u), F(v)
    x = rho / rho0 - 1
    g = G0 * (1 - rho0 / rho)
    f0 = (1 + (s / 3 - 2) * x + q * x**2 + r * x**3) / (1 - s * x)
    f1 = (s / 3 - 2 + 2 * q * x + 3 * r * x**2 + s * f0) / (1 - s * x)
    f2 = (2 * q + 6 * r * x + 2 * s * f1) / (1 - s * x)
    eps_k0 = eps0 - Cv0 * T0 * (1 + g) + F(1, 2) * (K0 / rho0) * x**2 * f0
    pk0 = -Cv0 * T0 * G0 * rho0 + F(1, 2) * K0 * x * (1 + x)**2 * (2 * f0 + x * f1)
    pk0prime = -F(1, 2) * K0 * (1 + x)**3 * rho0 * (2 * (1 + 3 * x) * f0 + 2 * x * (2 + 3 * x) * f1
                                                   + x**2 * (1 + x) * f2)
    e = E - F(1, 2) * (u**2 + v**2)
    p = pk0 + G0 * rho0 * (e - eps_k0)
    c2 = (G0 * rho0 * (p - pk0) - pk0prime) / rho**2
    return float(p), math.sqrt(float(c2))


def test_bizarrium_reference_density(po, oracle):
    """At rho = rho0 the EOS reduces to x = 0: pk0 = -Cv0*T0*G0*rho0 = -4.5e9, eps_k0 = -3e5,
    pk0' = -K0*rho0 = -1e15, so p = -4.5e9 + 1.5e4*(e + 3e5) and c = sqrt(1.5e4*(p + 4.5e9) + 1e15)/1e4."""
    E = 31250.0
    p, c, _ = _eos_point(po, oracle, "bizarrium", 1e4, E, 0.0, 0.0)
    p_exact = -4.5e9 + 1.5e4 * (E + 3e5)
    assert ulps(p, p_exact) <= 2
    assert ulps(c, math.sqrt(1.5e4 * (p_exact + 4.5e9) + 1e15) / 1e4) <= 4


@pytest.mark.parametrize("rho,E,u,v", [
    (1.42857142857e4, 4.48657821135e6, 0.0, 0.0),    # Bizarrium left state (init_test.cpp:25-35)
    (1.0e4, 31250.0, 250.0, 0.0),                    # right state
    (1.2e4, 1.0e6, 100.0, -50.0),
    (0.95e4, 2.0e5, -30.0, 80.0),
])
def test_bizarrium_exact_rational(po, oracle, rho, E, u, v):
    p, c, _ = _eos_point(po, oracle, "bizarrium", rho, E, u, v)
    p_ex, c_ex = _bizarrium_exact(rho, E, u, v)
    # the double evaluation cancels large terms (pk0 ~ -4.5e9 against G0*rho0*(e - eps_k0)):
    # allow a relative error of 1e-10 of the largest term, far below any transcription error
    scale = 4.5e9 + abs(p_ex)
    assert abs(p - p_ex) <= 1e-10 * scale
    assert abs(c - c_ex) <= 1e-9 * c_ex


@pytest.mark.parametrize("L,table", [
    (0, [(-1.0, 1.0), (0.0, 1.0), (0.5, 1.0), (3.0, 1.0)]),
    (1, [(-1.0, 0.0), (-0.0, 0.0), (0.0, 0.0), (0.25, 0.25), (1.0, 1.0), (7.0, 1.0)]),
    (2, [(-1.0, 0.0), (0.0, 0.0), (0.25, 0.5), (0.5, 1.0), (0.75, 1.0), (1.5, 1.5), (2.0, 2.0), (3.0, 2.0)]),
])
def test_limiter_tables(oracle, L, table):
    """None = 1; minmod = max(0, min(1, r)); superbee = max(0, min(2r, 1), min(r, 2))."""
    for r, phi in table:
        out = oracle.o_limiter(L, r)
        assert out == phi and math.copysign(1, out) == 1, (L, r, out)


def test_correctly_rounded_powers(oracle):
    """armon_cr_pow3/4 are the correctly rounded cube and fourth power (the exact rational
    rounded once). The reference's Kokkos::pow(y, 3|4) is build dependent: glibc pow() in an
    IEEE build (at most 1 ulp from the correctly rounded value, in a small fraction of
    inputs), y*y*y under its -ffast-math Release flags. The correctly rounded value is the
    build-independent choice; this pins how far it is from glibc pow()."""
    rng = random.Random(1234)
    ys = [rng.uniform(0.3, 2.5) for _ in range(20000)] + [rng.uniform(1e3, 3e4) for _ in range(5000)]
    ys += [1.0, 0.5, 2.0, 1.0 + 2**-52, 1.4285714285700001]
    mism = 0
    for y in ys:
        c3, c4 = oracle.armon_cr_pow3(y), oracle.armon_cr_pow4(y)
        assert c3 == float(Fraction(y) ** 3)
        assert c4 == float(Fraction(y) ** 4)
        for c, n in ((c3, 3), (c4, 4)):
            g = math.pow(y, n)
            if c != g:
                mism += 1
                assert ulps(c, g) <= 1.0, (y, n)
    assert mism <= 0.005 * 2 * len(ys)


# ---- invariants ---------------------------------------------------------------------------

def _field(d, name, p):
    return d.array(name).reshape(p.col_length, p.row_length)


def test_reflective_wall_zero_velocity(po, oracle):
    """At a reflective wall the mirrored ghost makes the Godunov velocity exactly zero, so the
    advection flux through the wall vanishes exactly (X_only, Sod walls on Left/Right)."""
    p = po.reference_params(oracle, "Sod", axis_splitting=po.SPLIT_X_ONLY, max_cycles=3)
    st, d, _, _, _ = po.run(oracle, p)
    assert st == 0
    g = p.nb_ghosts
    us = _field(d, "ustar", p)[g:g + p.ny]
    assert np.all(us[:, g] == 0.0)              # left wall face
    assert np.all(us[:, g + p.nx] == 0.0)       # right wall face
    for w in ("work_array_1", "work_array_2", "work_array_3", "work_array_4"):
        f = _field(d, w, p)[g:g + p.ny]
        assert np.all(f[:, g] == 0.0) and np.all(f[:, g + p.nx] == 0.0)


def test_uniform_state_is_steady(po, oracle):
    """A uniform state at rest stays uniform: fluxes equal the state, net fluxes cancel."""
    p = po.make_params(oracle, "Sod", scheme=po.SCHEME_GAD, projection=po.PROJ_EULER_2ND, nb_ghosts=4,
                       nx=16, ny=12, max_cycles=5, verbose=5, domain_size=(1.0, 1.0))
    d = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(d.c), False)
    d.array("rho")[:] = 1.0
    d.array("Emat")[:] = 2.5
    d.array("umat")[:] = 0.0
    d.array("vmat")[:] = 0.0
    lt, fdt, cyc = C.c_double(), C.c_double(), C.c_int()
    assert oracle.o_time_loop(C.byref(p), C.byref(d.c), C.byref(lt), C.byref(fdt), C.byref(cyc)) == 0
    m = d.array("domain_mask") == 1
    for f, val in (("rho", 1.0), ("Emat", 2.5)):
        a = d.array(f)[m]
        assert np.max(np.abs(a - val)) <= 4 * np.spacing(val)
    assert np.max(np.abs(d.array("umat")[m])) <= 1e-15


def test_dt_cfl_brute_force(po, oracle):
    rng = np.random.default_rng(7)
    p = po.make_params(oracle, "Sod", nx=30, ny=20, nb_ghosts=4)
    n = p.nb_cells
    u, v = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    c = rng.uniform(0.1, 3, n)
    mask = np.zeros(n)
    d = oracle.o_real_domain(C.byref(p))
    r, ir = po.ORange(), po.OInnerRange1D()
    oracle.o_dr_iter1d(C.byref(d), C.byref(r), C.byref(ir))
    g = p.nb_ghosts
    idx = np.array([[p.index_start + j * p.row_length + i for i in range(p.nx)] for j in range(p.ny)]).ravel()
    mask[idx] = 1.0
    dx, dy = 1.0 / 30, 1.0 / 20
    got = oracle.o_dt_CFL(C.byref(r), C.byref(ir), dx, dy, po.ptr(u), po.ptr(v), po.ptr(c), po.ptr(mask))
    want = min(np.min(dx / np.maximum(np.abs(u[idx] + c[idx]), np.abs(u[idx] - c[idx]))),
               np.min(dy / np.maximum(np.abs(v[idx] + c[idx]), np.abs(v[idx] - c[idx]))))
    assert got == want
    assert g == 4


def test_sod_x_y_transposition(po, oracle):
    """Sod along x and Sod along y are transposes of each other up to the X-then-Y sweep order.
    Not bitwise, and not ulp-close either: the lagged CFL step reads the sound speed of the last
    sweep's input (armon_2D.cpp:307 vs :314), which is the active sweep for Sod_y but the trivial
    one for Sod, so the dt sequences differ at first order (about 1e-4 relative)."""
    px = po.reference_params(oracle, "Sod")
    py = po.reference_params(oracle, "Sod_y")
    _, dx_, cx, _, _ = po.run(oracle, px)
    _, dy_, cy, _, _ = po.run(oracle, py)
    assert cx == cy == 45
    g = px.nb_ghosts
    rx = _field(dx_, "rho", px)[g:g + px.ny, g:g + px.nx]
    ry = _field(dy_, "rho", py)[g:g + py.ny, g:g + py.nx]
    assert np.max(np.abs(rx - ry.T)) <= 1e-3
    assert np.mean(np.abs(rx - ry.T)) <= 1e-4


def test_stale_ghost_layers_keep_init_values(po, oracle):
    """Godunov + Euler 1st order fills one ghost layer (stencil width 1) but reads two: the
    second layer keeps its init_test values, p = 0 and c = 1 (ranges.cpp:14-30, init_test.cpp:83-102)."""
    p = po.make_params(oracle, "Sod", scheme=po.SCHEME_GODUNOV, projection=po.PROJ_EULER, cfl=0.6,
                       nx=40, ny=40, verbose=5)
    assert p.stencil_width == 1 and p.nb_ghosts == 3
    st, d, cycles, _, _ = po.run(oracle, p)
    assert st == 0 and cycles > 0
    g = p.nb_ghosts
    pm = _field(d, "pmat", p)
    cm = _field(d, "cmat", p)
    assert np.all(pm[g:g + p.ny, g - 2] == 0.0) and np.all(cm[g:g + p.ny, g - 2] == 1.0)
    assert np.all(pm[g:g + p.ny, g + p.nx + 1] == 0.0) and np.all(cm[g:g + p.ny, g + p.nx + 1] == 1.0)


def test_throughput_build_follows_the_host_cpu(po, monkeypatch):
    """The CPU baseline loads the x86-64-v4 (AVX-512) build only on hosts whose CPU has every
    AVX-512 subset of that level, else the x86-64-v3 build; the checker is always the strict one."""
    import os

    monkeypatch.setattr(po, "host_has_x86_64_v4", lambda: False)
    assert os.path.basename(po.lib_path(False)) == "liboracle_fast.so"
    monkeypatch.setattr(po, "host_has_x86_64_v4", lambda: True)
    v4 = os.path.basename(po.lib_path(False))
    assert v4 == ("liboracle_fast_v4.so" if os.path.exists(os.path.join(os.path.dirname(po.__file__), "lib",
                                                                       "liboracle_fast_v4.so"))
                  else "liboracle_fast.so")
    assert os.path.basename(po.lib_path(True)) == "liboracle.so"


def test_boundary_factor_override(po, oracle):
    """The boundary factors given in the parameters replace the test's. On the corner Sedov blast
    (outside the reference's tests): with reflective walls the wall fluxes vanish exactly and the
    mass stays constant to round-off; with the Sedov test's own free-flow walls mass crosses the
    corner walls; and overriding with the test's own factors changes nothing."""
    from conftest import SEDOV_CORNER, sedov_corner_fields

    cfg = dict(scheme=po.SCHEME_GAD, projection=po.PROJ_EULER_2ND, limiter=po.LIM_MINMOD, nb_ghosts=4,
               nx=48, ny=48, cfl=0.7, max_cycles=120, max_time=1.0, verbose=5)

    def mass(p, fields=None, **kw):
        q = po.make_params(oracle, "Sedov", **dict(cfg, **kw))
        st, d, cycles, _, _ = po.run(oracle, q, init_fields=sedov_corner_fields(q) if fields is None else fields)
        assert st == 0 and cycles == 120
        m = d.array("domain_mask") == 1
        return float(np.sum(d.array("rho")[m])), d.array("rho").copy()

    walls = mass(None, **SEDOV_CORNER)
    free = mass(None, domain_origin=SEDOV_CORNER["domain_origin"], domain_size=SEDOV_CORNER["domain_size"])
    same = mass(None, domain_origin=SEDOV_CORNER["domain_origin"], domain_size=SEDOV_CORNER["domain_size"],
                bc_factors=((1.0, 1.0),) * 4)
    assert abs(walls[0] - 48 * 48) <= 1e-12 * 48 * 48
    assert abs(free[0] - 48 * 48) > 1e-6 * 48 * 48
    assert np.array_equal(same[1].view(np.uint64), free[1].view(np.uint64))


def test_sod_circ_radius_fields_reproduce_init_test(po, oracle):
    """The initial fields the circular Sod radius override writes are those of init_test at the
    reference radius 0.3, ghost cells included."""
    from conftest import sod_circ_radius_fields

    p = po.make_params(oracle, "Sod_circ", nx=96, ny=80, nb_ghosts=4, verbose=5)
    d = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(d.c), False)
    f = sod_circ_radius_fields(p, 0.3)
    for name in ("rho", "umat", "vmat", "Emat"):
        assert np.array_equal(f[name], d.array(name)), name


def test_bizarrium_fields_reproduce_init_test(po, oracle):
    """The Bizarrium initial fields of the override helper, in the reference's orientation, are
    those of init_test, ghost cells included."""
    from conftest import bizarrium_fields

    p = po.make_params(oracle, "Bizarrium", nx=90, ny=70, nb_ghosts=4, verbose=5)
    d = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(d.c), False)
    f = bizarrium_fields(p, left_moving=False)
    for name in ("rho", "umat", "vmat", "Emat"):
        assert np.array_equal(f[name], d.array(name)), name


# ---- border arrays (halo_exchange.cpp:39-92) ---------------------------------------------

BORDER_FIELDS = ("rho", "umat", "vmat", "pmat", "cmat", "gmat", "Emat")   # value_array order, :56-62


def _distinct_fields(po, oracle, p, offset=0.0):
    """Arrays of `p` whose every value names its field and cell: field f at index i holds
    (f + 1) * 1e7 + i + offset."""
    d = po.Data(oracle, p.nb_cells)
    for f, name in enumerate(BORDER_FIELDS):
        d.array(name)[:] = (f + 1) * 1e7 + np.arange(p.nb_cells) + offset
    return d


def _band_range(po, p, first_col):
    """The band of nghost columns starting at column `first_col` over the ny real rows, iterated
    as the border routines expect: inner index = column within the band (row_range_length =
    nghost), outer index = position along the side (one row per step)."""
    g = p.nb_ghosts
    r = po.ORange(0, p.ny * g)
    ir = po.OInnerRange2D(p.index_start + first_col, p.row_length, 0, g)
    band = p.index_start + np.arange(p.ny)[:, None] * p.row_length + first_col + np.arange(g)[None, :]
    return r, ir, band   # band[position, layer] -> array index


def _border(oracle, fn, p, r, ir, d, values):
    getattr(oracle, fn)(C.byref(r), C.byref(ir), p.ny, p.nb_ghosts,
                        *[getattr(d.c, n) for n in BORDER_FIELDS], values.ctypes.data_as(C.POINTER(C.c_double)))


def test_read_border_array_layout(po, oracle):
    """read_border_array fills every slot of the 7 x nghost x side_length value array exactly once,
    as layer-major blocks: value[layer][position][field], the fields in the reference's order."""
    p = po.make_params(oracle, "Sod", nx=20, ny=14, nb_ghosts=4)
    d = _distinct_fields(po, oracle, p)
    r, ir, band = _band_range(po, p, p.nx - p.nb_ghosts)
    values = np.full(7 * p.nb_ghosts * p.ny, np.nan)
    _border(oracle, "o_read_border_array", p, r, ir, d, values)
    assert not np.isnan(values).any()
    expect = np.stack([d.array(n)[band] for n in BORDER_FIELDS])        # (field, position, layer)
    assert np.array_equal(values.reshape(p.nb_ghosts, p.ny, 7), expect.transpose(2, 1, 0))
    assert len(set(values.tolist())) == values.size                       # every cell and field once


def test_border_array_round_trip_and_exchange(po, oracle):
    """write(read(x)) restores the band and touches nothing else; and the halo exchange the two
    routines exist for: the real band read on one domain's right side and written into another
    domain's left ghost band puts real column nx - nghost + l of the first into ghost column
    -nghost + l of the second."""
    p = po.make_params(oracle, "Sod", nx=20, ny=14, nb_ghosts=4)
    g = p.nb_ghosts
    a = _distinct_fields(po, oracle, p)
    before = {n: a.array(n).copy() for n in BORDER_FIELDS}
    r, ir, band = _band_range(po, p, p.nx - g)
    values = np.zeros(7 * g * p.ny)
    _border(oracle, "o_read_border_array", p, r, ir, a, values)
    for n in BORDER_FIELDS:
        a.array(n)[band] = -1.0
    _border(oracle, "o_write_border_array", p, r, ir, a, values)
    for n in BORDER_FIELDS:
        assert np.array_equal(a.array(n), before[n]), n
    b = _distinct_fields(po, oracle, p, offset=0.5)
    b_before = {n: b.array(n).copy() for n in BORDER_FIELDS}
    rg, irg, ghost_band = _band_range(po, p, -g)
    _border(oracle, "o_write_border_array", p, rg, irg, b, values)
    for n in BORDER_FIELDS:
        got = b.array(n)
        assert np.array_equal(got[ghost_band], before[n][band]), n
        untouched = np.ones(p.nb_cells, bool)
        untouched[ghost_band.ravel()] = False
        assert np.array_equal(got[untouched], b_before[n][untouched]), n


# ---- axis splitting (parameters.cpp:212-255) -----------------------------------------------

def test_split_axes_properties(po, oracle):
    """Per cycle, each axis's time factors sum to one (or zero for the axis a one-axis splitting
    leaves out); Sequential is always X then Y; SequentialSym and Strang start with Y on even
    cycles and with X on odd ones; Strang is a palindrome of half steps around a full one."""
    for split in range(5):
        for cycle in range(6):
            p = po.make_params(oracle, "Sod", axis_splitting=split)
            p.cycle = cycle
            axes, f = (C.c_int * 3)(), (C.c_double * 3)()
            n = oracle.o_split_axes(C.byref(p), axes, f)
            seq = [(axes[k], f[k]) for k in range(n)]
            sx = sum(w for a, w in seq if a == po.AXIS_X)
            sy = sum(w for a, w in seq if a == po.AXIS_Y)
            assert (sx, sy) == ((1.0, 0.0) if split == po.SPLIT_X_ONLY else
                                (0.0, 1.0) if split == po.SPLIT_Y_ONLY else (1.0, 1.0)), (split, cycle, seq)
            first = po.AXIS_Y if cycle % 2 == 0 else po.AXIS_X
            if split == po.SPLIT_SEQUENTIAL:
                assert [a for a, _ in seq] == [po.AXIS_X, po.AXIS_Y]
            elif split in (po.SPLIT_SEQUENTIAL_SYM, po.SPLIT_STRANG):
                assert seq[0][0] == first and seq[1][0] != first
            if split == po.SPLIT_STRANG:
                assert seq[0] == seq[2] == (first, 0.5) and seq[1][1] == 1.0


@pytest.mark.parametrize("scheme,proj", [(1, 2), (0, 1)])
def test_one_axis_splittings_transpose(po, oracle, scheme, proj):
    """X_only on Sod (along x, reflective left/right) and Y_only on Sod_y (along y, reflective
    top/bottom) are exact transposes on a square domain: same cycle count, same final dt, every
    field bit for bit transposed (u and v exchanged). With one sweep per cycle the lagged CFL reads
    the sound speed of the same sweep in both, unlike the two-axis splittings."""
    cfg = dict(scheme=scheme, projection=proj, limiter=1, nb_ghosts=5, nx=48, ny=48, max_cycles=400, verbose=5)
    px = po.make_params(oracle, "Sod", axis_splitting=po.SPLIT_X_ONLY, **cfg)
    py = po.make_params(oracle, "Sod_y", axis_splitting=po.SPLIT_Y_ONLY, **cfg)
    sx, dx_, cx, fx, _ = po.run(oracle, px)
    sy, dy_, cy, fy, _ = po.run(oracle, py)
    assert sx == sy == 0 and cx == cy and cx > 20 and fx == fy
    g = px.nb_ghosts
    for a, b in (("rho", "rho"), ("Emat", "Emat"), ("pmat", "pmat"), ("umat", "vmat"), ("vmat", "umat")):
        fa = _field(dx_, a, px)[g:g + px.ny, g:g + px.nx]
        fb = _field(dy_, b, py)[g:g + py.ny, g:g + py.nx]
        assert np.array_equal(fa, fb.T), (a, b)


# ---- time step rules (armon_2D.cpp:212-252) -----------------------------------------------

def _brute_cfl(d, p):
    """min over the real cells of min(dx / max|u +- c|, dy / max|v +- c|) (reductions.cpp:18-20)."""
    m = d.array("domain_mask") == 1
    u, v, c = d.array("umat")[m], d.array("vmat")[m], d.array("cmat")[m]
    dx, dy = p.domain_size[0] / p.nx, p.domain_size[1] / p.ny
    return min(np.min(dx / np.maximum(np.abs(u + c), np.abs(u - c))),
               np.min(dy / np.maximum(np.abs(v + c), np.abs(v - c))))


def _stepped_run(po, oracle, p):
    """The time loop of armon_2D.cpp:327-386 cycle by cycle: the dt each cycle used and the brute
    force CFL bound of the state each cycle started from. Ends where o_time_loop ends."""
    d = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(d.c), False)
    p.cycle, p.time, p.current_cycle_dt, p.next_cycle_dt = 0, 0.0, 0.0, 0.0
    oracle.o_update_axis(C.byref(p), po.AXIS_X)
    oracle.o_update_EOS(C.byref(p), C.byref(d.c))
    dts, bounds = [], []
    while p.time < p.max_time and p.cycle < p.max_cycles:
        bounds.append(_brute_cfl(d, p))
        assert oracle.o_solver_cycle(C.byref(p), C.byref(d.c)) == 0
        dts.append(p.current_cycle_dt)
        p.time += p.current_cycle_dt
        p.cycle += 1
    return dts, bounds, d


CYCLE_CFG = dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=40, ny=36, max_cycles=40, verbose=5)


def test_lagged_cfl_time_step(po, oracle):
    """Every cycle: dt(0) = dt(1) = cfl * CFL(state 0), then dt(c + 1) = min(cfl * CFL(state c),
    1.05 dt(c)), the bound taken from the state the cycle starts from; the same run as o_time_loop."""
    p = po.make_params(oracle, "Sedov", **CYCLE_CFG)
    dts, bounds, d = _stepped_run(po, oracle, p)
    assert dts[0] == dts[1] == p.cfl * bounds[0]
    for c in range(1, len(dts) - 1):
        assert dts[c + 1] == min(p.cfl * bounds[c], 1.05 * dts[c]), c
    q = po.make_params(oracle, "Sedov", **CYCLE_CFG)
    st, od, cycles, final_dt, _ = po.run(oracle, q)
    assert st == 0 and cycles == len(dts) and np.array_equal(od.array("rho"), d.array("rho"))


def test_dt_on_even_cycles(po, oracle):
    """dt_on_even_cycles: the bound is taken on even cycles only and the dt changes only between
    an even cycle and the next: dt(2k + 2) = dt(2k + 1), dt(2k + 1) = min(cfl * CFL(state 2k),
    1.05 dt(2k)) for k >= 1, and dt(0) = dt(1) = dt(2) = cfl * CFL(state 0)."""
    p = po.make_params(oracle, "Sedov", dt_on_even_cycles=1, **CYCLE_CFG)
    dts, bounds, _ = _stepped_run(po, oracle, p)
    assert len(dts) == CYCLE_CFG["max_cycles"]
    assert dts[0] == dts[1] == dts[2] == p.cfl * bounds[0]
    for k in range(1, (len(dts) - 1) // 2):
        assert dts[2 * k + 1] == min(p.cfl * bounds[2 * k], 1.05 * dts[2 * k]), k
        if 2 * k + 2 < len(dts):
            assert dts[2 * k + 2] == dts[2 * k + 1], k
    assert len(set(dts)) > 5   # the bound does move


def test_constant_time_step(po, oracle):
    """cst_dt: every cycle uses Dt, so a run to t_final takes ceil(t_final / Dt) cycles (Dt a power
    of two: the time sums exactly)."""
    Dt = 2.0 ** -10
    p = po.make_params(oracle, "Sod", cst_dt=1, Dt=Dt, scheme=1, projection=2, nb_ghosts=4, nx=32, ny=24,
                       max_cycles=10000, verbose=5)
    dts, _, _ = _stepped_run(po, oracle, p)
    assert all(dt == Dt for dt in dts)
    assert len(dts) == int(np.ceil(p.max_time / Dt)) == 205


def test_md_iteration_build(po):
    """The throughput build iterating as the reference's USE_MD_ITER + BALANCE_MD_ITER build (row
    tiles balanced over the threads) covers the same cells as the 2D-iteration build: the Sod KAT
    takes the reference's 45 cycles and both builds' states agree to fast-math rounding."""
    md = po.load(strict=False, iteration="md")
    fast = po.load(strict=False)
    assert md is not fast and "md" in os.path.basename(po.lib_path(False, "md"))
    out = {}
    for name, lib in (("md", md), ("2d", fast)):
        p = po.reference_params(lib, "Sod")
        st, d, cycles, _, _ = po.run(lib, p)
        assert st == 0 and cycles == 45, (name, cycles)
        out[name] = d.array("rho")[d.array("domain_mask") == 1].copy()
    assert np.max(np.abs(out["md"] - out["2d"])) <= 1e-12 * np.max(np.abs(out["2d"]))
