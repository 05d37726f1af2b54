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
