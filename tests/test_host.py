"""CPU tests of the product library: it loads, exports every symbol its headers declare,
its host-side parameters/domains reproduce the reference's KATs and agree with the oracle,
and argument errors are reported through raise_exception_handler before any GPU work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in ("kernels.h", "armon.h"):
        text = open(os.path.join(ROOT, "include", "armon", h)).read()
        text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
        names += re.findall(r"ARMON_EXPORT\s+extern\s+\w+\s*\(\s*\*\s*(\w+)\s*\)", text)
        names += re.findall(r"ARMON_EXPORT\s+(?!extern)[\w\s\*]*?\b(\w+)\s*\(", text)
    return sorted(set(names))


def test_headers_declare_the_reference_entry_points():
    names = declared_symbols()
    ref = ["acoustic", "acoustic_GAD", "perfect_gas_EOS", "bizarrium_EOS", "cell_update", "euler_projection",
           "advection_first_order", "advection_second_order", "boundary_conditions", "read_border_array",
           "write_border_array", "init_test", "dt_CFL", "conservation_vars", "idx_size", "is_uidx_signed",
           "flt_size", "raise_exception_handler"]
    for r in ref:
        assert r in names, r


def test_library_exports_every_declared_symbol(armon):
    lib = armon.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing
    # the binding knows every declared function
    known = set(armon.KERNEL_SYMBOLS) | set(armon.EXTENSION_SYMBOLS) | {"raise_exception_handler"}
    assert set(declared_symbols()) <= known


def test_float_library_exports_the_kernel_abi(armon):
    """lib/libarmon_f32.so (flt_t = float): every kernels.h symbol, the plumbing, no solver."""
    lib = armon.load_f32()
    kernel_names = [n for n in declared_symbols() if n in armon.KERNEL_SYMBOLS or n == "raise_exception_handler"]
    assert len(kernel_names) == len(armon.KERNEL_SYMBOLS) + 1
    missing = [n for n in kernel_names + list(armon.PLUMBING_SYMBOLS) if not hasattr(lib, n)]
    assert not missing
    assert not hasattr(lib, "armon_solver_create")
    assert lib.flt_size() == 4 and lib.idx_size() == 8 and lib.is_uidx_signed() is False


def test_selftest_library_is_separate_from_the_solver(armon):
    """lib/libarmon_selftest.so holds the device self-tests of the shortened FP64 sequences (the
    shared-divisor quotients and the square root); the solver library exports neither."""
    path = os.path.join(ROOT, "armon-kokkos_amd", "lib", "libarmon_selftest.so")
    selftest = C.CDLL(path)
    for name in ("armon_selftest_shared_divisor", "armon_selftest_sqrt"):
        assert hasattr(selftest, name), name
        assert not hasattr(armon.load(), name), name


def test_abi_queries(armon):
    lib = armon.load()
    assert lib.idx_size() == 8
    assert lib.flt_size() == 8
    assert lib.is_uidx_signed() is False
    assert C.sizeof(armon.Range) == 16 and C.sizeof(armon.InnerRange2D) == 32
    assert C.sizeof(armon.View) == 24


def _default(armon):
    # armon_2D_tests.cpp:61-80
    return armon.make_params("Sod", scheme=armon.SCHEME_GAD, limiter=1, projection=armon.PROJ_EULER_2ND,
                             axis_splitting=0, max_cycles=100, verbose=1, nx=100, ny=5, nb_ghosts=5,
                             engine=armon.ENGINE_STAGED)


def test_indexing_kats(armon):
    lib = armon.load()
    p = _default(armon)
    P = C.byref(p)
    assert lib.armon_params_check(P) == 1
    assert lib.armon_real_domain(P).as_tuple() == (555, 110, 995, 0, 1, 99)
    assert lib.armon_domain_fluxes(P).as_tuple() == (555, 110, 995, -2, 1, 102)
    assert lib.armon_domain_advection(P).as_tuple() == (555, 110, 995, 0, 1, 100)
    assert lib.armon_complete_domain(P).as_tuple() == (5, 110, 1545, -5, 1, 104)
    d = lib.armon_real_domain(P)
    r, i2 = armon.Range(), armon.InnerRange2D()
    lib.armon_domain_iter2D(C.byref(d), C.byref(r), C.byref(i2))
    assert (r.start, r.end, i2.main_range_start, i2.main_range_step, i2.row_range_start, i2.row_range_length) == \
        (0, 500, 555, 110, 0, 100)
    i1 = armon.InnerRange1D()
    lib.armon_domain_iter1D(C.byref(d), C.byref(r), C.byref(i1))
    assert (r.start, r.end, i1.start) == (0, 540, 555)
    disp = C.c_int()
    exp = {armon.SIDE_BOTTOM: (445, 544, 100, 110), armon.SIDE_LEFT: (554, 994, 5, 1),
           armon.SIDE_RIGHT: (655, 1095, 5, -1), armon.SIDE_TOP: (1105, 1204, 100, -110)}
    for side, (b, e, n, ds) in exp.items():
        dd = lib.armon_boundary_conditions_domain(P, side, C.byref(disp))
        assert (lib.armon_domain_begin(C.byref(dd)), lib.armon_domain_end(C.byref(dd)),
                lib.armon_domain_length(C.byref(dd)), disp.value) == (b, e, n, ds)
    assert [lib.armon_index_1D(P, i, j) for i, j in ((-1, 0), (0, -1), (100, 0), (0, 5))] == [554, 445, 655, 1105]


@pytest.mark.parametrize("test", ["Sod", "Sod_y", "Sod_circ", "Bizarrium", "Sedov"])
@pytest.mark.parametrize("scheme,proj", [(0, 1), (1, 1), (0, 2), (1, 2)])
def test_params_match_oracle(armon, po, oracle, test, scheme, proj):
    """Params::init, min_ghosts, update_axis and every iteration domain agree with the oracle."""
    lib = armon.load()
    kw = dict(scheme=scheme, projection=proj, nx=37, ny=23, nb_ghosts=5)
    p = armon.make_params(test, engine=armon.ENGINE_STAGED, **kw)
    q = po.make_params(oracle, test, **kw)
    for f in ("cfl", "max_time", "stencil_width", "row_length", "col_length", "nb_cells", "ideb", "ifin",
              "index_start", "s", "dx"):
        assert getattr(p, f) == getattr(q, f), f
    assert tuple(p.domain_size) == tuple(q.domain_size) and tuple(p.domain_origin) == tuple(q.domain_origin)
    assert lib.armon_min_ghosts(C.byref(p)) == oracle.o_min_ghosts(C.byref(q))
    for axis in (0, 1):
        lib.armon_update_axis(C.byref(p), axis)
        oracle.o_update_axis(C.byref(q), axis)
        assert (p.s, p.dx) == (q.s, q.dx)
        for fa, fo in (("armon_real_domain", "o_real_domain"), ("armon_domain_fluxes", "o_domain_fluxes"),
                       ("armon_domain_cell_update", "o_domain_cell_update"),
                       ("armon_domain_advection", "o_domain_advection"),
                       ("armon_complete_domain", "o_complete_domain")):
            assert getattr(lib, fa)(C.byref(p)).as_tuple() == getattr(oracle, fo)(C.byref(q)).as_tuple()
    for side in range(4):
        da, do = C.c_int(), C.c_int()
        a = lib.armon_boundary_conditions_domain(C.byref(p), side, C.byref(da))
        b = oracle.o_boundary_conditions_domain(C.byref(q), side, C.byref(do))
        assert a.as_tuple() == b.as_tuple() and da.value == do.value
        ua, va, uo, vo = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        lib.armon_test_boundary_condition(p.test, side, C.byref(ua), C.byref(va))
        oracle.o_test_bc(q.test, side, C.byref(uo), C.byref(vo))
        assert (ua.value, va.value) == (uo.value, vo.value)
    for split in range(5):
        for cyc in (0, 1, 2):
            p.axis_splitting = q.axis_splitting = split
            p.cycle = q.cycle = cyc
            aa, fa_ = (C.c_int * 3)(), (C.c_double * 3)()
            ao, fo_ = (C.c_int * 3)(), (C.c_double * 3)()
            na = lib.armon_split_axes(C.byref(p), aa, fa_)
            no = oracle.o_split_axes(C.byref(q), ao, fo_)
            assert na == no and list(aa)[:na] == list(ao)[:no] and list(fa_)[:na] == list(fo_)[:no]


def test_params_check_errors(armon):
    lib = armon.load()
    p = armon.make_params("Sod", projection=armon.PROJ_NONE, engine=armon.ENGINE_STAGED)
    assert lib.armon_params_check(C.byref(p)) == 0
    p = armon.make_params("Sod", cst_dt=1, Dt=0.0, engine=armon.ENGINE_STAGED)
    assert lib.armon_params_check(C.byref(p)) == 0
    p = armon.make_params("Sod", stencil_width=6, nb_ghosts=5, engine=armon.ENGINE_STAGED)
    assert lib.armon_params_check(C.byref(p)) == 0
    # the fused engine needs the ghost layers the stencil reads
    p = armon.make_params("Sod", scheme=1, projection=2, nb_ghosts=3, stencil_width=3, engine=armon.ENGINE_FUSED)
    assert lib.armon_params_check(C.byref(p)) == 0
    p = armon.make_params("Sod", scheme=1, projection=2, nb_ghosts=4, engine=armon.ENGINE_FUSED)
    assert lib.armon_params_check(C.byref(p)) == 1
    p = armon.make_params("Sod", scheme=1, projection=2, nb_ghosts=4, engine=armon.ENGINE_FUSED, y_run_rows=-1)
    assert lib.armon_params_check(C.byref(p)) == 0


def test_params_init_rejects_int_overflow(armon):
    """The indexing fields are ints: a domain of 2^31 cells or more (ghosts included) is refused
    by armon_params_init instead of wrapping around."""
    armon.make_params("Sod", nx=46332, ny=46332, nb_ghosts=4)          # 46340^2 < 2^31
    with pytest.raises(armon.ArmonError):
        armon.make_params("Sod", nx=46333, ny=46333, nb_ghosts=4)      # 46341^2 >= 2^31
    with pytest.raises(armon.ArmonError):
        armon.make_params("Sod", nx=1 << 20, ny=1 << 12, nb_ghosts=4)


def test_sedov_test_option(armon):
    """test_cases.h:118-120: r = hypot(dx, dy)/sqrt(2), i.e. dx for square cells."""
    lib = armon.load()
    p = armon.make_params("Sedov", nx=64, ny=64, nb_ghosts=4)
    assert abs(lib.armon_test_option(C.byref(p)) - 2.0 / 64) <= 2e-17


# ---- error paths (validated on the host, nothing is launched) ------------------------------

def _views(armon, n, names, extent=None):
    arrs = {k: np.zeros(1) for k in names}
    return {k: armon.View(arrs[k].ctypes.data, n if extent is None else extent, f"Armon_{k}".encode())
            for k in names}, arrs


def test_invalid_limiter_reports(armon):
    lib = armon.load()
    armon.errors()
    v, keep = _views(armon, 100, ["ustar", "pstar", "rho", "umat", "pmat", "cmat"])
    r, ir = armon.Range(0, 10), armon.InnerRange2D(20, 10, 0, 10)
    lib.acoustic_GAD(C.byref(r), C.byref(ir), 1, 0.1, 0.1, C.byref(v["ustar"]), C.byref(v["pstar"]),
                     C.byref(v["rho"]), C.byref(v["umat"]), C.byref(v["pmat"]), C.byref(v["cmat"]), 7)
    errs = armon.errors()
    assert len(errs) == 1 and errs[0][0] == "acoustic_GAD"
    assert "std::out_of_range" in errs[0][1] and "Invalid limiter index: 7" in errs[0][1]


def test_invalid_test_reports(armon):
    lib = armon.load()
    armon.errors()
    names = ["x", "y", "rho", "Emat", "umat", "vmat", "domain_mask", "pmat", "cmat", "gmat", "ustar", "pstar"]
    v, keep = _views(armon, 100, names)
    r, ir = armon.Range(0, 10), armon.InnerRange2D(0, 10, 0, 10)
    lib.init_test(C.byref(r), C.byref(ir), 10, 0, 10, 1, 1.0, 1.0, 0.0, 0.0, 0, 0, 10, 1,
                  *[C.byref(v[k]) for k in names], 9, False, 0.0)
    errs = armon.errors()
    assert len(errs) == 1 and "Invalid test case index: 9" in errs[0][1]


def test_wrong_view_order_reports(armon):
    """utils.h:80-93: the view label must end with the parameter name."""
    lib = armon.load()
    armon.errors()
    v, keep = _views(armon, 100, ["ustar", "pstar", "rho", "umat", "pmat", "cmat"])
    r, ir = armon.Range(0, 10), armon.InnerRange2D(20, 10, 0, 10)
    # pstar passed where ustar is expected
    lib.acoustic(C.byref(r), C.byref(ir), 1, C.byref(v["pstar"]), C.byref(v["ustar"]), C.byref(v["rho"]),
                 C.byref(v["umat"]), C.byref(v["pmat"]), C.byref(v["cmat"]))
    errs = armon.errors()
    assert len(errs) == 1 and "std::invalid_argument" in errs[0][1] and "Wrong view order" in errs[0][1]


def test_out_of_bounds_reports(armon):
    """Every index a kernel would touch is checked against the view extent before launch."""
    lib = armon.load()
    armon.errors()
    v, keep = _views(armon, 29, ["ustar", "pstar", "rho", "umat", "pmat", "cmat"])
    r, ir = armon.Range(0, 10), armon.InnerRange2D(20, 10, 0, 10)   # touches 19..29
    lib.acoustic(C.byref(r), C.byref(ir), 1, C.byref(v["ustar"]), C.byref(v["pstar"]), C.byref(v["rho"]),
                 C.byref(v["umat"]), C.byref(v["pmat"]), C.byref(v["cmat"]))
    errs = armon.errors()
    assert len(errs) == 1 and "out of bounds" in errs[0][1]
    # reading i - s below zero
    v, keep = _views(armon, 1000, ["ustar", "pstar", "rho", "umat", "pmat", "cmat"])
    r, ir = armon.Range(0, 10), armon.InnerRange2D(0, 10, 0, 10)
    lib.acoustic(C.byref(r), C.byref(ir), 1, C.byref(v["ustar"]), C.byref(v["pstar"]), C.byref(v["rho"]),
                 C.byref(v["umat"]), C.byref(v["pmat"]), C.byref(v["cmat"]))
    errs = armon.errors()
    assert len(errs) == 1 and "out of bounds" in errs[0][1] and "[-1," in errs[0][1]


def test_float_library_error_paths(armon):
    """The float build validates on the host like the FP64 one and reports through its own
    raise_exception_handler: an invalid limiter, a wrong view order, an out-of-bounds range."""
    lib = armon.load_f32()
    armon.errors()
    names = ["ustar", "pstar", "rho", "umat", "pmat", "cmat"]
    v, keep = _views(armon, 100, names)
    r, ir = armon.Range(0, 10), armon.InnerRange2D(20, 10, 0, 10)
    lib.acoustic_GAD(C.byref(r), C.byref(ir), 1, 0.1, 0.1, *[C.byref(v[k]) for k in names], 7)
    errs = armon.errors()
    assert len(errs) == 1 and errs[0][0] == "acoustic_GAD" and "Invalid limiter index: 7" in errs[0][1]
    lib.acoustic(C.byref(r), C.byref(ir), 1, C.byref(v["pstar"]), C.byref(v["ustar"]),
                 *[C.byref(v[k]) for k in names[2:]])
    errs = armon.errors()
    assert len(errs) == 1 and "Wrong view order" in errs[0][1]
    v, keep = _views(armon, 29, names)
    lib.acoustic(C.byref(r), C.byref(ir), 1, *[C.byref(v[k]) for k in names])
    errs = armon.errors()
    assert len(errs) == 1 and "out of bounds" in errs[0][1]


def test_boundary_conditions_bounds(armon):
    lib = armon.load()
    armon.errors()
    names = ["rho", "umat", "vmat", "pmat", "cmat", "gmat", "Emat"]
    v, keep = _views(armon, 100, names)
    # 4 layers outwards from index 2 with disp 1: writes 2, 1, 0, -1 -> out of bounds
    r, ir = armon.Range(0, 1), armon.InnerRange1D(2, 1)
    lib.boundary_conditions(C.byref(r), C.byref(ir), 4, 1, -1.0, 1.0, *[C.byref(v[k]) for k in names])
    errs = armon.errors()
    assert len(errs) == 1 and "out of bounds" in errs[0][1]


def test_cli_help():
    import subprocess
    import __graft_entry__ as g
    exe = os.path.join(g.PKG_DIR, "bin", "armon_hip_2D")
    out = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "--engine" in out.stdout and "--cells" in out.stdout
    out = subprocess.run([exe, "--bogus"], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0 and "Wrong option" in out.stderr


def test_fused_engine_refuses_ghost_output(armon):
    """The fused engine keeps no ghost cells in memory, so writing them (write_ghosts with
    write_output) is refused for it and accepted for the staged engine."""
    lib = armon.load()
    for engine, ok in ((armon.ENGINE_FUSED, 0), (armon.ENGINE_STAGED, 1)):
        p = armon.make_params("Sod", nx=32, ny=32, nb_ghosts=4, engine=engine, write_output=1, write_ghosts=1)
        assert lib.armon_params_check(C.byref(p)) == ok
        p.write_ghosts = 0
        assert lib.armon_params_check(C.byref(p)) == 1


def test_single_precision_switch_selects_the_float_library(tmp_path):
    """ARMON_SINGLE_PRECISION (the reference's USE_SINGLE_PRECISION, common.h:8-12) makes flt_t
    float; a C caller built that way and linked to lib/libarmon_f32.so, or built without it and
    linked to lib/libarmon.so, sees its own flt_t size in flt_size()."""
    import subprocess

    src = tmp_path / "use.c"
    src.write_text('#include <stdio.h>\n#include "armon/kernels.h"\n'
                   'int main(void) { printf("%d %d\\n", (int) sizeof(flt_t), flt_size());'
                   ' return flt_size() == (int) sizeof(flt_t) ? 0 : 1; }\n')
    inc = os.path.join(ROOT, "include")
    libdir = os.path.join(ROOT, "armon-kokkos_amd", "lib")
    for flags, lib, size in (([], "armon", 8), (["-DARMON_SINGLE_PRECISION=1"], "armon_f32", 4)):
        exe = tmp_path / f"use_{lib}"
        cc = subprocess.run(["gcc", "-std=c11", *flags, "-I", inc, str(src), "-o", str(exe), "-L", libdir,
                             f"-l{lib}", f"-Wl,-rpath,{libdir}"], capture_output=True, text=True)
        assert cc.returncode == 0, cc.stderr
        run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
        assert run.returncode == 0 and run.stdout.split() == [str(size), str(size)], run.stdout + run.stderr
