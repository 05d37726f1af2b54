"""ctypes binding of the CPU oracle (oracle/armon_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``. The product library never uses
it. ``load(strict=True)`` gives the IEEE-strict checker, ``strict=False`` the
fast-math throughput build used as the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

# Enumerations (same values as the reference)
LIM_NONE, LIM_MINMOD, LIM_SUPERBEE = 0, 1, 2
TEST_SOD, TEST_SOD_Y, TEST_SOD_CIRC, TEST_BIZARRIUM, TEST_SEDOV = 0, 1, 2, 3, 4
SCHEME_GODUNOV, SCHEME_GAD = 0, 1
PROJ_NONE, PROJ_EULER, PROJ_EULER_2ND = 0, 1, 2
SPLIT_SEQUENTIAL, SPLIT_SEQUENTIAL_SYM, SPLIT_STRANG, SPLIT_X_ONLY, SPLIT_Y_ONLY = 0, 1, 2, 3, 4
AXIS_X, AXIS_Y = 0, 1
SIDE_LEFT, SIDE_RIGHT, SIDE_TOP, SIDE_BOTTOM = 0, 1, 2, 3

TEST_NAMES = {"Sod": 0, "Sod_y": 1, "Sod_circ": 2, "Bizarrium": 3, "Sedov": 4}


class ORange(C.Structure):
    _fields_ = [("start", C.c_int64), ("end", C.c_int64)]


class OInnerRange1D(C.Structure):
    _fields_ = [("start", C.c_int64), ("step", C.c_int64)]


class OInnerRange2D(C.Structure):
    _fields_ = [("main_range_start", C.c_int64), ("main_range_step", C.c_int64),
                ("row_range_start", C.c_int64), ("row_range_length", C.c_uint64)]


class ODomainRange(C.Structure):
    _fields_ = [("col_start", C.c_int64), ("col_step", C.c_int64), ("col_end", C.c_int64),
                ("row_start", C.c_int64), ("row_step", C.c_int64), ("row_end", C.c_int64)]

    def as_tuple(self):
        return (self.col_start, self.col_step, self.col_end, self.row_start, self.row_step, self.row_end)


class OParams(C.Structure):
    _fields_ = [
        ("test", C.c_int), ("scheme", C.c_int), ("riemann", C.c_int), ("projection", C.c_int),
        ("limiter", C.c_int), ("current_axis", C.c_int), ("axis_splitting", C.c_int),
        ("nb_ghosts", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
        ("dx", C.c_double), ("cfl", C.c_double), ("Dt", C.c_double),
        ("stencil_width", C.c_int), ("cst_dt", C.c_int), ("dt_on_even_cycles", C.c_int),
        ("domain_size", C.c_double * 2), ("domain_origin", C.c_double * 2),
        ("row_length", C.c_int), ("col_length", C.c_int), ("nb_cells", C.c_int), ("ideb", C.c_int),
        ("ifin", C.c_int), ("index_start", C.c_int), ("idx_row", C.c_int), ("idx_col", C.c_int),
        ("s", C.c_int),
        ("cycle", C.c_int), ("cycle_dt", C.c_double), ("current_cycle_dt", C.c_double),
        ("next_cycle_dt", C.c_double), ("time", C.c_double),
        ("max_cycles", C.c_int), ("max_time", C.c_double), ("verbose", C.c_int),
        ("bc_override", C.c_int), ("bc_factors", (C.c_double * 2) * 4),
    ]


_DATA_FIELDS = ["x", "y", "rho", "umat", "vmat", "Emat", "pmat", "cmat", "gmat", "ustar", "pstar",
                "work_array_1", "work_array_2", "work_array_3", "work_array_4", "domain_mask"]


class OData(C.Structure):
    _fields_ = [(f, C.POINTER(C.c_double)) for f in _DATA_FIELDS]


_P = C.POINTER(C.c_double)
_libs: dict = {}


class OTestState(C.Structure):
    """High/low region states of a test case (init_test.cpp:12-48)."""
    _fields_ = [(n, C.c_double) for n in ("high_rho", "low_rho", "high_E", "low_E",
                                          "high_u", "low_u", "high_v", "low_v")]


def _proto(lib):
    i64, d = C.c_int64, C.c_double
    R, I1, I2 = C.POINTER(ORange), C.POINTER(OInnerRange1D), C.POINTER(OInnerRange2D)
    sig = {
        "o_scale_index_1d": (i64, [I1, C.c_uint64]),
        "o_scale_index_2d": (i64, [I2, C.c_uint64]),
        "armon_cr_pow3": (d, [d]),
        "armon_cr_pow4": (d, [d]),
        "o_limiter": (d, [C.c_int, d]),
        "o_acoustic": (None, [R, I2, i64, _P, _P, _P, _P, _P, _P]),
        "o_acoustic_GAD": (C.c_int, [R, I2, i64, d, d, _P, _P, _P, _P, _P, _P, C.c_int]),
        "o_perfect_gas_EOS": (None, [R, I2, d, _P, _P, _P, _P, _P, _P, _P]),
        "o_bizarrium_EOS": (None, [R, I2, _P, _P, _P, _P, _P, _P, _P]),
        "o_cell_update": (None, [R, I2, i64, d, d, _P, _P, _P, _P, _P]),
        "o_euler_projection": (None, [R, I2, i64, d, d, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
        "o_advection_first_order": (None, [R, I2, i64, d, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
        "o_advection_second_order": (None, [R, I2, i64, d, d, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
        "o_boundary_conditions": (None, [R, I1, i64, i64, d, d, _P, _P, _P, _P, _P, _P, _P]),
        "o_read_border_array": (None, [R, I2, i64, i64, _P, _P, _P, _P, _P, _P, _P, _P]),
        "o_write_border_array": (None, [R, I2, i64, i64, _P, _P, _P, _P, _P, _P, _P, _P]),
        "o_init_test": (C.c_int, [R, I2, i64, i64, i64, i64, d, d, d, d, i64, i64, i64, i64,
                                  _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int, C.c_bool, d]),
        "o_reference_test_state": (None, [C.c_int, d, C.POINTER(OTestState)]),
        "o_init_test_states": (C.c_int, [R, I2, i64, i64, i64, i64, d, d, d, d, i64, i64, i64, i64,
                                         _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int, C.c_bool, d,
                                         C.POINTER(OTestState)]),
        "o_dt_CFL": (d, [R, I1, d, d, _P, _P, _P, _P]),
        "o_conservation_vars": (None, [R, I1, d, _P, _P, _P, _P, _P]),
        "o_params_defaults": (None, [C.POINTER(OParams)]),
        "o_min_ghosts": (C.c_int, [C.POINTER(OParams)]),
        "o_params_init": (None, [C.POINTER(OParams)]),
        "o_params_check": (C.c_int, [C.POINTER(OParams)]),
        "o_update_axis": (None, [C.POINTER(OParams), C.c_int]),
        "o_split_axes": (C.c_int, [C.POINTER(OParams), C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "o_test_bc": (None, [C.c_int, C.c_int, _P, _P]),
        "o_index_1D": (C.c_int64, [C.POINTER(OParams), C.c_int, C.c_int]),
        "o_real_domain": (ODomainRange, [C.POINTER(OParams)]),
        "o_domain_fluxes": (ODomainRange, [C.POINTER(OParams)]),
        "o_domain_cell_update": (ODomainRange, [C.POINTER(OParams)]),
        "o_domain_advection": (ODomainRange, [C.POINTER(OParams)]),
        "o_complete_domain": (ODomainRange, [C.POINTER(OParams)]),
        "o_boundary_conditions_domain": (ODomainRange, [C.POINTER(OParams), C.c_int, C.POINTER(C.c_int)]),
        "o_dr_col_length": (C.c_uint64, [C.POINTER(ODomainRange)]),
        "o_dr_row_length": (C.c_uint64, [C.POINTER(ODomainRange)]),
        "o_dr_length": (C.c_uint64, [C.POINTER(ODomainRange)]),
        "o_dr_begin": (C.c_int64, [C.POINTER(ODomainRange)]),
        "o_dr_end": (C.c_int64, [C.POINTER(ODomainRange)]),
        "o_dr_iter1d": (None, [C.POINTER(ODomainRange), R, I1]),
        "o_dr_direct_iter1d": (None, [C.POINTER(ODomainRange), R, I1]),
        "o_dr_iter2d": (None, [C.POINTER(ODomainRange), R, I2]),
        "o_data_alloc": (C.c_int, [C.POINTER(OData), C.c_int64]),
        "o_data_free": (None, [C.POINTER(OData)]),
        "o_test_option": (d, [C.POINTER(OParams)]),
        "o_init_test_params": (None, [C.POINTER(OParams), C.POINTER(OData), C.c_bool]),
        "o_init_test_params_states": (None, [C.POINTER(OParams), C.POINTER(OData), C.POINTER(OTestState)]),
        "o_update_EOS": (None, [C.POINTER(OParams), C.POINTER(OData)]),
        "o_boundary_conditions_all": (None, [C.POINTER(OParams), C.POINTER(OData)]),
        "o_numerical_fluxes": (None, [C.POINTER(OParams), C.POINTER(OData)]),
        "o_cell_update_params": (None, [C.POINTER(OParams), C.POINTER(OData)]),
        "o_projection_remap": (None, [C.POINTER(OParams), C.POINTER(OData)]),
        "o_local_time_step": (d, [C.POINTER(OParams), C.POINTER(OData), d]),
        "o_time_step": (C.c_int, [C.POINTER(OParams), C.POINTER(OData)]),
        "o_solver_cycle": (C.c_int, [C.POINTER(OParams), C.POINTER(OData)]),
        "o_conservation_vars_params": (None, [C.POINTER(OParams), C.POINTER(OData), _P, _P]),
        "o_time_loop": (C.c_int, [C.POINTER(OParams), C.POINTER(OData), _P, _P, C.POINTER(C.c_int)]),
        "o_num_threads": (C.c_int, []),
        "o_set_num_threads": (None, [C.c_int]),
        "o_is_strict": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


# x86-64-v4 = AVX-512 F, BW, CD, DQ and VL (on top of x86-64-v3, which AVX-512 hosts have)
_V4_FLAGS = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")


def host_has_x86_64_v4() -> bool:
    """True when /proc/cpuinfo lists every AVX-512 subset of x86-64-v4."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    return all(x in flags for x in _V4_FLAGS)
    except OSError:
        pass
    return False


def lib_path(strict: bool = True, iteration: str = "2d") -> str:
    """The strict checker, or the throughput build for this host's CPU (x86-64-v4 when the
    host has it and the library was built, else x86-64-v3) iterating as the reference's
    USE_2D_ITER build (``iteration="2d"``) or its USE_MD_ITER + BALANCE_MD_ITER build ("md")."""
    if strict:
        return os.path.join(_HERE, "lib", "liboracle.so")
    stem = {"2d": "liboracle_fast", "md": "liboracle_fast_md"}[iteration]
    v4 = os.path.join(_HERE, "lib", stem + "_v4.so")
    if host_has_x86_64_v4() and os.path.exists(v4):
        return v4
    return os.path.join(_HERE, "lib", stem + ".so")


def load(strict: bool = True, iteration: str = "2d"):
    """Load (once) the strict checker or a fast-math baseline build."""
    key = (bool(strict), "strict" if strict else iteration)
    if key not in _libs:
        path = lib_path(strict, iteration)
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library missing: {path} (run `make -C oracle`)")
        _libs[key] = _proto(C.CDLL(path))
    return _libs[key]


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def make_params(lib, test="Sod", **kw) -> OParams:
    """Params with the reference defaults (parameters.h:34-97), overridden by
    keyword arguments, then ``init()`` (parameters.cpp:21-34)."""
    p = OParams()
    lib.o_params_defaults(C.byref(p))
    p.test = TEST_NAMES[test] if isinstance(test, str) else int(test)
    for k, v in kw.items():
        if k in ("domain_size", "domain_origin"):
            getattr(p, k)[0], getattr(p, k)[1] = v
        elif k == "bc_factors":   # ((u, v) for Left, Right, Top, Bottom): overrides the test's
            for side, (uf, vf) in enumerate(v):
                p.bc_factors[side][0], p.bc_factors[side][1] = uf, vf
            p.bc_override = 1
        else:
            setattr(p, k, v)
    lib.o_params_init(C.byref(p))
    return p


def reference_params(lib, test, **kw) -> OParams:
    """tests/reference.cpp:33-64: 100x100, 5 ghosts, GAD + Minmod, Euler 2nd
    order, Sequential splitting, 1000 cycles, test-default CFL and end time."""
    base = dict(scheme=SCHEME_GAD, projection=PROJ_EULER_2ND, limiter=LIM_MINMOD,
                axis_splitting=SPLIT_SEQUENTIAL, nb_ghosts=5, nx=100, ny=100,
                max_cycles=1000, max_time=0.0, verbose=5)
    base.update(kw)
    return make_params(lib, test, **base)


class Data:
    """The 16 solver arrays (data.h:15-39), owned by the oracle library."""

    def __init__(self, lib, nb_cells: int):
        self.lib = lib
        self.n = int(nb_cells)
        self.c = OData()
        if lib.o_data_alloc(C.byref(self.c), self.n) != 0:
            raise MemoryError("oracle data allocation failed")

    def array(self, name: str) -> np.ndarray:
        return np.ctypeslib.as_array(getattr(self.c, name), shape=(self.n,))

    def __del__(self):
        try:
            self.lib.o_data_free(C.byref(self.c))
        except Exception:
            pass


def run(lib, p: OParams, debug_indexes: bool = False, states: "OTestState | None" = None,
        init_fields: "dict | None" = None):
    """init_test + time_loop (armon_2D.cpp:389-407). Returns (status, data, cycles, final_dt, loop_time).
    ``states`` replaces the test's reference high/low states (a batched-ensemble domain);
    ``init_fields`` (field name -> array of nb_cells values) overwrites fields after init_test,
    for initial conditions outside the reference's tests."""
    d = Data(lib, p.nb_cells)
    if states is not None:
        lib.o_init_test_params_states(C.byref(p), C.byref(d.c), C.byref(states))
    else:
        lib.o_init_test_params(C.byref(p), C.byref(d.c), bool(debug_indexes))
    for name, values in (init_fields or {}).items():
        d.array(name)[:] = np.asarray(values, dtype=np.float64).reshape(-1)
    loop_time, final_dt, cycles = C.c_double(), C.c_double(), C.c_int()
    st = lib.o_time_loop(C.byref(p), C.byref(d.c), C.byref(loop_time), C.byref(final_dt), C.byref(cycles))
    return st, d, cycles.value, final_dt.value, loop_time.value
