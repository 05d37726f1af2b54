"""Python binding of the Armon MI355X library (``lib/libarmon.so``).

The library is C++/HIP; this module only loads it with ctypes and mirrors its C-ABI
(``include/armon/kernels.h`` and ``include/armon/armon.h``) so that tests and the
benchmark can drive it. It never computes anything itself: if the shared library is
missing, :func:`load` raises.

Names follow the reference: ``Range``, ``InnerRange1D/2D``, ``DomainRange``, the 14
kernel entry points (src/kernels/kernels.h:10-86) and the Params / time loop of
src/parameters.h and src/armon_2D.cpp.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libarmon.so")
LIB_F32_PATH = os.path.join(_HERE, "lib", "libarmon_f32.so")
CLI_PATH = os.path.join(_HERE, "bin", "armon_hip_2D")

# ---- enumerations (same integer values as the reference) -------------------------------
LIMITERS = {"None": 0, "Minmod": 1, "Superbee": 2}
TESTS = {"Sod": 0, "Sod_y": 1, "Sod_circ": 2, "Bizarrium": 3, "Sedov": 4}
SCHEME_GODUNOV, SCHEME_GAD = 0, 1
PROJ_NONE, PROJ_EULER, PROJ_EULER_2ND = 0, 1, 2
SPLIT = {"Sequential": 0, "SequentialSym": 1, "Strang": 2, "X_only": 3, "Y_only": 4}
AXIS_X, AXIS_Y = 0, 1
SIDE_LEFT, SIDE_RIGHT, SIDE_TOP, SIDE_BOTTOM = 0, 1, 2, 3
ENGINE_STAGED, ENGINE_FUSED = 0, 1
PREC_FP64, PREC_FP32_MIXED = 0, 1

FIELDS = ["x", "y", "rho", "umat", "vmat", "Emat", "pmat", "cmat", "gmat", "ustar", "pstar",
          "work_array_1", "work_array_2", "work_array_3", "work_array_4", "domain_mask"]
FIELD = {name: i for i, name in enumerate(FIELDS)}


# ---- C structures ----------------------------------------------------------------------
class Range(C.Structure):
    _fields_ = [("start", C.c_int64), ("end", C.c_int64)]


class InnerRange1D(C.Structure):
    _fields_ = [("start", C.c_int64), ("step", C.c_int64)]


class InnerRange2D(C.Structure):
    _fields_ = [("main_range_start", C.c_int64), ("main_range_step", C.c_int64),
                ("row_range_start", C.c_int64), ("row_range_length", C.c_uint64)]


class View(C.Structure):
    """Device array handle: {data, extent, label} (replaces Kokkos::View<flt_t*>)."""
    _fields_ = [("data", C.c_void_p), ("extent", C.c_int64), ("label", C.c_char_p)]


class DomainRange(C.Structure):
    _fields_ = [("col_start", C.c_int64), ("col_step", C.c_int64), ("col_end", C.c_int64),
                ("row_start", C.c_int64), ("row_step", C.c_int64), ("row_end", C.c_int64)]

    def as_tuple(self):
        return (self.col_start, self.col_step, self.col_end, self.row_start, self.row_step, self.row_end)


class ArmonParams(C.Structure):
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
        ("max_cycles", C.c_int), ("max_time", C.c_double),
        ("verbose", C.c_int),
        ("write_output", C.c_int), ("write_ghosts", C.c_int), ("write_throughput", C.c_int),
        ("output_precision", C.c_int), ("output_file", C.c_char_p),
        ("engine", C.c_int), ("precision", C.c_int), ("graph_chunk", C.c_int),
        ("compare", C.c_int), ("comparison_tolerance", C.c_double),
        ("y_run_rows", C.c_int),
        ("bc_override", C.c_int), ("bc_factors", (C.c_double * 2) * 4),
    ]


class ArmonResult(C.Structure):
    _fields_ = [("status", C.c_int), ("cycles", C.c_int), ("time", C.c_double),
                ("current_dt", C.c_double), ("final_dt", C.c_double), ("loop_time", C.c_double),
                ("grind_time", C.c_double), ("initial_mass", C.c_double), ("initial_energy", C.c_double),
                ("final_mass", C.c_double), ("final_energy", C.c_double)]


class ArmonTestState(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("high_rho", "low_rho", "high_E", "low_E",
                                          "high_u", "low_u", "high_v", "low_v")]


class ArmonDomainSpec(C.Structure):
    _fields_ = [("test", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("max_time", C.c_double),
                ("state", ArmonTestState)]


class ArmonBatchTuning(C.Structure):
    _fields_ = [("y_run_rows", C.c_int), ("grid_x", C.c_int), ("grid_y", C.c_int), ("items_x", C.c_int),
                ("items_y", C.c_int), ("ms_x", C.c_double), ("ms_y", C.c_double), ("tile_x", C.c_int),
                ("x_cells_per_thread", C.c_int)]


class ArmonKernelTimes(C.Structure):
    _fields_ = [("launches", C.c_int * 4), ("total_ms", C.c_double * 4)]


HANDLER_T = C.CFUNCTYPE(None, C.c_char_p, C.c_char_p)

# Every symbol the headers declare: name -> (restype, argtypes)
_R, _I1, _I2, _V = C.POINTER(Range), C.POINTER(InnerRange1D), C.POINTER(InnerRange2D), C.POINTER(View)
_i64, _d = C.c_int64, C.c_double
_PP = C.POINTER(ArmonParams)
KERNEL_SYMBOLS = {
    "acoustic": (None, [_R, _I2, _i64, _V, _V, _V, _V, _V, _V]),
    "acoustic_GAD": (None, [_R, _I2, _i64, _d, _d, _V, _V, _V, _V, _V, _V, C.c_int]),
    "perfect_gas_EOS": (None, [_R, _I2, _d, _V, _V, _V, _V, _V, _V, _V]),
    "bizarrium_EOS": (None, [_R, _I2, _V, _V, _V, _V, _V, _V, _V]),
    "cell_update": (None, [_R, _I2, _i64, _d, _d, _V, _V, _V, _V, _V]),
    "euler_projection": (None, [_R, _I2, _i64, _d, _d, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "advection_first_order": (None, [_R, _I2, _i64, _d, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "advection_second_order": (None, [_R, _I2, _i64, _d, _d, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "boundary_conditions": (None, [_R, _I1, _i64, _i64, _d, _d, _V, _V, _V, _V, _V, _V, _V]),
    "read_border_array": (None, [_R, _I2, _i64, _i64, _V, _V, _V, _V, _V, _V, _V, _V]),
    "write_border_array": (None, [_R, _I2, _i64, _i64, _V, _V, _V, _V, _V, _V, _V, _V]),
    "init_test": (None, [_R, _I2, _i64, _i64, _i64, _i64, _d, _d, _d, _d, _i64, _i64, _i64, _i64,
                         _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, C.c_int, C.c_bool, _d]),
    "dt_CFL": (_d, [_R, _I1, _d, _d, _V, _V, _V, _V]),
    "conservation_vars": (None, [_R, _I1, _d, _V, _V, _V, C.POINTER(_d), C.POINTER(_d)]),
    "idx_size": (C.c_int, []),
    "is_uidx_signed": (C.c_bool, []),
    "flt_size": (C.c_int, []),
}
EXTENSION_SYMBOLS = {
    "armon_set_stream": (C.c_int, [C.c_void_p]),
    "armon_get_stream": (C.c_void_p, []),
    "armon_device_count": (C.c_int, []),
    "armon_malloc": (C.c_void_p, [C.c_size_t]),
    "armon_free": (C.c_int, [C.c_void_p]),
    "armon_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "armon_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "armon_memcpy_dtod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "armon_synchronize": (C.c_int, []),
    "armon_last_error": (C.c_char_p, []),
    "armon_clear_error": (None, []),
    "armon_version": (C.c_char_p, []),
    "armon_copy_bandwidth": (C.c_double, [C.c_size_t, C.c_int]),
    "armon_reference_test_state": (None, [C.c_int, _d, C.POINTER(ArmonTestState)]),
    "init_test_states": (None, [_R, _I2, _i64, _i64, _i64, _i64, _d, _d, _d, _d, _i64, _i64, _i64, _i64,
                                _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V, C.c_int, _d,
                                C.POINTER(ArmonTestState)]),
    "dt_CFL_device": (None, [_R, _I1, _d, _d, _V, _V, _V, _V, C.c_void_p]),
    "conservation_vars_noalloc": (None, [_R, _I1, _d, _V, _V, _V, C.POINTER(_d), C.POINTER(_d)]),
    "armon_params_default": (None, [_PP]),
    "armon_params_init": (C.c_int, [_PP]),
    "armon_params_check": (C.c_int, [_PP]),
    "armon_min_ghosts": (C.c_int, [_PP]),
    "armon_update_axis": (None, [_PP, C.c_int]),
    "armon_split_axes": (C.c_int, [_PP, C.POINTER(C.c_int), C.POINTER(_d)]),
    "armon_test_boundary_condition": (None, [C.c_int, C.c_int, C.POINTER(_d), C.POINTER(_d)]),
    "armon_test_option": (_d, [_PP]),
    "armon_domain_col_length": (C.c_uint64, [C.POINTER(DomainRange)]),
    "armon_domain_row_length": (C.c_uint64, [C.POINTER(DomainRange)]),
    "armon_domain_length": (C.c_uint64, [C.POINTER(DomainRange)]),
    "armon_domain_begin": (C.c_int64, [C.POINTER(DomainRange)]),
    "armon_domain_end": (C.c_int64, [C.POINTER(DomainRange)]),
    "armon_domain_iter1D": (None, [C.POINTER(DomainRange), _R, _I1]),
    "armon_domain_direct_iter1D": (None, [C.POINTER(DomainRange), _R, _I1]),
    "armon_domain_iter2D": (None, [C.POINTER(DomainRange), _R, _I2]),
    "armon_index_1D": (C.c_int64, [_PP, C.c_int, C.c_int]),
    "armon_real_domain": (DomainRange, [_PP]),
    "armon_domain_fluxes": (DomainRange, [_PP]),
    "armon_domain_cell_update": (DomainRange, [_PP]),
    "armon_domain_advection": (DomainRange, [_PP]),
    "armon_complete_domain": (DomainRange, [_PP]),
    "armon_boundary_conditions_domain": (DomainRange, [_PP, C.c_int, C.POINTER(C.c_int)]),
    "armon_solver_create": (C.c_void_p, [_PP]),
    "armon_solver_destroy": (None, [C.c_void_p]),
    "armon_solver_get_params": (C.c_int, [C.c_void_p, _PP]),
    "armon_solver_init_test": (C.c_int, [C.c_void_p, C.c_bool]),
    "armon_solver_time_loop": (C.c_int, [C.c_void_p, C.POINTER(ArmonResult)]),
    "armon_solver_begin": (C.c_int, [C.c_void_p]),
    "armon_solver_advance": (C.c_int, [C.c_void_p, C.c_int]),
    "armon_solver_finish": (C.c_int, [C.c_void_p, C.POINTER(ArmonResult)]),
    "armon_solver_kernel_times": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(ArmonKernelTimes)]),
    "armon_solver_stage_count": (C.c_int, [C.c_void_p]),
    "armon_solver_stage_time": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(_d),
                                          C.POINTER(C.c_int64)]),
    "armon_solver_read_field": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(_d), C.c_int64]),
    "armon_solver_write_field": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(_d), C.c_int64]),
    "armon_solver_conservation": (C.c_int, [C.c_void_p, C.POINTER(_d), C.POINTER(_d)]),
    "armon_solver_stream": (C.c_void_p, [C.c_void_p]),
    "armon_run": (C.c_int, [_PP, C.POINTER(ArmonResult)]),
    "armon_batch_random_specs": (C.c_int, [C.c_uint64, C.c_int, C.c_int64, C.c_int64, _d,
                                           C.POINTER(ArmonDomainSpec)]),
    "armon_batch_create": (C.c_void_p, [_PP, C.c_int, C.POINTER(ArmonDomainSpec)]),
    "armon_batch_destroy": (None, [C.c_void_p]),
    "armon_batch_count": (C.c_int, [C.c_void_p]),
    "armon_batch_cells": (C.c_int64, [C.c_void_p]),
    "armon_batch_domain_params": (C.c_int, [C.c_void_p, C.c_int, _PP]),
    "armon_batch_stream": (C.c_void_p, [C.c_void_p]),
    "armon_batch_init": (C.c_int, [C.c_void_p]),
    "armon_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(ArmonBatchTuning)]),
    "armon_batch_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(ArmonBatchTuning)]),
    "armon_batch_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(ArmonBatchTuning)]),
    "armon_batch_begin": (C.c_int, [C.c_void_p]),
    "armon_batch_advance": (C.c_int, [C.c_void_p, C.c_int]),
    "armon_batch_run": (C.c_int, [C.c_void_p]),
    "armon_batch_active": (C.c_int, [C.c_void_p]),
    "armon_batch_finish": (C.c_int, [C.c_void_p]),
    "armon_batch_domain_result": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(ArmonResult)]),
    "armon_batch_read_field": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(_d), C.c_int64]),
    "armon_batch_kernel_times": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(ArmonKernelTimes)]),
}

_lib = None
_errors: list = []
_handler_ref = None


def load(path: Optional[str] = None):
    """Loads ``lib/libarmon.so`` once and installs a raise_exception_handler that records
    errors (``errors()``) instead of exiting the process. Raises if the library is missing."""
    global _lib, _handler_ref
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FileNotFoundError(f"Armon HIP library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for table in (KERNEL_SYMBOLS, EXTENSION_SYMBOLS):
        for name, (res, args) in table.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args

    def _handler(kernel, msg):
        _errors.append((kernel.decode(), msg.decode()))

    _handler_ref = HANDLER_T(_handler)
    C.c_void_p.in_dll(lib, "raise_exception_handler").value = C.cast(_handler_ref, C.c_void_p).value
    _lib = lib
    return lib


# The float build (flt_t = float, the reference's USE_SINGLE_PRECISION): the kernel table with
# every flt_t argument and result as float, and the memory plumbing (no flt_t).
_PD, _PF = C.POINTER(C.c_double), C.POINTER(C.c_float)
KERNEL_SYMBOLS_F32 = {
    name: (C.c_float if res is _d else res, [C.c_float if a is _d else _PF if a is _PD else a for a in args])
    for name, (res, args) in KERNEL_SYMBOLS.items()
}
PLUMBING_SYMBOLS = {k: EXTENSION_SYMBOLS[k] for k in (
    "armon_set_stream", "armon_get_stream", "armon_device_count", "armon_malloc", "armon_free",
    "armon_memcpy_htod", "armon_memcpy_dtoh", "armon_memcpy_dtod", "armon_synchronize", "armon_last_error",
    "armon_clear_error", "armon_version")}

_lib_f32 = None
_handler_ref_f32 = None


def load_f32(path: Optional[str] = None):
    """Loads ``lib/libarmon_f32.so``, the kernel C-ABI built with flt_t = float (common.h:8-12),
    with the same error recording as :func:`load`. Raises if the library is missing."""
    global _lib_f32, _handler_ref_f32
    if _lib_f32 is not None:
        return _lib_f32
    path = path or LIB_F32_PATH
    if not os.path.exists(path):
        raise FileNotFoundError(f"Armon HIP float library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for table in (KERNEL_SYMBOLS_F32, PLUMBING_SYMBOLS):
        for name, (res, args) in table.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args

    def _handler(kernel, msg):
        _errors.append((kernel.decode(), msg.decode()))

    _handler_ref_f32 = HANDLER_T(_handler)
    C.c_void_p.in_dll(lib, "raise_exception_handler").value = C.cast(_handler_ref_f32, C.c_void_p).value
    _lib_f32 = lib
    return lib


def errors(clear: bool = True) -> list:
    """Errors reported through raise_exception_handler since the last call."""
    out = list(_errors)
    if clear:
        _errors.clear()
    return out


class ArmonError(RuntimeError):
    pass


def _check(status, what, lib=None):
    if status:
        lib = lib or _lib
        msg = lib.armon_last_error().decode() if lib else ""
        raise ArmonError(f"{what} failed ({status}): {msg} {errors()}")


# ---- device memory ---------------------------------------------------------------------
class DeviceArray:
    """A device buffer of flt_t owned by Python, exposed to the kernels as a View: float64 for
    ``lib/libarmon.so``, float32 with ``single=True`` for ``lib/libarmon_f32.so`` (allocated,
    copied and synchronised through that library, on its stream)."""

    def __init__(self, n: int, label: Optional[str] = None, host: Optional[np.ndarray] = None,
                 single: bool = False):
        self._load = load_f32 if single else load
        self.dtype = np.float32 if single else np.float64
        lib = self._load()
        self.n = int(n)
        self.ptr = lib.armon_malloc(max(self.n, 1) * np.dtype(self.dtype).itemsize)
        if not self.ptr:
            raise MemoryError(lib.armon_last_error().decode())
        self._label = label.encode() if label else None
        if host is not None:
            self.upload(host)

    @property
    def view(self) -> View:
        return View(self.ptr, self.n, self._label)

    def upload(self, host: np.ndarray):
        a = np.ascontiguousarray(host, dtype=self.dtype)
        assert a.size == self.n
        lib = self._load()
        _check(lib.armon_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes), "armon_memcpy_htod", lib)

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        lib = self._load()
        _check(lib.armon_synchronize(), "armon_synchronize", lib)
        _check(lib.armon_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes), "armon_memcpy_dtoh", lib)
        return out

    def free(self):
        if self.ptr:
            self._load().armon_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---- parameters and domains ------------------------------------------------------------
def make_params(test="Sod", **kw) -> ArmonParams:
    """Params with the reference defaults (parameters.h:34-97), keyword overrides, then
    ``init()`` (parameters.cpp:21-34)."""
    lib = load()
    p = ArmonParams()
    lib.armon_params_default(C.byref(p))
    p.test = TESTS[test] if isinstance(test, str) else int(test)
    for k, v in kw.items():
        if k in ("domain_size", "domain_origin"):
            getattr(p, k)[0], getattr(p, k)[1] = v
        elif k == "bc_factors":   # ((u, v) for Left, Right, Top, Bottom): overrides the test's
            for side, (uf, vf) in enumerate(v):
                p.bc_factors[side][0], p.bc_factors[side][1] = uf, vf
            p.bc_override = 1
        elif k == "output_file":
            p.output_file = v.encode() if isinstance(v, str) else v
        else:
            setattr(p, k, v)
    _check(lib.armon_params_init(C.byref(p)), "armon_params_init")
    return p


def reference_params(test, **kw) -> ArmonParams:
    """tests/reference.cpp:33-64: 100x100, 5 ghosts, GAD + Minmod, Euler 2nd order,
    Sequential splitting, 1000 cycles, the test's default CFL and end time."""
    base = dict(scheme=SCHEME_GAD, projection=PROJ_EULER_2ND, limiter=1, axis_splitting=0,
                nb_ghosts=5, nx=100, ny=100, max_cycles=1000, max_time=0.0, verbose=5)
    base.update(kw)
    return make_params(test, **base)


# ---- solver ----------------------------------------------------------------------------
class Solver:
    """The host driver (src/armon_2D.cpp) with its device data."""

    def __init__(self, params: ArmonParams):
        lib = load()
        self.p = params
        self.h = lib.armon_solver_create(C.byref(params))
        if not self.h:
            raise ArmonError(f"armon_solver_create failed: {lib.armon_last_error().decode()} {errors()}")
        self.n = params.nb_cells

    def init_test(self, debug_indexes: bool = False):
        _check(load().armon_solver_init_test(self.h, bool(debug_indexes)), "init_test")

    def time_loop(self) -> ArmonResult:
        r = ArmonResult()
        st = load().armon_solver_time_loop(self.h, C.byref(r))
        r.status = st
        return r

    def begin(self):
        _check(load().armon_solver_begin(self.h), "begin")

    def advance(self, n: int):
        _check(load().armon_solver_advance(self.h, int(n)), "advance")

    def finish(self) -> ArmonResult:
        r = ArmonResult()
        r.status = load().armon_solver_finish(self.h, C.byref(r))
        return r

    def kernel_times(self, n: int) -> ArmonKernelTimes:
        t = ArmonKernelTimes()
        _check(load().armon_solver_kernel_times(self.h, int(n), C.byref(t)), "kernel_times")
        return t

    def stage_times(self) -> dict:
        """{label: (ms, calls)} of the per-stage device times (recorded when verbose < 3)."""
        lib = load()
        out = {}
        for k in range(lib.armon_solver_stage_count(self.h)):
            label, ms, calls = C.c_char_p(), C.c_double(), C.c_int64()
            _check(lib.armon_solver_stage_time(self.h, k, C.byref(label), C.byref(ms), C.byref(calls)),
                   "stage_time")
            out[label.value.decode()] = (ms.value, calls.value)
        return out

    def read(self, field) -> np.ndarray:
        f = FIELD[field] if isinstance(field, str) else int(field)
        out = np.empty(self.n, dtype=np.float64)
        _check(load().armon_solver_read_field(self.h, f, out.ctypes.data_as(C.POINTER(C.c_double)), self.n),
               "read_field")
        return out

    def write(self, field, values: np.ndarray):
        f = FIELD[field] if isinstance(field, str) else int(field)
        a = np.ascontiguousarray(values, dtype=np.float64)
        _check(load().armon_solver_write_field(self.h, f, a.ctypes.data_as(C.POINTER(C.c_double)), self.n),
               "write_field")

    def params(self) -> ArmonParams:
        p = ArmonParams()
        load().armon_solver_get_params(self.h, C.byref(p))
        return p

    def close(self):
        if self.h:
            load().armon_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(params: ArmonParams, init_fields: Optional[dict] = None):
    """init_test + time_loop on a new solver. Returns (result, solver)."""
    s = Solver(params)
    s.init_test()
    for k, v in (init_fields or {}).items():
        s.write(k, v)
    r = s.time_loop()
    return r, s


# ---- batched ensemble ------------------------------------------------------------------
def reference_state(test, test_option: float = 0.0) -> ArmonTestState:
    """The reference high/low states of a test (init_test.cpp:12-48)."""
    st = ArmonTestState()
    load().armon_reference_test_state(TESTS[test] if isinstance(test, str) else int(test), float(test_option),
                                      C.byref(st))
    return st


def random_specs(seed: int, n: int, min_area: int = 64 * 64, max_area: int = 1024 * 1024,
                 max_aspect: float = 4.0):
    """``n`` domain specs drawn from splitmix64(seed) (armon_batch_random_specs)."""
    arr = (ArmonDomainSpec * n)()
    _check(load().armon_batch_random_specs(C.c_uint64(seed), int(n), int(min_area), int(max_area),
                                           float(max_aspect), arr), "armon_batch_random_specs")
    return arr


def batch_params(**kw) -> ArmonParams:
    """Common parameters of a batch: the reference defaults plus overrides (no test-specific
    init; cfl 0 means each domain's test default)."""
    p = ArmonParams()
    load().armon_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Batch:
    """Many independent domains advanced together (armon_batch_*)."""

    def __init__(self, common: ArmonParams, specs):
        lib = load()
        n = len(specs)
        if not isinstance(specs, C.Array):
            arr = (ArmonDomainSpec * n)()
            for k, sp in enumerate(specs):
                arr[k] = sp
            specs = arr
        self.specs = specs
        self.h = lib.armon_batch_create(C.byref(common), n, specs)
        if not self.h:
            raise ArmonError(f"armon_batch_create failed: {lib.armon_last_error().decode()} {errors()}")
        self.n = n

    def init(self):
        _check(load().armon_batch_init(self.h), "armon_batch_init")

    def autotune(self) -> ArmonBatchTuning:
        t = ArmonBatchTuning()
        _check(load().armon_batch_autotune(self.h, C.byref(t)), "armon_batch_autotune")
        return t

    def tuning(self) -> ArmonBatchTuning:
        t = ArmonBatchTuning()
        load().armon_batch_get_tuning(self.h, C.byref(t))
        return t

    def set_tuning(self, y_run_rows: int, grid_x: int = 0, grid_y: int = 0, tile_x: int = 0,
                   x_cells_per_thread: int = 0):
        t = ArmonBatchTuning(y_run_rows=y_run_rows, grid_x=grid_x, grid_y=grid_y, tile_x=tile_x,
                             x_cells_per_thread=x_cells_per_thread)
        _check(load().armon_batch_set_tuning(self.h, C.byref(t)), "armon_batch_set_tuning")

    def begin(self):
        _check(load().armon_batch_begin(self.h), "armon_batch_begin")

    def advance(self, n: int):
        _check(load().armon_batch_advance(self.h, int(n)), "armon_batch_advance")

    def run(self):
        _check(load().armon_batch_run(self.h), "armon_batch_run")

    def active(self) -> int:
        return load().armon_batch_active(self.h)

    def finish(self) -> int:
        return load().armon_batch_finish(self.h)

    def result(self, d: int) -> ArmonResult:
        r = ArmonResult()
        _check(load().armon_batch_domain_result(self.h, int(d), C.byref(r)), "armon_batch_domain_result")
        return r

    def params(self, d: int) -> ArmonParams:
        p = ArmonParams()
        _check(load().armon_batch_domain_params(self.h, int(d), C.byref(p)), "armon_batch_domain_params")
        return p

    def cells(self) -> int:
        return load().armon_batch_cells(self.h)

    def read(self, d: int, field) -> np.ndarray:
        f = FIELD[field] if isinstance(field, str) else int(field)
        n = self.params(d).nb_cells
        out = np.empty(n, dtype=np.float64)
        _check(load().armon_batch_read_field(self.h, int(d), f, out.ctypes.data_as(C.POINTER(C.c_double)), n),
               "armon_batch_read_field")
        return out

    def kernel_times(self, n: int) -> ArmonKernelTimes:
        t = ArmonKernelTimes()
        _check(load().armon_batch_kernel_times(self.h, int(n), C.byref(t)), "armon_batch_kernel_times")
        return t

    def stream(self):
        return load().armon_batch_stream(self.h)

    def close(self):
        if self.h:
            load().armon_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
