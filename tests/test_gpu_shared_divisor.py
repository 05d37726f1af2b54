"""The shortened FP64 sequences of the sweeps against the compiler's own, on the device, bit for
bit (self-test kernels in armon-kokkos_amd/csrc/selftest_division.hip):
- the shared-divisor quotients (armon::with_divisor, device_math.hpp: one refined reciprocal
  per divisor for Godunov's u* and p*, the projection's u, v, E and Bizarrium's f0..f3) against
  IEEE division: a divisor and three numerators per sample, over every exponent of both signs
  with subnormals, the bounds of the ranges where the shared sequence is taken, zeros of both
  signs, infinities and NaNs, and mostly the range of physical states;
- the square root (armon::sqrt_t<double>, which skips the library's scaling and special-value
  select for finite operands of at least 2^-767) against __builtin_sqrt: every exponent with
  subnormals and a few negatives, the neighbourhood of 2^-767, zeros, infinities, NaN, the
  largest finite values, and mostly the range of physical states."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "armon-kokkos_amd", "lib", "libarmon_selftest.so")


def _check(seed, n):
    lib = ctypes.CDLL(LIB)
    fn = lib.armon_selftest_shared_divisor
    fn.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64),
                   ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    bad, shared = ctypes.c_uint64(0), ctypes.c_uint64(0)
    first = (ctypes.c_double * 4)()
    rc = fn(seed, n, ctypes.byref(bad), ctypes.byref(shared), first)
    assert rc == 0, f"HIP error {rc}"
    return bad.value, shared.value, list(first)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 20240917])
def test_shared_divisor_quotients_are_the_ieee_quotients(seed):
    n = 1 << 32
    print(f"shared divisor self-test: seed {seed}, {n} samples", flush=True)
    bad, shared, first = _check(seed, n)
    print(f"mismatches {bad}, shared-sequence samples {shared}", flush=True)
    assert bad == 0, f"{bad} quotients differ from a / b; one sample (b, a0, a1, a2): {first}"
    # both sequences ran: most samples lie in the ranges (physical states), many outside them
    assert 0.3 * n < shared < 0.9 * n


def _check_sqrt(seed, n):
    lib = ctypes.CDLL(LIB)
    fn = lib.armon_selftest_sqrt
    fn.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64),
                   ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    bad, fast = ctypes.c_uint64(0), ctypes.c_uint64(0)
    first = (ctypes.c_double * 1)()
    rc = fn(seed, n, ctypes.byref(bad), ctypes.byref(fast), first)
    assert rc == 0, f"HIP error {rc}"
    return bad.value, fast.value, first[0]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [3, 77031])
def test_short_square_root_is_the_library_square_root(seed):
    n = 1 << 32
    print(f"square root self-test: seed {seed}, {n} samples", flush=True)
    bad, fast, first = _check_sqrt(seed, n)
    print(f"mismatches {bad}, short-sequence samples {fast}", flush=True)
    assert bad == 0, f"{bad} roots differ from __builtin_sqrt; one operand: {first!r}"
    # both sequences ran: most operands are finite and at least 2^-767, many are not
    assert 0.5 * n < fast < 0.95 * n
