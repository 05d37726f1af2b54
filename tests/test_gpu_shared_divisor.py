"""The shared-divisor quotients of the FP64 sweeps (armon::with_divisor, device_math.hpp: one
refined reciprocal per divisor for Godunov's u* and p*, the projection's u, v, E and Bizarrium's
f0..f3) against the compiler's IEEE division on the device, bit for bit. The self-test kernel
(armon-kokkos_amd/csrc/selftest_division.hip) draws a divisor and three numerators per sample:
every exponent of both signs with subnormals, the bounds of the ranges where the shared sequence
is taken, zeros of both signs, infinities and NaNs, and mostly the range of physical states."""
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
