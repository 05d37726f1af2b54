"""The Julia calling convention (Armon.jl's ccall through Kokkos.jl, reference README.md:5-7)
exercised from C by tests/abi/ccall_check.c: dlopen/dlsym with no header, Range/InnerRange/view
structs by pointer, the error handler replaced through the raise_exception_handler data symbol
(utils.cpp:16-17), a returning and an unwinding (longjmp) handler; acoustic_GAD and dt_CFL
bit for bit against the oracle opened the same way."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "abi", "bin", "ccall_check")


def test_ccall_convention_from_c():
    lib = os.path.join(ROOT, "armon-kokkos_amd", "lib", "libarmon.so")
    olib = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
    print("case ccall_check", flush=True)
    out = subprocess.run([EXE, lib, olib], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PASSED: 0 failure(s)" in out.stdout
    assert out.stdout.count("\nok ") + out.stdout.startswith("ok ") >= 12
