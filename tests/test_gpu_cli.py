"""The command-line driver on the GPU: a reference-style run with an output file (io.cpp
format) that matches the oracle's final state, and the batched-ensemble mode."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "armon-kokkos_amd", "bin", "armon_hip_2D")


def run(args, cwd):
    print("case cli", " ".join(args), flush=True)
    return subprocess.run([EXE] + args, capture_output=True, text=True, timeout=120, cwd=cwd)


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_cli_run_writes_the_oracle_state(po, oracle, tmp_path, engine):
    out = run(["-t", "Sod_circ", "-s", "GAD", "--projection", "euler_2nd", "--cells", "40,32", "--nghost", "4",
               "--cycle", "25", "--time", "1000", "--write-output", "1", "--precision", "17", "--output", str(tmp_path / "final"),
               "--engine", engine, "--verbose", "3"], tmp_path)
    assert out.returncode == 0, out.stderr
    assert "Cycles:     25" in out.stdout
    got = np.loadtxt(tmp_path / "final", delimiter=",")
    p = po.make_params(oracle, "Sod_circ", scheme=1, projection=2, nx=40, ny=32, nb_ghosts=4, max_cycles=25,
                       max_time=1000.0, verbose=5)
    st, d, cycles, _, _ = po.run(oracle, p)
    assert st == 0 and cycles == 25
    rl, g = p.row_length, p.nb_ghosts
    rows = []
    for j in range(p.ny):
        for i in range(p.nx):
            idx = p.index_start + j * rl + i
            rows.append([d.array(f)[idx] for f in ("x", "y", "rho", "umat", "vmat", "pmat")])
    ref = np.array(rows)
    # 17 significant digits in the file: the round trip is exact
    assert np.array_equal(got, ref)


def test_cli_write_ghosts(po, oracle, tmp_path):
    """--write-ghosts 1 (io.cpp:12-15): the ghost layers too, which only the staged engine keeps in
    memory, so the CLI runs it when no engine is named; every cell of the file, ghosts and corners
    included, equals the oracle's arrays."""
    args = ["-t", "Sod_circ", "-s", "GAD", "--projection", "euler_2nd", "--cells", "40,32", "--nghost", "4",
            "--cycle", "25", "--time", "1000", "--write-output", "1", "--write-ghosts", "1", "--precision", "17",
            "--output", str(tmp_path / "final"), "--verbose", "2"]
    out = run(args, tmp_path)
    assert out.returncode == 0, out.stderr
    assert "engine:     staged" in out.stdout
    got = np.loadtxt(tmp_path / "final", delimiter=",")
    p = po.make_params(oracle, "Sod_circ", scheme=1, projection=2, nx=40, ny=32, nb_ghosts=4, max_cycles=25,
                       max_time=1000.0, verbose=5)
    st, d, cycles, _, _ = po.run(oracle, p)
    assert st == 0 and cycles == 25
    g = p.nb_ghosts
    rows = []
    for j in range(-g, p.ny + g):
        for i in range(-g, p.nx + g):
            idx = p.index_start + j * p.row_length + i
            rows.append([d.array(f)[idx] for f in ("x", "y", "rho", "umat", "vmat", "pmat")])
    assert got.shape == (len(rows), 6)
    assert np.array_equal(got, np.array(rows))
    # the fused engine refuses to write ghost cells it does not keep
    out = run(args + ["--engine", "fused"], tmp_path)
    assert out.returncode != 0 and "staged engine" in (out.stdout + out.stderr)


def test_cli_batch_mode(tmp_path):
    out = run(["--batch", "12", "--seed", "7", "--cycle", "15", "--nghost", "4", "--projection", "euler_2nd",
               "--verbose", "1"], tmp_path)
    assert out.returncode == 0, out.stderr
    lines = [l for l in out.stdout.splitlines() if l.startswith("domain")]
    assert len(lines) == 12
    assert "Batch of 12 domains" in out.stdout and "Autotuned" in out.stdout
