"""Per-step comparison mode (armon_2D.cpp:255-284, io.cpp:55-171) on the staged engine.

Reference step files are written from the oracle, stage by stage, in the reference's text
format (io.cpp:10-44) with 17 significant digits; the library's run with `compare` must find
no difference at tolerance 0, and must stop at the first step whose file is altered, writing
the state of that step to <file>_diff."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAIN = ("x", "y", "rho", "umat", "vmat", "pmat")
CFG = dict(scheme=1, projection=2, limiter=1, nb_ghosts=4, nx=20, ny=16, max_cycles=3, verbose=5)


def write_state(path, p, d, ghosts):
    g = p.nb_ghosts if ghosts else 0
    arrays = [d.array(f) for f in MAIN]
    with open(path, "w") as f:
        for j in range(-g, p.ny + g):
            for i in range(-g, p.nx + g):
                idx = p.index_start + j * p.row_length + i
                f.write(", ".join(f"{a[idx]:23.16e}" for a in arrays) + "\n")
            f.write("\n")


def write_reference_steps(po, oracle, prefix, test, ghosts):
    """The oracle's loop (armon_2D.cpp:304-361) with a file after every step."""
    p = po.make_params(oracle, test, **CFG)
    d = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(d.c), False)
    axis_name = lambda: "X" if p.current_axis == 0 else "Y"
    name = lambda label: f"{prefix}_{p.cycle:03d}_{label}_{axis_name()}"
    write_state(name("init_test"), p, d, ghosts)
    p.cycle, p.time, p.current_cycle_dt, p.next_cycle_dt = 0, 0.0, 0.0, 0.0
    oracle.o_update_axis(C.byref(p), 0)
    oracle.o_update_EOS(C.byref(p), C.byref(d.c))
    write_state(name("EOS_init"), p, d, ghosts)
    steps = 2
    while p.time < p.max_time and p.cycle < p.max_cycles:
        assert oracle.o_time_step(C.byref(p), C.byref(d.c)) == 0
        with open(name("time_step"), "w") as f:
            f.write(f"{p.current_cycle_dt:.17g}\n")
        axes, factors = (C.c_int * 3)(), (C.c_double * 3)()
        for k in range(oracle.o_split_axes(C.byref(p), axes, factors)):
            oracle.o_update_axis(C.byref(p), axes[k])
            p.cycle_dt = p.current_cycle_dt * factors[k]
            for label, fn in (("EOS", oracle.o_update_EOS), ("boundary_conditions", oracle.o_boundary_conditions_all),
                              ("numerical_fluxes", oracle.o_numerical_fluxes),
                              ("cell_update", oracle.o_cell_update_params),
                              ("projection_remap", oracle.o_projection_remap)):
                fn(C.byref(p), C.byref(d.c))
                write_state(name(label), p, d, ghosts)
                steps += 1
        p.time += p.current_cycle_dt
        p.cycle += 1
    return steps


def run_compare(armon, prefix, test, ghosts, tol=0.0):
    p = armon.make_params(test, engine=0, compare=1, comparison_tolerance=tol, output_file=str(prefix),
                          write_ghosts=int(ghosts), **CFG)
    r = armon.ArmonResult()
    status = armon.load().armon_run(C.byref(p), C.byref(r))
    return status, r


@pytest.mark.parametrize("ghosts", [False, True])
@pytest.mark.parametrize("test", ["Sod_circ", "Sedov"])
def test_compare_all_steps_match(armon, po, oracle, tmp_path, test, ghosts):
    prefix = str(tmp_path / "ref")
    n = write_reference_steps(po, oracle, prefix, test, ghosts)
    print(f"case compare {test} ghosts={ghosts}: {n} step files", flush=True)
    status, r = run_compare(armon, prefix, test, ghosts)
    assert status == 0, armon.errors()
    assert r.cycles == CFG["max_cycles"]
    assert not [f for f in os.listdir(tmp_path) if f.endswith("_diff")]


def test_compare_stops_at_altered_step(armon, po, oracle, tmp_path):
    prefix = str(tmp_path / "ref")
    write_reference_steps(po, oracle, prefix, "Sod_circ", False)
    target = f"{prefix}_001_cell_update_Y"
    lines = open(target).read().split("\n")
    cols = lines[5].split(",")
    cols[2] = f"{float(cols[2]) * (1 + 1e-9):23.16e}"    # rho of one cell
    lines[5] = ",".join(cols)
    open(target, "w").write("\n".join(lines))
    print("case compare altered cell_update", flush=True)
    status, r = run_compare(armon, prefix, "Sod_circ", False, tol=1e-13)
    assert status == 3
    assert os.path.exists(target + "_diff")
    assert not os.path.exists(f"{prefix}_002_time_step_Y_diff")
    # the diff file holds the library's state of that step: it equals the unaltered reference
    got = np.loadtxt(target + "_diff", delimiter=",")
    write_reference_steps(po, oracle, str(tmp_path / "clean"), "Sod_circ", False)
    ref = np.loadtxt(str(tmp_path / "clean") + "_001_cell_update_Y", delimiter=",")
    assert np.allclose(got, ref, rtol=1e-8, atol=0)


def test_compare_time_step_and_missing_file(armon, po, oracle, tmp_path):
    prefix = str(tmp_path / "ref")
    write_reference_steps(po, oracle, prefix, "Sedov", False)
    dt_file = f"{prefix}_001_time_step_Y"
    dt = float(open(dt_file).read())
    open(dt_file, "w").write(f"{dt * (1 + 1e-10):.17g}\n")
    print("case compare altered time step", flush=True)
    status, _ = run_compare(armon, prefix, "Sedov", False, tol=1e-13)
    assert status == 3 and os.path.exists(dt_file + "_diff")
    os.remove(f"{prefix}_000_EOS_init_X")
    print("case compare missing file", flush=True)
    status, r = run_compare(armon, prefix, "Sedov", False)
    assert status == 3 and r.cycles == 0
