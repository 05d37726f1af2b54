"""CPU checks of the whole-run oracle digests (tests/golden/full_runs.json) and of the helper
that extracts the real cells they are taken over."""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_full_runs import CASES, FIELDS, digest, real_cells  # noqa: E402


def load_cases():
    with open(os.path.join(HERE, "golden", "full_runs.json")) as f:
        return json.load(f)["cases"]


def whole_runs():
    return {k: c for k, c in load_cases().items() if c.get("kind") != "batch"}


def test_fixture_holds_every_case():
    cases = load_cases()
    assert set(cases) == set(CASES)
    for name, c in cases.items():
        if c.get("kind") == "batch":
            want = CASES[name]
            assert (c["seed"], c["params"], len(c["domains"])) == (want["seed"], want["params"], want["domains"])
            for k, d in enumerate(c["domains"]):
                assert len(d["sha256"]) == 64 and int(d["sha256"], 16) >= 0, (name, k)
                assert float.fromhex(d["final_dt"]) > 0 and d["rho_sum"] > 0, (name, k)
            assert len({d["sha256"] for d in c["domains"]}) == len(c["domains"])
            continue
        assert c["test"] == CASES[name]["test"] and c["params"] == CASES[name]["params"], name
        assert c.get("init") == CASES[name].get("init"), name
        assert float.fromhex(c["final_dt"]) > 0
        assert set(c["fields"]) == set(FIELDS)
        for f, d in c["fields"].items():
            assert len(d["sha256"]) == 64 and int(d["sha256"], 16) >= 0, (name, f)
            assert np.isfinite(d["sum"]) and d["abs_sum"] >= abs(d["sum"]), (name, f)


def test_fixed_step_cases_ran_their_steps():
    cases = load_cases()
    for name, c in cases.items():
        mc = c["params"]["max_cycles"]
        if c.get("kind") == "batch":   # end times pushed out: every domain runs max_cycles
            assert all(d["cycles"] == mc for d in c["domains"]), name
        elif mc < 10**6:
            assert c["cycles"] == mc, name


def test_sedov_mass_leaves_only_through_the_boundary(oracle, po):
    """By t = 1 the Sedov blast reaches the free-flow boundaries and mass leaves through them;
    none is created. The final mass lies below the initial one, by a few percent."""
    c = load_cases()["sedov_1024_to_end_time"]
    p = po.make_params(oracle, c["test"], **dict(c["params"], max_cycles=0))
    d = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(d.c), False)
    m0 = float(np.sum(real_cells(d.array("rho"), p.nx, p.ny, p.row_length, p.index_start)))
    m1 = c["fields"]["rho"]["sum"]
    assert 0.9 * m0 < m1 < m0, (m0, m1)


def test_sod_circ_conserves_mass():
    """After 200 steps at 8192^2 the circular Sod waves are far from the boundary (and the
    vortex of the swirling variant is tangent to the reflective walls), so the mass is the
    initial one to rounding: cells whose centre lies within the radius (0.3, the reference's;
    0.2 for the benchmark list's variants) of (0.5, 0.5) hold rho = 1, the others
    0.125 (test_cases.h:73-76, init_test.cpp:12-22), counted here from the cell centres without
    the oracle."""
    for name, radius in (("sod_circ_8192_200_steps", 0.3), ("sod_circ_r02_8192_200_steps", 0.2),
                         ("sod_circ_r02_swirl_8192_200_steps", 0.2)):
        c = load_cases()[name]
        n = c["params"]["nx"]
        centre = (np.arange(n) + 0.5) / n
        inside = 0
        for j in range(0, n, 1024):
            y = centre[j:j + 1024, None]
            inside += int(np.count_nonzero((centre[None, :] - 0.5) ** 2 + (y - 0.5) ** 2 <= radius * radius))
        m0 = inside * 1.0 + (n * n - inside) * 0.125
        assert abs(c["fields"]["rho"]["sum"] - m0) <= 1e-12 * m0, name


def test_positive_fields():
    for name, c in whole_runs().items():
        for f in ("rho", "cmat"):
            d = c["fields"][f]
            assert d["sum"] == d["abs_sum"] > 0, (name, f)


def test_symmetric_blasts_have_no_net_velocity():
    """The Sedov and circular Sod states are mirror-symmetric in x and y, and the swirling one
    is symmetric under the half turn about (0.5, 0.5) (which negates u and v), so each velocity
    component sums to zero up to rounding (not exactly: the splitting evaluates the two sides
    of a mirror in different orders)."""
    cases = load_cases()
    for name in ("sedov_1024_to_end_time", "sod_circ_8192_200_steps", "sod_circ_r02_8192_200_steps",
                 "sod_circ_r02_swirl_8192_200_steps"):
        for f in ("umat", "vmat"):
            d = cases[name]["fields"][f]
            assert d["abs_sum"] > 0 and abs(d["sum"]) <= 1e-4 * d["abs_sum"], (name, f, d)


def test_bizarrium_tube_has_no_y_velocity():
    """The Bizarrium tube varies along x only, between reflective walls in y: v stays (signed) zero,
    whichever side moves."""
    for name in ("bizarrium_4096_superbee_500_steps", "bizarrium_left_moving_4096_superbee_500_steps"):
        d = load_cases()[name]["fields"]["vmat"]
        assert d["sum"] == 0 and d["abs_sum"] == 0, name


def test_real_cells_follow_the_domain_mask(oracle, po):
    """real_cells returns exactly the cells the oracle marks real, in row-major order."""
    p = po.make_params(oracle, "Sod", nx=13, ny=7, nb_ghosts=4, max_cycles=0)
    d = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(d.c), True)
    mask = d.array("domain_mask")
    assert np.all(real_cells(mask, p.nx, p.ny, p.row_length, p.index_start) == 1)
    assert int(np.sum(mask == 1)) == p.nx * p.ny
    x = np.arange(p.nb_cells, dtype=np.float64)
    assert np.array_equal(real_cells(x, p.nx, p.ny, p.row_length, p.index_start).ravel(), x[mask == 1])


def test_digest_is_over_the_exact_bits():
    a = np.array([[0.0, 1.0]])
    b = np.array([[-0.0, 1.0]])
    assert digest(a)["sha256"] != digest(b)["sha256"]
    assert digest(a)["sum"] == digest(b)["sum"]


def test_corner_sedov_keeps_its_mass():
    """The corner Sedov blast is closed by reflective walls on all four sides, through which no
    mass passes (the wall face velocity is exactly zero): at t = 1 the mass is still the initial
    nx * ny * 1 to rounding."""
    c = load_cases()["sedov_corner_1024_to_end_time"]
    m0 = float(c["params"]["nx"] * c["params"]["ny"])
    assert abs(c["fields"]["rho"]["sum"] - m0) <= 1e-12 * m0


def test_benchmark_vortex_state_is_the_fixture_state():
    """bench.py's developed flow starts from the initial state of the swirling circular Sod
    case (the same arrays, bit for bit), so the fixture's digests pin its FP64 side; the
    vortex is tangent to the walls and the pressure is the test's."""
    import types

    sys.path.insert(0, os.path.dirname(HERE))
    import bench
    from conftest import sod_circ_swirl_fields

    n, g = 96, 4
    p = types.SimpleNamespace(nx=n, ny=n, nb_ghosts=g, row_length=n + 2 * g, col_length=n + 2 * g,
                              nb_cells=(n + 2 * g) ** 2, domain_origin=(0.0, 0.0), domain_size=(1.0, 1.0))
    a, b = bench.swirl_fields(p, 0.2, 0.5), sod_circ_swirl_fields(p, 0.2, 0.5)
    for f in ("rho", "umat", "vmat", "Emat"):
        assert np.array_equal(a[f], b[f]), f
    rl = p.row_length
    u = a["umat"].reshape(-1, rl)[g:g + n, g:g + n]
    v = a["vmat"].reshape(-1, rl)[g:g + n, g:g + n]
    assert np.all(np.abs(u) + np.abs(v) > 0)
    # wall-normal velocities of the first and last real cells are O(dx), tangent to the walls
    assert np.max(np.abs(u[:, 0])) < 0.5 * np.pi / n and np.max(np.abs(v[0, :])) < 0.5 * np.pi / n
    rho = a["rho"].reshape(-1, rl)[g:g + n, g:g + n]
    e = a["Emat"].reshape(-1, rl)[g:g + n, g:g + n] - 0.5 * (u * u + v * v)
    assert np.allclose(0.4 * rho * e, np.where(rho == 1.0, 1.0, 0.1), rtol=1e-14)
