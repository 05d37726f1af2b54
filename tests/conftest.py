import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def armon():
    import __graft_entry__ as g

    mod = g.import_package()
    mod.load()
    mod.errors()
    return mod


@pytest.fixture(scope="session")
def oracle():
    import pyoracle as po

    return po.load(strict=True)


@pytest.fixture(scope="session")
def po():
    import pyoracle

    return pyoracle


@pytest.fixture(scope="session")
def kats():
    """Known answers of the reference's own tests (tests/golden/reference_kats.json)."""
    import json

    with open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")) as f:
        return json.load(f)


def reference_cycle_counts():
    """(test, cycles) pairs of the reference's NaN-check suite, for parametrize."""
    import json

    with open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")) as f:
        counts = json.load(f)["cycle_counts"]
    return [(k, v) for k, v in counts.items() if k != "source"]


# A configuration outside the reference's tests, through the explicit overrides: the Sedov
# blast at a corner of the domain, at the origin, with reflective walls and a background
# internal energy of 1e-5 (the quarter-symmetric Sedov problem). Domain [0, 1]^2; the quarter
# of the reference blast energy (0.851072, init_test.cpp:38-48) in the 2x2 corner cells.
SEDOV_CORNER = dict(domain_origin=(0.0, 0.0), domain_size=(1.0, 1.0),
                    bc_factors=((-1.0, 1.0), (-1.0, 1.0), (1.0, -1.0), (1.0, -1.0)))


def sedov_corner_fields(p):
    """Initial rho, umat, vmat, Emat over all nb_cells of params `p` (ghosts included)."""
    import numpy as np

    n = p.nb_cells
    rho, u, v, E = np.ones(n), np.zeros(n), np.zeros(n), np.full(n, 1e-5)
    dx, dy = p.domain_size[0] / p.nx, p.domain_size[1] / p.ny
    for j in (0, 1):
        for i in (0, 1):
            E[p.index_start + j * p.row_length + i] = 0.851072 / 4 / (4 * dx * dy)
    return {"rho": rho, "umat": u, "vmat": v, "Emat": E}


def sod_circ_radius_fields(p, radius):
    """Initial rho, umat, vmat, Emat of the circular Sod test (test_cases.h:73-76,
    init_test.cpp:12-22) with another radius about (0.5, 0.5), over all nb_cells of params `p`
    (ghost cells by the same test on their centres, as init_test does)."""
    import numpy as np

    dx, dy = p.domain_size[0] / p.nx, p.domain_size[1] / p.ny
    g = p.nb_ghosts
    i = np.arange(p.row_length) - g
    j = np.arange(p.col_length) - g
    x = p.domain_origin[0] + (i + 0.5) * dx
    y = p.domain_origin[1] + (j + 0.5) * dy
    inside = ((x[None, :] - 0.5) ** 2 + (y[:, None] - 0.5) ** 2 <= radius * radius).reshape(-1)
    n = p.nb_cells
    rho = np.where(inside, 1.0, 0.125)[:n]
    E = np.where(inside, 2.5, 2.0)[:n]
    return {"rho": rho, "umat": np.zeros(n), "vmat": np.zeros(n), "Emat": E}


def sod_circ_swirl_fields(p, radius, speed):
    """The circular Sod state of sod_circ_radius_fields set in a vortex: u = speed sin(pi x)
    cos(pi y), v = -speed cos(pi x) sin(pi y) over the unit square, which vanishes normal to its
    reflective walls; E adds the kinetic energy, so the pressure stays the test's. Every cell
    moves from the first cycle, in every direction."""
    import numpy as np

    f = sod_circ_radius_fields(p, radius)
    dx, dy = p.domain_size[0] / p.nx, p.domain_size[1] / p.ny
    g = p.nb_ghosts
    x = p.domain_origin[0] + (np.arange(p.row_length) - g + 0.5) * dx
    y = p.domain_origin[1] + (np.arange(p.col_length) - g + 0.5) * dy
    n = p.nb_cells
    u = (speed * np.sin(np.pi * x)[None, :] * np.cos(np.pi * y)[:, None]).reshape(-1)[:n]
    v = (-speed * np.cos(np.pi * x)[None, :] * np.sin(np.pi * y)[:, None]).reshape(-1)[:n]
    f["umat"], f["vmat"] = u, v
    f["Emat"] = f["Emat"] + 0.5 * (u * u + v * v)
    return f


def bizarrium_fields(p, left_moving):
    """Initial rho, umat, vmat, Emat of the Bizarrium tube over all nb_cells of params `p`: the
    reference's (init_test.cpp:25-35: x <= 0.5 at rho 1.42857142857e4 and internal energy
    4.48657821135e6 at rest; the rest at rho 1e4, internal energy 0, moving at u = 250, so
    E = 31250) when `left_moving` is false, else the same internal energies with the left
    state moving at u = 250 and the right one at rest (E = e + u^2 / 2)."""
    import numpy as np

    dx = p.domain_size[0] / p.nx
    g = p.nb_ghosts
    x = p.domain_origin[0] + (np.arange(p.row_length) - g + 0.5) * dx
    left = np.broadcast_to(x[None, :] <= 0.5, (p.col_length, p.row_length)).reshape(-1)[:p.nb_cells]
    n = p.nb_cells
    u_left, u_right = (250.0, 0.0) if left_moving else (0.0, 250.0)
    rho = np.where(left, 1.42857142857e4, 1e4)
    u = np.where(left, u_left, u_right)
    E = np.where(left, 4.48657821135e6 + 0.5 * u_left * u_left, 0.0 + 0.5 * u_right * u_right)
    return {"rho": rho, "umat": u, "vmat": np.zeros(n), "Emat": E}
