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
