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
