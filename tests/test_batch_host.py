"""Host side of the batched ensemble: the seeded domain generator and the oracle's init with
caller-given states (no GPU needed)."""
import ctypes as C
import math

import numpy as np
import pytest

TEST_IDS = {0, 2, 4}   # Sod, Sod_circ, Sedov


def spec_tuple(s):
    st = s.state
    return (s.test, s.nx, s.ny, s.max_time, st.high_rho, st.low_rho, st.high_E, st.low_E)


def test_random_specs_deterministic(armon):
    a = [spec_tuple(s) for s in armon.random_specs(1234, 64)]
    b = [spec_tuple(s) for s in armon.random_specs(1234, 64)]
    c = [spec_tuple(s) for s in armon.random_specs(1235, 64)]
    assert a == b
    assert a != c


def test_random_specs_distribution(armon):
    """256 domains: log-uniform areas in [64^2, 1024^2], aspect in [1, 4] both ways, the three
    tests, and the reference states perturbed by factors in [0.8, 1.2]."""
    specs = armon.random_specs(2024, 256)
    areas = np.array([s.nx * s.ny for s in specs], dtype=np.float64)
    aspect = np.array([max(s.nx, s.ny) / min(s.nx, s.ny) for s in specs])
    tests = [s.test for s in specs]
    assert set(tests) == TEST_IDS
    assert areas.min() >= 0.95 * 64 * 64 and areas.max() <= 1.05 * 1024 * 1024
    assert aspect.max() <= 4.0 * 1.05
    assert any(s.nx > s.ny for s in specs) and any(s.ny > s.nx for s in specs)
    # log-uniform: the mean of log(area) sits near the middle of the log range
    mid = 0.5 * (math.log(64 * 64) + math.log(1024 * 1024))
    assert abs(np.log(areas).mean() - mid) < 0.35
    # about 48M real cells in expectation
    assert 2.5e7 < areas.sum() < 8e7
    for s in specs:
        q = armon.make_params(int(s.test), nx=s.nx, ny=s.ny)
        ref = armon.reference_state(int(s.test), armon.load().armon_test_option(C.byref(q)))
        st = s.state
        for hl in ("high", "low"):
            f_rho = getattr(st, f"{hl}_rho") / getattr(ref, f"{hl}_rho")
            # p = (gamma - 1) rho E at rest: the pressure factor is the product of both ratios
            f_p = f_rho * getattr(st, f"{hl}_E") / getattr(ref, f"{hl}_E")
            assert 0.8 - 1e-12 <= f_rho <= 1.2 + 1e-12
            assert 0.8 - 1e-12 <= f_p <= 1.2 + 1e-12
        assert st.high_u == st.low_u == st.high_v == st.low_v == 0.0


def test_random_specs_rejects_bad_arguments(armon):
    lib = armon.load()
    out = (armon.ArmonDomainSpec * 1)()
    assert lib.armon_batch_random_specs(1, 1, 100, 1000, 2.0, out) == 1        # min_area below 16^2
    assert lib.armon_batch_random_specs(1, 1, 4096, 1024, 2.0, out) == 1       # max < min
    assert lib.armon_batch_random_specs(1, 1, 4096, 8192, 0.5, out) == 1       # aspect below 1


@pytest.mark.parametrize("test", ["Sod", "Sod_circ", "Bizarrium", "Sedov"])
def test_oracle_init_with_reference_states_is_init_test(po, oracle, test):
    """o_init_test_states with the reference states reproduces o_init_test exactly."""
    p = po.make_params(oracle, test, nx=37, ny=29, nb_ghosts=4)
    a = po.Data(oracle, p.nb_cells)
    b = po.Data(oracle, p.nb_cells)
    oracle.o_init_test_params(C.byref(p), C.byref(a.c), False)
    st = po.OTestState()
    oracle.o_reference_test_state(p.test, oracle.o_test_option(C.byref(p)), C.byref(st))
    oracle.o_init_test_params_states(C.byref(p), C.byref(b.c), C.byref(st))
    for f in ("x", "y", "rho", "umat", "vmat", "Emat", "pmat", "cmat", "gmat", "domain_mask"):
        assert np.array_equal(a.array(f).view(np.uint64), b.array(f).view(np.uint64)), f


def test_oracle_init_states_regions(po, oracle):
    """Caller-given states land in the test's high and low regions (Sod: x_mid <= 0.5)."""
    p = po.make_params(oracle, "Sod", nx=40, ny=8, nb_ghosts=4)
    d = po.Data(oracle, p.nb_cells)
    st = po.OTestState(high_rho=3.0, low_rho=0.5, high_E=7.0, low_E=1.5, high_u=0.0, low_u=0.0,
                       high_v=0.0, low_v=0.0)
    oracle.o_init_test_params_states(C.byref(p), C.byref(d.c), C.byref(st))
    rl = p.row_length
    rho = d.array("rho").reshape(-1, rl)
    E = d.array("Emat").reshape(-1, rl)
    g = p.nb_ghosts
    row = rho[g + 3, g:g + p.nx]
    assert np.all(row[:20] == 3.0) and np.all(row[20:] == 0.5)
    assert np.all(E[g + 3, g:g + 20] == 7.0) and np.all(E[g + 3, g + 20:g + p.nx] == 1.5)


def test_batch_symbols_exported(armon):
    lib = armon.load()
    for name in armon.EXTENSION_SYMBOLS:
        if name.startswith("armon_batch"):
            assert getattr(lib, name) is not None


def test_batch_rejects_mixed_eos(armon):
    """The domains of a batch share one EOS; the check runs on the host before any allocation."""
    specs = (armon.ArmonDomainSpec * 2)()
    specs[0].test, specs[0].nx, specs[0].ny = 3, 32, 32
    specs[1].test, specs[1].nx, specs[1].ny = 0, 32, 32
    with pytest.raises(armon.ArmonError, match="EOS"):
        armon.Batch(armon.batch_params(scheme=1, projection=2, nb_ghosts=4), specs)


def test_batch_rejects_small_domain(armon):
    specs = (armon.ArmonDomainSpec * 1)()
    specs[0].test, specs[0].nx, specs[0].ny = 0, 2, 32
    with pytest.raises(armon.ArmonError):
        armon.Batch(armon.batch_params(scheme=1, projection=2, nb_ghosts=4), specs)
