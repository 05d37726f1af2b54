"""conservation_vars_noalloc (include/armon/armon.h): the allocation-free variant of
conservation_vars, whose partial sums live in the library's static device scratch. It must
give exactly conservation_vars' results (same kernels, same block count), call after call,
whatever the previous call left in the scratch."""
import ctypes as C
from math import fsum

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 20241203


def _case(armon, nx, ny, seed):
    p = armon.make_params("Sod", nx=nx, ny=ny, nb_ghosts=4, engine=armon.ENGINE_STAGED)
    rng = np.random.default_rng(seed)
    n = p.nb_cells
    host = {"rho": rng.uniform(0.1, 2.0, n), "Emat": rng.uniform(0.5, 5.0, n),
            "domain_mask": (rng.uniform(0, 1, n) < 0.9).astype(np.float64)}
    dev = {k: armon.DeviceArray(n, f"Armon_{k}", v) for k, v in host.items()}
    lib = armon.load()
    d = lib.armon_real_domain(C.byref(p))
    r, ir = armon.Range(), armon.InnerRange1D()
    lib.armon_domain_iter1D(C.byref(d), C.byref(r), C.byref(ir))
    return p, host, dev, r, ir


def _sums(armon, fn, p, dev, r, ir):
    m, e = C.c_double(-1.0), C.c_double(-1.0)
    getattr(armon.load(), fn)(C.byref(r), C.byref(ir), p.dx, *[C.byref(dev[k].view) for k in
                                                               ("rho", "Emat", "domain_mask")], C.byref(m), C.byref(e))
    assert not armon.errors(), fn
    return m.value, e.value


@pytest.mark.parametrize("order", ["noalloc_first", "reference_first"])
def test_noalloc_matches_conservation_vars(armon, order):
    """Two calls in a row, the second on a smaller domain (fewer reduction blocks than the first
    left in the scratch): each bit-identical to conservation_vars, and within the compensated
    sum's bound of the exact sums."""
    for k, (nx, ny) in enumerate([(2000, 1500), (300, 200)]):
        p, host, dev, r, ir = _case(armon, nx, ny, SEED + k)
        if order == "noalloc_first":
            got = _sums(armon, "conservation_vars_noalloc", p, dev, r, ir)
            want = _sums(armon, "conservation_vars", p, dev, r, ir)
        else:
            want = _sums(armon, "conservation_vars", p, dev, r, ir)
            got = _sums(armon, "conservation_vars_noalloc", p, dev, r, ir)
        print(f"call {k}: {nx}x{ny} mass {got[0]!r} energy {got[1]!r}", flush=True)
        assert np.array_equal(np.array(got).view(np.uint64), np.array(want).view(np.uint64)), (got, want)
        idx = ir.start + np.arange(r.start, r.end) * ir.step
        cm = host["rho"][idx] * host["domain_mask"][idx] * (p.dx * p.dx)
        ce = cm * host["Emat"][idx]
        assert abs(got[0] - fsum(cm)) <= 2e-16 * abs(fsum(cm))
        assert abs(got[1] - fsum(ce)) <= 2e-16 * abs(fsum(ce))


def test_noalloc_empty_range_and_errors(armon):
    """An empty range returns zeros (the Sum identity), like conservation_vars; a view too short
    for the range is refused through raise_exception_handler before any launch."""
    p, host, dev, r, ir = _case(armon, 64, 48, SEED + 7)
    r.end = r.start
    assert _sums(armon, "conservation_vars_noalloc", p, dev, r, ir) == (0.0, 0.0)
    p, host, dev, r, ir = _case(armon, 64, 48, SEED + 8)
    short = armon.DeviceArray(10, "Armon_rho", np.ones(10))
    m, e = C.c_double(-1.0), C.c_double(-1.0)
    armon.load().conservation_vars_noalloc(C.byref(r), C.byref(ir), p.dx, C.byref(short.view),
                                           C.byref(dev["Emat"].view), C.byref(dev["domain_mask"].view),
                                           C.byref(m), C.byref(e))
    errs = armon.errors()
    assert errs and errs[0][0] == "conservation_vars_noalloc" and "out of bounds" in errs[0][1]
    assert (m.value, e.value) == (-1.0, -1.0)
