"""scl_moments_noalloc, the allocation-free variant of scl_moments (called
through the C-ABI): the same FP64 mass and momentum, bit for bit, in every
storage form and phase, on two calls in a row (the second on a smaller
lattice, so the shared device array holds the first call's partial sums
beyond the rows the second one writes); lattices of more than 32768 rows are
refused."""
import numpy as np
import pytest

from sc_lbm_amd import native as nat
from sc_lbm_amd import workloads as wl
from tests import gpu_helpers as gh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    nat.init(0)
    yield
    nat.synchronize()


def _both(case, prec, steps):
    g, fl = gh.make(case, precision=prec)
    if steps:
        nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], steps)
    out = []
    for f in fl:
        a = f.moments_noalloc()     # before scl_moments, whose staging buffer it does not use
        b = f.moments()
        out.append((a, b))
    for f in fl:
        f.close()
    g.close()
    return out


@pytest.mark.parametrize("prec", [nat.FP64, nat.FP32_DEV, nat.FP16_DEV])
def test_noalloc_moments_equal_scl_moments(prec):
    big = wl.porous(700, 480, porosity=0.75, seed=17, rmin=5, rmax=12)
    small = wl.porous(260, 96, porosity=0.8, seed=18, rmin=4, rmax=8)
    for case, steps in ((big, 0), (big, 7), (small, 5), (small, 0)):
        for a, b in _both(case, prec, steps):
            assert np.all(np.isfinite(a))
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (a, b)


def test_noalloc_moments_refuse_more_rows_than_the_array_holds():
    mask = wl.open_mask(64, 32769)
    g = nat.Geometry(mask)
    f = nat.Field(g, 1.0)
    with pytest.raises(nat.SclError, match="at most 32768"):
        f.moments_noalloc()
    f.close()
    g.close()
