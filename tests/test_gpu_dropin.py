"""The C++ drop-in (include/lbm.h, fluid.h, geometry.h over libsclbm.so)
driven the way the reference's drivers drive LBM / Fluid / Geometry
(tests/drivers/sc_driver.cpp), compared bit for bit with the golden vectors
of the compiled reference and with the oracle: the reference loop (fused),
the same loop with host reads in the middle of a step (operator by
operator), host round trips of the populations, and the reference's
out_of_range error when the fluid-solid force was never computed."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from sc_lbm_amd import workloads as wl
from tests import golden_cases as gc
from tests.drivers import case_io

pytestmark = pytest.mark.gpu

FIELD = {"f": po.F, "feq": po.FEQ, "rho": po.RHO, "ux": po.UX, "uy": po.UY, "ueqx": po.UEQX,
         "ueqy": po.UEQY, "frx": po.FRX, "fry": po.FRY, "fsx": po.FSX, "fsy": po.FSY}


@pytest.fixture(scope="module", autouse=True)
def _driver():
    if not os.path.exists(case_io.BIN):
        pytest.fail("tests/drivers/bin/sc_driver is not built (run __graft_entry__.build())")


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("name", gc.names())
def test_driver_matches_reference_golden(name, mode, tmp_path):
    case = gc.load(name)
    if mode == 2 and case["steps"] < 8:
        pytest.skip("round trip needs more than 7 steps")
    out = case_io.run(case, case["steps"], mode, str(tmp_path))
    for key in case["out"]:
        got, want = out[key], case["out"][key]
        assert np.array_equal(got, want, equal_nan=True), "%s mode %d: %s[%d] differs" % (
            name, mode, key[0], key[1])


@pytest.mark.parametrize("mode", [0, 1])
def test_driver_porous_matches_oracle(mode, tmp_path):
    case = wl.porous(160, 96, porosity=0.75, seed=31, rmin=4, rmax=7)
    steps = 300
    out = case_io.run(case, steps, mode, str(tmp_path))
    s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"])
    for c in range(2):
        s.set_f(c, case["f"][c])
    s.solid_surface_force()
    s.run(steps, case["force"][0], case["force"][1])
    for (nm, c), got in out.items():
        want = s.get(c, FIELD[nm])
        assert np.array_equal(got, want), "%s[%d] differs" % (nm, c)


def test_driver_reports_out_of_range_without_solid_force(tmp_path):
    case = wl.droplet(32, 6, seed=2)
    rc, text = case_io.run(case, 1, 3, str(tmp_path))
    assert rc == 3 and "out_of_range" in text
