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


def _dropin_refs_exe(name):
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "dropin", name)


@pytest.mark.parametrize("name", ["fluid_test_main", "two_fluids_test_main"])
def test_reference_fluid_unit_tests_pass_on_the_dropin(name, tmp_path):
    """The reference's own Fluid unit tests (tests/fluid/*.cpp of the
    reference), compiled unchanged against include/ and libsclbm.so by
    oracle/dropin_refs.py, run on the GPU: initialisers, macroscopic
    quantities, equilibrium distributions and text output, checked by the
    reference's own code."""
    import subprocess
    exe = _dropin_refs_exe(name)
    if not os.path.exists(exe):
        pytest.skip("reference test programs not built here (oracle/_ref/dropin)")
    (tmp_path / "test_data").mkdir()
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300,
                       cwd=str(tmp_path))
    assert p.returncode == 0, p.stdout[-3000:]
    assert p.stdout.count("test passed") == 3 and "failed" not in p.stdout, p.stdout[-3000:]


def _force_args(f):
    return [float(v) for v in (f if f is not None else np.zeros(9))]


@pytest.mark.parametrize("name", ["dry_rect_48x32_s60", "init_droplet_40x36_s1"])
def test_builders_initialisers_and_loop_match_reference(name, tmp_path):
    """Geometry builders, LBM initialiser and the step loop, all through the
    drop-in classes, against the reference's golden run of the same calls
    (tests/golden/make_golden.py cases 5 and 6)."""
    case = gc.load(name)
    if name.startswith("dry_rect"):
        head = ["G", 48, 32, "W", 1, "y", "F", 1.3, 0.8, "IR", 2.0, 2.0, 0.0, 0.0, 20.0, 14.0, 9.0, 6.0]
    else:
        head = ["G", 40, 36, "C", 9, 12, 10, "F", 1.0, 1.0, "ID", 1.0, 0.9, 0.05, 0.04, 20.5, 18.0, 9.5]
    loop = ["S", float(case["g_solid"][0]), float(case["g_solid"][1]), float(case["g_fluid"][0]),
            float(case["g_fluid"][1]), case["steps"]] + _force_args(case["force"][0]) + _force_args(case["force"][1])
    st, mask, fs, out = case_io.run_dropin_host(str(tmp_path), *(head + loop))
    assert st == 0, out
    assert np.array_equal(mask, case["mask"])
    for c in range(2):
        assert np.array_equal(fs[c], case["out"][("f", c)]), "component %d" % c


def test_text_restart_after_gpu_steps_rounds_oracle_state(tmp_path):
    """write_f after 60 GPU steps, restart_from_file into new fluids: the
    populations read back are the oracle's after 60 steps rounded to 6
    significant digits (the reference's text precision)."""
    case = gc.load("dry_rect_48x32_s60")
    head = ["G", 48, 32, "W", 1, "y", "F", 1.3, 0.8, "IR", 2.0, 2.0, 0.0, 0.0, 20.0, 14.0, 9.0, 6.0]
    loop = ["S", 0.1, -0.1, 0.9, 0.9, 60] + _force_args(case["force"][0]) + _force_args(case["force"][1])
    st, _, fs, out = case_io.run_dropin_host(str(tmp_path), *(head + loop + ["T", "f_after"]))
    assert st == 0, out
    s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"])
    for c in range(2):
        s.set_f(c, case["f0"][c])
    s.solid_surface_force()
    s.run(60, case["force"][0], case["force"][1])
    for c in range(2):
        want = np.char.mod("%g", s.get_f(c)).astype(np.float64)
        assert np.array_equal(fs[c], want), "component %d" % c


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_deferred_steps_equal_steps_issued_one_by_one(precision, tmp_path):
    """The drop-in queues identical fused steps and issues them as one
    scl_step2(n) at the next host read (tests/drivers/dropin_bench.cpp, the
    reference's phase-separation loop): with no read until the end, a read
    after every step (each step issued alone), a read every 5 steps and a
    flush after every step, the populations are the same bits, in FP64 and
    both deviation forms."""
    if not os.path.exists(case_io.DROPIN_BENCH):
        case_io.build_dropin_bench()
    runs = {}
    for every in (0, 1, 5, -1):
        print("dropin_bench precision %d read_every %d" % (precision, every), flush=True)
        res, fs = case_io.run_dropin_bench(str(tmp_path), 96, 23, precision, every)
        runs[every] = fs
    for every in (1, 5, -1):
        for c in range(2):
            assert np.array_equal(runs[0][c], runs[every][c]), (every, c)
    assert np.all(np.isfinite(runs[0][0])) and np.all(np.isfinite(runs[0][1]))
