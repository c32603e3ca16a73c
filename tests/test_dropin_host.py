"""The drop-in classes' host code, run on the CPU through a driver written
against include/geometry.h, lbm.h and fluid.h (tests/drivers/dropin_host.cpp),
compared with what the compiled reference wrote for the same calls:

- Geometry builders: add_walls, add_circle, add_rectangle
  (tests/golden/geometry_cases.npz), add_ellipse, add_square and add_array in
  both forms (tests/golden/geometry_objects.npz), including the bounds errors
  and their exception types (src/geometry.cpp:34-240);
- LBM initialisers: initialize_droplet and initialize_fluid_rectangle give the
  initial populations of init_droplet_40x36_s1 and dry_rect_48x32_s60
  (src/lbm.cpp:11-99) bit for bit;
- Fluid::simple_ini writes rho/9 on fluid sites and 0 on solids
  (src/fluid.cpp:67-104; the reference's own check is
  tests/fluid/two_fluids_test_main.cpp:176-233);
- Fluid::initialize_randomly_perturbed_density with a seeded RNG equals an
  independent restatement: numpy's MT19937 (legacy 32-bit seeding, the same
  generator as std::mt19937), libstdc++'s generate_canonical (two 32-bit
  draws, (a + b*2^32) / 2^64) and uniform_real_distribution (U*(b-a) + a), in
  the reference's x-outer, y-inner order (src/fluid.cpp:106-151);
- write_f / restart_from_file: the text files are byte-identical to the
  compiled reference's write_f of the same populations, and a restart reads
  back each value rounded to 6 significant digits (src/fluid.cpp:285-353).
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.drivers import case_io

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _driver():
    if not os.path.exists(case_io.DROPIN_HOST):
        case_io.build_dropin_host()


def _geometry_cmds(spec):
    nx, ny, dh, wx = (int(v) for v in spec[:4])
    cmds = ["G", nx, ny]
    if dh > 0:
        cmds += ["W", dh, "x" if wx else "y"]
    circ = [int(v) for v in spec[4:]]
    for k in range(0, len(circ), 3):
        cmds += ["C"] + circ[k:k + 3]
    return cmds


# the reference's exception of each status: 1 runtime_error, 2 invalid_argument
REF_TO_HOST = {0: 0, 1: 1, 2: 2}


@pytest.mark.parametrize("k", range(6))
def test_walls_and_circles_match_reference(k, tmp_path):
    g = np.load(os.path.join(GOLDEN, "geometry_cases.npz"))
    st, mask, _, _ = case_io.run_dropin_host(str(tmp_path), *_geometry_cmds(g["spec_%d" % k]))
    assert st == REF_TO_HOST[int(g["rc_%d" % k])]
    if st == 0:
        assert np.array_equal(mask, g["mask_%d" % k])


@pytest.mark.parametrize("k", range(4))
def test_rectangles_match_reference(k, tmp_path):
    g = np.load(os.path.join(GOLDEN, "geometry_cases.npz"))
    nx, ny, lx, ly, xc, yc = (int(v) for v in g["rect_spec_%d" % k])
    st, mask, _, _ = case_io.run_dropin_host(str(tmp_path), "G", nx, ny, "R", lx, ly, xc, yc)
    assert st == int(g["rect_rc_%d" % k])
    if st == 0:
        assert np.array_equal(mask, g["rect_mask_%d" % k])


def _object_cases():
    return len([k for k in np.load(os.path.join(GOLDEN, "geometry_objects.npz")).files if k.startswith("rc_")])


@pytest.mark.parametrize("k", range(_object_cases()))
def test_ellipses_squares_and_arrays_match_reference(k, tmp_path):
    """Status codes of the reference: 0 ok, 1 runtime_error, 2
    invalid_argument, 3 out_of_range; the driver reports the same codes."""
    g = np.load(os.path.join(GOLDEN, "geometry_objects.npz"))
    spec = [int(v) for v in g["spec_%d" % k]]
    nx, ny, kind, u, v = spec[:5]
    bounds, props = spec[5:9], spec[9:]
    typ = str(g["type_%d" % k])
    cmds = ["G", nx, ny]
    if kind == 0:
        cmds += ["E"] + props
    elif kind == 1:
        cmds += ["Q"] + props
    else:
        cmds += ["AD" if kind == 2 else "AN", typ, len(props)] + props + bounds + [u, v]
    st, mask, _, out = case_io.run_dropin_host(str(tmp_path), *cmds)
    assert st == int(g["rc_%d" % k]), out
    if st == 0:
        assert np.array_equal(mask, g["mask_%d" % k])


def test_initialize_droplet_matches_reference(tmp_path):
    d = np.load(os.path.join(GOLDEN, "init_droplet_40x36_s1.npz"))
    # Geometry(40, 36) with a disk (9, 12, 10); the initialiser's arguments
    # are those tests/golden/make_golden.py gave the reference
    st, mask, fs, _ = case_io.run_dropin_host(
        str(tmp_path), "G", 40, 36, "C", 9, 12, 10, "F", 1.0, 1.0,
        "ID", 1.0, 0.9, 0.05, 0.04, 20.5, 18.0, 9.5)
    assert st == 0
    assert np.array_equal(mask, d["mask"])
    assert np.array_equal(fs[0], d["f0_0"])
    assert np.array_equal(fs[1], d["f0_1"])


def test_initialize_fluid_rectangle_matches_reference(tmp_path):
    d = np.load(os.path.join(GOLDEN, "dry_rect_48x32_s60.npz"))
    st, mask, fs, _ = case_io.run_dropin_host(
        str(tmp_path), "G", 48, 32, "W", 1, "y", "F", 1.3, 0.8,
        "IR", 2.0, 2.0, 0.0, 0.0, 20.0, 14.0, 9.0, 6.0)
    assert st == 0
    assert np.array_equal(mask, d["mask"])
    assert np.array_equal(fs[0], d["f0_0"])
    assert np.array_equal(fs[1], d["f0_1"])


def test_simple_ini_writes_rho_over_nine_on_fluid_sites(tmp_path):
    st, mask, fs, _ = case_io.run_dropin_host(str(tmp_path), "G", 20, 10, "W", 1, "x", "C", 5, 10, 5,
                                              "F", 1.0, 1.0, "IS", 2.0, 0.7)
    assert st == 0
    fluid = np.tile(mask.reshape(-1) == 1, 9)
    for f, rho in zip(fs, (2.0, 0.7)):
        assert np.array_equal(f[fluid], np.full(fluid.sum(), rho / 9.0))
        assert not np.any(f[~fluid])
    assert np.array_equal(fs[0][fluid], np.full(fluid.sum(), 2.0 / 9))   # the reference KAT's 2/9


def _mt_uniform(seed, n, lo, hi):
    """n draws of std::uniform_real_distribution<double>(lo, hi) from
    std::mt19937(seed), restated with numpy's MT19937."""
    raw = np.random.RandomState(seed)._bit_generator.random_raw(2 * n).astype(np.float64)
    u = (raw[0::2] + raw[1::2] * 2.0 ** 32) / 2.0 ** 64
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return u * (hi - lo) + lo


def test_randomly_perturbed_density_restated(tmp_path):
    nx, ny, seed = 24, 14, 12345
    st, mask, fs, _ = case_io.run_dropin_host(str(tmp_path), "G", nx, ny, "W", 1, "x", "C", 5, 12, 7,
                                              "F", 1.0, 1.0, "IP", 1.0, 0.5, seed)
    assert st == 0
    # x outer, y inner, fluid sites only; A draws first, then B from the same generator
    order = [(x, y) for x in range(nx) for y in range(ny) if mask[y, x] == 1]
    draws = _mt_uniform(seed, 2 * len(order), 0.0, 1.0)
    for c, rho0 in enumerate((1.0, 0.5)):
        d = draws[c * len(order):(c + 1) * len(order)]
        rho = np.zeros((ny, nx))
        for (x, y), u in zip(order, d):
            rho[y, x] = rho0 + (u * (rho0 / 1000.0 - 0.0) + 0.0)
        want = np.tile((rho / 9.0).reshape(-1), 9)
        want[np.tile(mask.reshape(-1) == 0, 9)] = 0.0
        assert np.array_equal(fs[c], want), "component %d" % c


def test_text_round_trip_rounds_to_six_digits(tmp_path):
    d = np.load(os.path.join(GOLDEN, "init_droplet_40x36_s1.npz"))
    st, _, fs, _ = case_io.run_dropin_host(
        str(tmp_path), "G", 40, 36, "C", 9, 12, 10, "F", 1.0, 1.0,
        "ID", 1.0, 0.9, 0.05, 0.04, 20.5, 18.0, 9.5, "T", "f_text")
    assert st == 0
    for c in range(2):
        want = np.char.mod("%g", d["f0_%d" % c]).astype(np.float64)
        assert np.array_equal(fs[c], want)
        assert not np.array_equal(fs[c], d["f0_%d" % c])    # 0.05/9 etc. do lose digits


@pytest.mark.skipif(not po.ref_available(), reason="compiled reference not built (oracle/_ref)")
def test_text_files_byte_identical_to_reference(tmp_path):
    d = np.load(os.path.join(GOLDEN, "dry_rect_48x32_s60.npz"))
    st, _, _, _ = case_io.run_dropin_host(
        str(tmp_path), "G", 48, 32, "W", 1, "y", "F", 1.3, 0.8,
        "IR", 2.0, 2.0, 0.0, 0.0, 20.0, 14.0, 9.0, 6.0, "T", "mine")
    assert st == 0
    s = po.System(d["mask"].astype(np.int32), list(d["tau"]), list(d["g_fluid"]), list(d["g_solid"]), kind="ref")
    for c, tag in enumerate("AB"):
        s.set_f(c, d["f0_%d" % c])
        assert po.ref_lib().ref_write_f(s.h, c, str(tmp_path / ("ref_" + tag)).encode()) == 0
        for i in range(9):
            mine = (tmp_path / ("mine_%s_%d.txt" % (tag, i))).read_bytes()
            ref = (tmp_path / ("ref_%s_%d.txt" % (tag, i))).read_bytes()
            assert mine == ref, "component %s direction %d" % (tag, i)


def test_restart_from_missing_file_raises_ios_failure(tmp_path):
    st, _, _, out = case_io.run_dropin_host(str(tmp_path), "G", 8, 8, "F", 1.0, 1.0, "IS", 1.0, 1.0,
                                            "T", "no_such_dir/f")
    assert st == 4 and "ios_base::failure" in out
