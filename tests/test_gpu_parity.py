"""Parity of the HIP path (libsclbm.so, called through its C-ABI) with the
reference: golden vectors of the compiled reference, the oracle on the same
seeded inputs, and size-independent properties at full size.

The FP64 form is held to bit-exact equality (values compared with ==; the
reference's operation order is reproduced exactly).  The precision forms are
held to the tolerances stated in each test."""
import numpy as np
import pytest

from oracle import pyoracle as po
from sc_lbm_amd import native as nat
from sc_lbm_amd import workloads as wl
from tests import golden_cases as gc
from tests import gpu_helpers as gh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    nat.init(0)
    yield
    nat.synchronize()


def _golden_check(name, fl, fields=True):
    case = gc.load(name)
    for (nm, c), want in case["out"].items():
        if nm != "f" and not fields:
            continue
        got = fl[c].download(gh.NAMES[nm])
        assert np.array_equal(got, want, equal_nan=True), "%s: %s[%d] differs, max rel %g" % (
            name, nm, c, gh.max_rel(got, want))


@pytest.mark.parametrize("name", gc.names())
def test_fused_matches_reference_golden(name):
    case = gc.load(name)
    _, fl = gh.run_fused(case)
    _golden_check(name, fl)


@pytest.mark.parametrize("name", gc.names())
def test_per_operator_matches_reference_golden(name):
    case = gc.load(name)
    _, fl = gh.run_perop(case)
    _golden_check(name, fl)


def _oracle_run(case, steps):
    s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"])
    for c in range(len(case["tau"])):
        s.set_f(c, case["f"][c])
    if len(case["tau"]) == 2:
        s.solid_surface_force()
    s.run(steps, case["force"][0], case["force"][1] if len(case["tau"]) == 2 else None)
    return s


CONFIGS = {
    # the 200x50 Poiseuille channel, full size, 1000 steps
    "poiseuille_200x50": lambda: wl.poiseuille(200, 50, tau=1.0, rho0=1.0, g=1e-6),
    # the static droplet at half resolution (128^2, r = 20)
    "droplet_128": lambda: wl.droplet(128, 20, seed=3),
    # the porous two-component case on a 256x128 slice (disk radii 4..8)
    "porous_256x128": lambda: wl.porous(256, 128, porosity=0.7, seed=5, rmin=4, rmax=8),
}


@pytest.mark.parametrize("cfg", sorted(CONFIGS))
def test_fused_fp64_matches_oracle_1000_steps(cfg):
    case = CONFIGS[cfg]()
    steps = 1000
    _, fl = gh.run_fused(case, steps)
    s = _oracle_run(case, steps)
    for c in range(len(fl)):
        got, want = fl[c].get_f(), s.get_f(c)
        # bit-exact; the stated bar is 1e-12 relative
        assert gh.max_rel(got, want) <= 1e-12
        assert np.array_equal(got, want), "%s comp %d: max rel %g" % (cfg, c, gh.max_rel(got, want))


def test_tile_shape_does_not_change_results():
    case = wl.porous(200, 96, porosity=0.75, seed=8, rmin=4, rmax=7)
    ref = None
    for tile in [(64, 4, 1), (128, 7, 2), (256, 16, 1), (256, 16, 2), (512, 3, 2), (64, 96, 1)]:
        _, fl = gh.run_fused(case, 60, tile=tile)
        got = [f.get_f() for f in fl]
        if ref is None:
            ref = got
        else:
            for c in range(2):
                assert np.array_equal(got[c], ref[c]), tile


def test_chunked_calls_and_reads_between_steps():
    case = wl.porous(96, 64, porosity=0.8, seed=2, rmin=4, rmax=6)
    _, one = gh.run_fused(case, 40)
    g, fl = gh.make(case)
    for k in range(4):
        nat.step2(g, fl[0], fl[1], case["force"][0], None, 10)
        # reading intermediates and populations between calls must not
        # disturb the run (replay + gather into staging)
        fl[0].download(nat.RHO)
        fl[1].download(nat.UEQX)
        fl[k % 2].get_f()
    for c in range(2):
        assert np.array_equal(fl[c].get_f(), one[c].get_f())
    s = _oracle_run(case, 40)
    for c in range(2):
        for nm in ("rho", "ux", "uy", "ueqx", "ueqy", "frx", "fry", "feq"):
            fld = {"rho": po.RHO, "ux": po.UX, "uy": po.UY, "ueqx": po.UEQX, "ueqy": po.UEQY,
                   "frx": po.FRX, "fry": po.FRY, "feq": po.FEQ}[nm]
            assert np.array_equal(fl[c].download(gh.NAMES[nm]), s.get(c, fld)), (nm, c)


def test_fused_and_per_operator_interleave():
    case = wl.droplet(64, 12, seed=4)
    g, fl = gh.make(case)
    nat.step2(g, fl[0], fl[1], None, None, 7)
    for _ in range(3):
        nat.compute_density(fl[0])
        nat.compute_density(fl[1])
        nat.fluid_repulsion(g, fl[0], fl[1])
        nat.equilibrium_velocities(g, fl[0], fl[1])
        nat.collide2(fl[0], fl[1])
        nat.stream2(g, fl[0], fl[1])
    nat.step2(g, fl[0], fl[1], None, None, 5)
    s = _oracle_run(case, 15)
    for c in range(2):
        assert np.array_equal(fl[c].get_f(), s.get_f(c))


def test_two_component_step_requires_solid_force():
    case = wl.droplet(32, 6, seed=1)
    g = nat.Geometry(case["mask"])
    a = nat.Field(g, 1.0, g_fluid=0.9)
    b = nat.Field(g, 1.0, g_fluid=0.9)
    with pytest.raises(nat.RangeError):
        nat.step2(g, a, b, None, None, 1)
    with pytest.raises(nat.RangeError):
        nat.equilibrium_velocities(g, a, b)


def test_label_swap_bitwise():
    case = wl.porous(128, 64, porosity=0.8, seed=9, rmin=4, rmax=6)
    case["g_solid"] = [0.1, 0.1]
    case["force"] = [None, None]
    sw = dict(case, f=[case["f"][1], case["f"][0]])
    _, a = gh.run_fused(case, 50)
    _, b = gh.run_fused(sw, 50)
    assert np.array_equal(a[0].get_f(), b[1].get_f())
    assert np.array_equal(a[1].get_f(), b[0].get_f())


def test_translation_equivariance_large():
    n, m = 1024, 512
    rng = np.random.default_rng(0)
    ra = 1.0 + 0.01 * (2 * rng.random((m, n)) - 1)
    rb = 1.0 + 0.01 * (2 * rng.random((m, n)) - 1)
    mask = wl.open_mask(n, m)
    f = [wl.f_from_rho(ra, mask), wl.f_from_rho(rb, mask)]
    case = dict(mask=mask, tau=[1.0, 1.0], g_fluid=[1.2, 1.2], g_solid=[0.0, 0.0], force=[None, None], f=f)
    sh = (37, 101)
    rolled = [np.roll(x.reshape(9, m, n), sh, axis=(1, 2)).ravel() for x in f]
    case2 = dict(case, f=rolled)
    _, a = gh.run_fused(case, 30)
    _, b = gh.run_fused(case2, 30)
    for c in range(2):
        ga = np.roll(a[c].get_f().reshape(9, m, n), sh, axis=(1, 2))
        assert np.array_equal(ga.ravel(), b[c].get_f())


def test_mass_conservation_4096():
    # the 4096^2 spinodal lattice at full size: FP64 mass is conserved to
    # rounding, momentum stays at rounding level (no force, periodic)
    case = wl.spinodal(4096, G=1.2, seed=777)
    g, fl = gh.make(case)
    m0 = [f.moments() for f in fl]
    nat.step2(g, fl[0], fl[1], None, None, 20)
    m1 = [f.moments() for f in fl]
    for c in range(2):
        assert abs(m1[c][0] - m0[c][0]) <= 1e-12 * abs(m0[c][0])
    # the pair forces cancel between the components, so only the total
    # momentum is conserved (it starts at zero: f = rho/9 in every direction)
    assert abs(m1[0][1] + m1[1][1]) < 1e-6 and abs(m1[0][2] + m1[1][2]) < 1e-6
    for f in fl:
        f.close()


def test_repeat_runs_are_bitwise_identical():
    case = wl.porous(1024, 512, porosity=0.7, seed=2024)
    _, a = gh.run_fused(case, 20)
    _, b = gh.run_fused(case, 20)
    for c in range(2):
        assert np.array_equal(a[c].get_f(), b[c].get_f())


@pytest.mark.parametrize("prec,tol", [(nat.FP32_DEV, 1e-5), (nat.FP16_DEV, 2e-2)])
def test_precision_forms_vs_fp64(prec, tol):
    # droplet 128^2, 1000 steps: max density error of the deviation forms
    # against the FP64 GPU run (itself bit-exact with the oracle)
    case = wl.droplet(128, 20, seed=3)
    _, ref = gh.run_fused(case, 1000)
    _, low = gh.run_fused(case, 1000, precision=prec, rho0=(0.5, 0.5))
    for c in range(2):
        r = ref[c].get_f().reshape(9, -1).sum(0)
        l = low[c].get_f().reshape(9, -1).sum(0)
        err = float(np.max(np.abs(r - l)))
        print("precision %d comp %d max |drho| %.3e" % (prec, c, err))
        assert err <= tol


def test_one_component_precision_forms_poiseuille():
    case = wl.poiseuille(200, 50, tau=1.0, rho0=1.0, g=1e-6)
    _, ref = gh.run_fused(case, 2000)
    r = ref[0].get_f().reshape(9, 50, 200)
    for prec, tol in ((nat.FP32_DEV, 1e-6), (nat.FP16_DEV, 1e-4)):
        _, low = gh.run_fused(case, 2000, precision=prec, rho0=(1.0,))
        l = low[0].get_f().reshape(9, 50, 200)
        ur = (r * wl.CX[:, None, None]).sum(0)
        ul = (l * wl.CX[:, None, None]).sum(0)
        err = float(np.max(np.abs(ur - ul)))
        print("precision %d max |d jx| %.3e (u_max ~ 1.7e-3)" % (prec, err))
        assert err <= tol


def test_destroy_geometry_before_fields():
    # a geometry destroyed while fields still use it stays alive until the
    # last field goes; the field destroy may replay the last step, which
    # reads the geometry's mask
    case = wl.porous(512, 256, porosity=0.8, seed=4, rmin=4, rmax=8)
    g, fl = gh.run_fused(case, 5)
    fl2 = [nat.Field(g, 1.0, g_fluid=0.9) for _ in range(2)]
    g.close()
    with pytest.raises(nat.SclError):
        nat.Field.__new__(nat.Field).__init__(g, 1.0)   # handle is gone on the Python side
    rho = fl[0].download(nat.RHO)                       # replay after the geometry handle was dropped
    assert np.isfinite(rho).all()
    fl[1].close()                                       # materializes nothing left; frees
    fl[0].close()
    for f in fl2:
        f.close()


def test_autotuned_shape_gives_same_results():
    case = wl.porous(300, 140, porosity=0.8, seed=12, rmin=4, rmax=8)
    _, ref = gh.run_fused(case, 30)
    g, fl = gh.make(case)
    shape = g.autotune(2, nat.FP64)
    assert shape == g.tile(2, nat.FP64)
    assert shape[0] in (64, 128, 256, 512) and shape[2] in (1, 2)
    nat.step2(g, fl[0], fl[1], case["force"][0], None, 30)
    for c in range(2):
        assert np.array_equal(fl[c].get_f(), ref[c].get_f())


def test_geometry_change_between_fused_steps():
    """Fused steps on one mask, then on another: the populations at the
    switch are gathered with the mask that produced them (its bounce-back
    ghost slots included), and the next steps start from that streamed
    state, as the reference's loop does when its Geometry changes."""
    case = wl.droplet(n=96, radius=20, G=0.9, seed=8)
    # the second mask only adds solids: a site that turns fluid starts with
    # zero density, which is singular for the Shan-Chen step (both the
    # reference and this library then produce NaN there)
    m1 = wl.add_circle(wl.open_mask(96, 96), 11, 20, 20)
    m2 = wl.add_circle(wl.add_circle(wl.walls_x(96, 96), 11, 20, 20), 15, 70, 60)
    f0 = [np.where(np.tile(m1.ravel(), 9) == 1, case["f"][c], 0.0) for c in range(2)]
    g1, g2 = nat.Geometry(m1), nat.Geometry(m2)
    fl = [nat.Field(g1, case["tau"][c], nat.FP64, 0.0, g_fluid=case["g_fluid"][c]) for c in range(2)]
    for c in range(2):
        fl[c].set_f(f0[c])
    nat.solid_surface_force(g1, fl[0], fl[1])
    nat.step2(g1, fl[0], fl[1], None, None, 40)
    nat.step2(g2, fl[0], fl[1], None, None, 35)
    got = [fl[c].get_f() for c in range(2)]
    s1 = po.System(m1, case["tau"], case["g_fluid"], [0.0, 0.0])
    for c in range(2):
        s1.set_f(c, f0[c])
    s1.solid_surface_force()
    s1.run(40)
    s2 = po.System(m2, case["tau"], case["g_fluid"], [0.0, 0.0])
    for c in range(2):
        s2.set_f(c, s1.get_f(c))
    s2.solid_surface_force()
    s2.run(35)
    for c in range(2):
        assert np.all(np.isfinite(got[c]))
        assert np.array_equal(got[c].reshape(-1), s2.get_f(c))


def _dry_case(nx=320, ny=96):
    """Component B absent (density exactly 0) outside a rectangle: exercises
    the zero-density rule (stale equilibrium velocity) of the deviation forms."""
    m = wl.add_circle(wl.walls_x(nx, ny), 21, nx // 3, ny // 2)
    ra = wl.noisy(np.full((ny, nx), 1.0), 3, 0.01)
    rb = np.zeros((ny, nx))
    rb[20:60, 150:260] = 0.8
    return dict(mask=m, tau=[1.0, 0.8], g_fluid=[0.9, 0.9], g_solid=[0.1, -0.1],
                f=[wl.f_from_rho(ra, m), wl.f_from_rho(rb, m)], force=[wl.body_force(2e-5), None])


@pytest.mark.parametrize("prec", [nat.FP32_DEV, nat.FP16_DEV])
@pytest.mark.parametrize("name", ["open", "open_force", "porous", "dry"])
def test_two_steps_per_launch_equal_single_steps(prec, name):
    """The two- and three-step kernels (the intermediate steps stay on chip)
    give the populations of single-step launches bit for bit, on lattices
    whose width is not a multiple of the strip, with solids, adhesion, a
    force and dry sites, and solid-free with and without a force."""
    if name == "open":
        case = wl.spinodal(300, G=1.2, seed=4)
        case["force"] = [None, None]
    elif name == "open_force":
        case = wl.spinodal(300, G=1.2, seed=7)
        case["force"] = [wl.body_force(1e-5), None]
    elif name == "porous":
        case = wl.porous(nx=700, ny=230, porosity=0.75, seed=6, rmin=5, rmax=12)
    else:
        case = _dry_case()
    out = []
    for fusion in (1, 2, 3):
        g = nat.Geometry(case["mask"])
        g.set_step_fusion(fusion, 24)
        assert g.step_fusion(prec)[0] == fusion
        fl = [nat.Field(g, case["tau"][c], prec, 1.0, g_solid=case["g_solid"][c], g_fluid=case["g_fluid"][c])
              for c in range(2)]
        for c in range(2):
            fl[c].set_f(case["f"][c])
        nat.solid_surface_force(g, fl[0], fl[1])
        for n in (1, 6, 9, 2, 5, 4):   # bootstrap step, then chunks of every remainder
            nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], n)
        out.append([fl[c].get_f() for c in range(2)])
        for f in fl:
            f.close()
        g.close()
    for c in range(2):
        assert np.all(np.isfinite(out[0][c]))
        assert np.array_equal(out[0][c], out[1][c])
        assert np.array_equal(out[0][c], out[2][c])


def _fp64_run(case, fusion, steps, watch=None, chunks=None):
    """FP64 populations after `steps` steps in calls of `chunks` steps, with
    `fusion` steps per launch, and the number of launches timed."""
    g = nat.Geometry(case["mask"])
    g.set_step_fusion(fusion, 24)
    if watch is not None:
        g.set_fp64_fusion_watch(watch)
    fl = [nat.Field(g, case["tau"][c], nat.FP64, 0.0, g_solid=case["g_solid"][c], g_fluid=case["g_fluid"][c])
          for c in range(2)]
    for c in range(2):
        fl[c].set_f(case["f"][c])
    nat.solid_surface_force(g, fl[0], fl[1])
    nat.timing(1)
    for n in chunks or [steps]:
        nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], n)
    _, launches = nat.timing(0)
    nat.timing(-1)
    out = [f.get_f() for f in fl]
    fused = g.step_fusion(nat.FP64)[0]
    for f in fl:
        f.close()
    g.close()
    return out, launches, fused


@pytest.mark.parametrize("name", ["open", "porous", "dry"])
def test_fp64_two_steps_per_launch_equal_single_steps(name):
    """FP64 with two steps per launch (k_step2d, the reference's arithmetic)
    gives the populations of single steps bit for bit, and so the oracle's:
    solid-free, porous with adhesion and a force on A, and a lattice with dry
    sites, where the zero-density rule hands every step back to single
    launches.  Calls of every length mix both kernels."""
    if name == "open":
        case = wl.spinodal(300, G=1.2, seed=4)
        case["force"] = [None, None]
    elif name == "porous":
        case = wl.porous(nx=700, ny=230, porosity=0.75, seed=6, rmin=5, rmax=12)
    else:
        case = _dry_case()
    chunks = [1, 6, 9, 2, 5, 4, 3, 31]
    want, n1, f1 = _fp64_run(case, 1, sum(chunks), chunks=chunks)
    got, n2, f2 = _fp64_run(case, 2, sum(chunks), chunks=chunks)
    assert (f1, f2) == (1, 2)
    for c in range(2):
        assert np.all(np.isfinite(got[c]))
        assert np.array_equal(got[c].view(np.uint64), want[c].view(np.uint64)), c
    assert n1 == sum(chunks)
    if name == "dry":
        assert n2 > n1 // 2   # handed back: launches of single steps
    else:
        assert n2 < n1 * 2 // 3
    if name == "porous":   # the oracle, on the same inputs
        s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"])
        for c in range(2):
            s.set_f(c, case["f"][c])
        s.solid_surface_force()
        s.run(sum(chunks), case["force"][0], case["force"][1])
        for c in range(2):
            assert np.array_equal(got[c].reshape(-1), s.get_f(c))


@pytest.mark.parametrize("name", ["open", "porous"])
def test_fp64_reads_after_a_call_ending_in_two_steps(name):
    """An FP64 call that ends in a two-step launch never stores its last
    step's input; a read of the reference's intermediate arrays (density,
    velocities, equilibrium velocities, repulsion, feq) recomputes it from
    the launch's input and replays the last step.  With reads after every
    call (the first call ends in a single step, the next two in two-step
    launches) the arrays and the populations equal those of single steps bit
    for bit, and the last ones the oracle's."""
    if name == "open":
        case = wl.spinodal(256, G=1.2, seed=5)
        case["force"] = [None, None]
    else:
        case = wl.porous(nx=300, ny=64, porosity=0.8, seed=3, rmin=4, rmax=7)
    names = ("rho", "ux", "uy", "ueqx", "ueqy", "frx", "fry", "feq")
    out = {}
    for fusion in (1, 2):
        g, fl = gh.make(case)
        g.set_step_fusion(fusion, 0)
        assert g.step_fusion(nat.FP64)[0] == fusion
        reads = []
        for n in (4, 6, 2):
            nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], n)
            reads.append([[fl[c].download(gh.NAMES[nm]) for nm in names] + [fl[c].get_f()] for c in range(2)])
        out[fusion] = reads
        for f in fl:
            f.close()
        g.close()
    for k in range(3):
        for c in range(2):
            for i, nm in enumerate(names + ("f",)):
                assert np.array_equal(out[1][k][c][i], out[2][k][c][i]), (k, c, nm)
    s = _oracle_run(case, 12)
    fields = {"rho": po.RHO, "ux": po.UX, "uy": po.UY, "ueqx": po.UEQX, "ueqy": po.UEQY,
              "frx": po.FRX, "fry": po.FRY, "feq": po.FEQ}
    for c in range(2):
        for i, nm in enumerate(names):
            assert np.array_equal(out[2][2][c][i], s.get(c, fields[nm])), (nm, c)
        assert np.array_equal(out[2][2][c][-1], s.get_f(c))


def test_fp64_two_steps_hand_over_in_the_middle_of_a_call():
    """A density falling below the hand-over threshold after some two-step
    launches (a spinodal lattice separating: its least density passes 0.5
    at step 38 of 80, threshold 0.5): the launch that met it and those
    queued after it are repeated one step per launch from its input, bit
    for bit the single-step result."""
    case = wl.spinodal(300, G=1.2, seed=8)
    case["force"] = [None, None]
    want, n1, _ = _fp64_run(case, 1, 80, chunks=[1, 79])
    clean, n_clean, _ = _fp64_run(case, 2, 80, chunks=[1, 79])
    got, n2, _ = _fp64_run(case, 2, 80, watch=0.5, chunks=[1, 79])
    for c in range(2):
        assert np.array_equal(clean[c].view(np.uint64), want[c].view(np.uint64))
        assert np.array_equal(got[c].view(np.uint64), want[c].view(np.uint64))
    # launches: the bootstrap, 39 two-step launches and the last step
    # (n_clean = 41); with the hand-over at launch j (1-based) also the
    # 79 - 2 (j - 1) steps from its input one per launch, 121 - 2 j in all;
    # j = 1 (a hand-over at the call's first launch) would give 119
    assert (n1, n_clean) == (80, 41)
    assert n_clean < n2 < 119 and (121 - n2) % 2 == 0
    rho = [want[c].reshape(9, -1).sum(0) for c in range(2)]
    assert min(float(r.min()) for r in rho) < 0.5


@pytest.mark.parametrize("prec", [nat.FP32_DEV, nat.FP16_DEV])
@pytest.mark.parametrize("name", ["spinodal", "porous"])
def test_multi_step_kernels_at_every_band_height_remainder(prec, name):
    """The multi-step tiles walk their rows in a loop unrolled by 3 or 6,
    with ring slots fixed per unrolled row: band heights 20..25 end the
    walk at every remainder (a band runs its height plus 8 or 10 rows), and
    the last band of the 300-row lattice is partial.  Two and three steps
    per launch equal single steps bit for bit at each, for the solid-free
    tiles (spinodal) and for the tiles with solids (porous: walls, disks,
    adhesion +-0.1 and a body force on A)."""
    if name == "spinodal":
        case = wl.spinodal(300, G=1.2, seed=9)
    else:
        case = wl.porous(nx=300, ny=300, porosity=0.75, seed=13, rmin=5, rmax=12)
    forces = case.get("force", [None, None])

    def run(fusion, band):
        g = nat.Geometry(case["mask"])
        g.set_step_fusion(fusion, band)
        fl = [nat.Field(g, case["tau"][c], prec, 1.0, g_solid=case["g_solid"][c], g_fluid=case["g_fluid"][c])
              for c in range(2)]
        for c in range(2):
            fl[c].set_f(case["f"][c])
        nat.solid_surface_force(g, fl[0], fl[1])
        nat.step2(g, fl[0], fl[1], forces[0], forces[1], 13)   # bootstrap step, then 12
        out = [f.get_f() for f in fl]
        for f in fl:
            f.close()
        g.close()
        return out

    want = run(1, 0)
    for band in range(20, 26):
        for fusion in (2, 3):
            got = run(fusion, band)
            for c in range(2):
                assert np.array_equal(got[c], want[c]), (fusion, band, c)


@pytest.mark.parametrize("prec", [nat.FP64, nat.FP32_DEV, nat.FP16_DEV])
@pytest.mark.parametrize("name", ["porous", "dry"])
def test_checkpoint_restore_continues_bit_for_bit(prec, name):
    """A checkpoint taken after 31 steps (post-collision phase, with solids,
    adhesion, a force and dry sites) restored into fresh fields continues
    bit for bit as the original: 26 more steps, in chunks that mix the
    multi-step kernels with single steps."""
    case = wl.porous(nx=300, ny=90, porosity=0.75, seed=3, rmin=5, rmax=10) if name == "porous" else _dry_case()

    def fields(g):
        fl = [nat.Field(g, case["tau"][c], prec, 1.0, g_solid=case["g_solid"][c], g_fluid=case["g_fluid"][c])
              for c in range(2)]
        nat.solid_surface_force(g, fl[0], fl[1])
        return fl

    g = nat.Geometry(case["mask"])
    fl = fields(g)
    for c in range(2):
        fl[c].set_f(case["f"][c])
    nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], 31)
    ck = [f.checkpoint() for f in fl]
    for n in (7, 19):
        nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], n)
    want = [f.get_f() for f in fl]
    g2 = nat.Geometry(case["mask"])
    fl2 = fields(g2)
    for c in range(2):
        fl2[c].restore(ck[c])
    for n in (7, 19):
        nat.step2(g2, fl2[0], fl2[1], case["force"][0], case["force"][1], n)
    for c in range(2):
        got = fl2[c].get_f()
        assert np.all(np.isfinite(got))
        assert np.array_equal(got.view(np.uint64), want[c].view(np.uint64))
    # a checkpoint refuses a field of another precision or size
    other = nat.Field(g2, 1.0, nat.FP32_DEV if prec != nat.FP32_DEV else nat.FP64, 1.0)
    with pytest.raises(nat.SclError):
        other.restore(ck[0])
    # ... and a field of the same size and precision on another mask (its
    # solid sites would hold the wrong bounce-back slots)
    mask3 = case["mask"].copy()
    fy, fx = np.argwhere(mask3 == 1)[len(np.argwhere(mask3 == 1)) // 2]
    mask3[fy, fx] = 0
    g3 = nat.Geometry(mask3)
    on_other_mask = nat.Field(g3, case["tau"][0], prec, 1.0)
    with pytest.raises(nat.SclError, match="another geometry"):
        on_other_mask.restore(ck[0])
    for f in fl + fl2 + [other, on_other_mask]:
        f.close()
    g.close()
    g2.close()
    g3.close()


def test_automatic_step_fusion_follows_the_lattice_size():
    """Untuned (fusion 0, no autotuning) the steps per launch follow the
    lattice size on this 256-CU part: on the 1024 x 512 porous lattice FP32
    storage takes two per launch (three-step bands would be under 16 rows),
    FP16 one (two-step bands under 10 rows with solids) and FP64 two, at
    bands of one whole wave (two would give 8 rows); on the 4096^2 open
    lattice FP32 three, FP16 two and FP64 two in bands of 16 rows.  Which one
    runs never changes the bits (the fused-against-single tests)."""
    small = nat.Geometry(wl.porous(nx=1024, ny=512, porosity=0.70, seed=2024)["mask"])
    big = nat.Geometry(np.ones((4096, 4096), dtype=np.int32))
    try:
        assert small.step_fusion(nat.FP32_DEV)[0] == 2
        assert small.step_fusion(nat.FP16_DEV)[0] == 1
        k, rows = small.step_fusion(nat.FP64)
        assert k == 2 and rows > 8
        assert big.step_fusion(nat.FP32_DEV)[0] == 3
        assert big.step_fusion(nat.FP16_DEV)[0] == 2
        assert big.step_fusion(nat.FP64) == (2, 16)
    finally:
        small.close()
        big.close()


@pytest.mark.parametrize("prec", [nat.FP64, nat.FP32_DEV, nat.FP16_DEV])
@pytest.mark.parametrize("shape", [(300, 5), (300, 6), (300, 8), (256, 40), (257, 33)])
def test_multi_step_kernels_on_edge_shapes(prec, shape):
    """Lattices where the multi-step kernels' halos wrap: 8 rows (the
    fewest a three-step tile takes: its band halo of 9 rows wraps the
    lattice), 6 (the fewest for two steps; three fall back to two) and 5
    (single steps), 256 columns (one two-step strip exactly) and 257 (a
    strip of one column).  Solids and adhesion included; fusion 1, 2 and 3
    give the same bits."""
    nx, ny = shape
    rng = np.random.default_rng(nx * 1000 + ny)
    m = np.ones((ny, nx), dtype=np.int32)
    m[rng.random((ny, nx)) < 0.05] = 0
    ra = 1.0 + 0.02 * rng.standard_normal((ny, nx))
    rb = np.where(rng.random((ny, nx)) < 0.5, 0.3, 0.9)
    out = []
    for fusion in (1, 2, 3):
        g = nat.Geometry(m)
        g.set_step_fusion(fusion, 0)
        # a K-step tile needs 2K + 2 rows; shorter lattices take fewer steps
        # per launch (FP64: at most two)
        kmax = 2 if prec == nat.FP64 else 3
        assert g.step_fusion(prec)[0] == max(k for k in (1, 2, 3)
                                             if k <= min(fusion, kmax) and (k == 1 or ny >= 2 * k + 2))
        fl = [nat.Field(g, (1.0, 0.8)[c], prec, 1.0, g_solid=(0.1, -0.1)[c], g_fluid=0.9) for c in range(2)]
        fl[0].set_f(wl.f_from_rho(ra, m))
        fl[1].set_f(wl.f_from_rho(rb, m))
        nat.solid_surface_force(g, fl[0], fl[1])
        for n in (1, 8, 6, 5):
            nat.step2(g, fl[0], fl[1], None, None, n)
        out.append([f.get_f() for f in fl])
        for f in fl:
            f.close()
        g.close()
    for c in range(2):
        assert np.all(np.isfinite(out[0][c]))
        assert np.array_equal(out[0][c], out[1][c])
        assert np.array_equal(out[0][c], out[2][c])


def test_zero_density_rule_without_held_velocity_is_reported():
    """The reference keeps a component's previous equilibrium velocity where
    |rho| <= 1e-16 (src/lbm.cpp:262-269).  The fused kernels hold it only where
    rho < 1e-8, so a density that falls from >= 1e-8 to <= 1e-16 within one
    step is reported (SCL_ERR_INEXACT) instead of silently reusing an older
    value.  Forced here: one site of component A holds f_E = a, f_W = b with
    (a - b)/(a + b) = u where 1.5 u^2 rounds to exactly 1, so at tau = 1 its
    rest population relaxes to feq_0 = (4/9) rho (1 - 1.5 u^2) = 0; every
    other population of A is 0 and B is empty, so after one step that site's
    density is exactly 0 after being 1.  The per-operator path keeps every
    site's velocity and reports nothing."""
    nx = ny = 16
    fa = np.zeros((9, ny, nx))
    fa[1, 8, 8] = 0.908248290463863
    fa[3, 8, 8] = 0.09175170953613698
    u = (fa[1, 8, 8] - fa[3, 8, 8]) / (fa[1, 8, 8] + fa[3, 8, 8])
    assert 1.0 - 1.5 * (u * u) == 0.0
    case = dict(mask=wl.open_mask(nx, ny), tau=[1.0, 1.0], g_fluid=[0.0, 0.0], g_solid=[0.0, 0.0],
                force=[None, None], f=[fa.ravel(), np.zeros(9 * nx * ny)])
    s = _oracle_run(case, 2)
    site = 8 * nx + 8
    assert s.get(0, po.RHO)[site] == 0.0 and s.get(0, po.UEQX)[site] == u   # the rule applied at step 2
    g, fl = gh.run_fused(case, 2)
    with pytest.raises(nat.InexactError, match="zero-density rule"):
        fl[0].get_f()
    # reported once; the populations are the reference's (the density there is 0)
    for c in range(2):
        assert np.array_equal(fl[c].get_f(), s.get_f(c))
    g2, pl = gh.run_perop(case, 2)
    for c in range(2):
        assert np.array_equal(pl[c].get_f(), s.get_f(c))
        assert np.array_equal(pl[c].download(nat.UEQX), s.get(c, po.UEQX))
    # one more fused step after the per-operator ones: every velocity was held
    nat.step2(g2, pl[0], pl[1], None, None, 1)
    s.run(1, None, None)
    for c in range(2):
        assert np.array_equal(pl[c].get_f(), s.get_f(c))
