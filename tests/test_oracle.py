"""Pins the CPU oracle (oracle/sc_oracle.cpp) before anything is checked
against it.

Pins, per oracle function:
  density, velocities, feq (both forms), repulsion, equilibrium velocities,
  collide (1 and 2 comp), volume force, stream, solid surface force
      -> golden vectors of the compiled reference (every field, every case),
         plus invariants: feq(u=0) = w*rho (tests/fluid/two_fluids_test_main.cpp:236-267),
         feq moments, stream = per-direction cyclic shift on an open periodic
         lattice, stream is a permutation of fluid slots, bounce-back at a flat
         wall, uniform psi gives zero repulsion, flat-wall adhesion S_y = -1/6,
         A<->B label swap gives bitwise swapped results, translation
         equivariance, mass conservation and momentum gain g per fluid site.
  add_walls / add_circle / add_rectangle -> geometry fixtures of the reference.
  init_droplet / init_rectangle -> golden initial populations.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from sc_lbm_amd import workloads as wl
from tests import golden_cases as gc

FIELD = {"f": po.F, "feq": po.FEQ, "rho": po.RHO, "ux": po.UX, "uy": po.UY, "ueqx": po.UEQX,
         "ueqy": po.UEQY, "frx": po.FRX, "fry": po.FRY, "fsx": po.FSX, "fsy": po.FSY}


def run_golden_on_oracle(case, kind="oracle"):
    s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"], kind=kind)
    for c in range(case["ncomp"]):
        s.set_f(c, case["f0"][c])
    if case["ncomp"] == 2:
        s.solid_surface_force()
    s.run(case["steps"], case["force"][0], case["force"][1] if case["ncomp"] == 2 else None)
    return s


@pytest.mark.parametrize("name", gc.names())
def test_oracle_matches_reference_golden(name):
    case = gc.load(name)
    s = run_golden_on_oracle(case)
    for (nm, c), want in case["out"].items():
        got = s.get(c, FIELD[nm])
        assert np.array_equal(got, want), "%s: field %s comp %d differs (max %g)" % (
            name, nm, c, np.max(np.abs(got - want)))


def test_geometry_matches_reference_fixtures():
    d = np.load(gc.GOLDEN + "/geometry_cases.npz")
    k = 0
    while "spec_%d" % k in d:
        spec = d["spec_%d" % k]
        nx, ny, dh, wx = [int(v) for v in spec[:4]]
        circ = spec[4:].reshape(-1, 3)
        m = np.ones((ny, nx), np.int32)
        rc = 0
        try:
            if dh > 0:
                m = po.add_walls(m, dh, "x" if wx else "y")
            for (dd, xc, yc) in circ:
                m = po.add_circle(m, int(dd), int(xc), int(yc))
        except RuntimeError:
            rc = 1
        assert rc == int(d["rc_%d" % k])
        if rc == 0:
            assert np.array_equal(m, d["mask_%d" % k])
            # the numpy rasteriser used by the workload builders agrees too
            m2 = np.ones((ny, nx), np.int32)
            if dh > 0:
                m2 = po.add_walls(m2, dh, "x" if wx else "y")
            for (dd, xc, yc) in circ:
                wl.add_circle(m2, int(dd), int(xc), int(yc))
            assert np.array_equal(m2, m)
        k += 1
    k = 0
    while "rect_spec_%d" % k in d:
        nx, ny, lx, ly, xc, yc = [int(v) for v in d["rect_spec_%d" % k]]
        try:
            m = po.add_rectangle(np.ones((ny, nx), np.int32), lx, ly, xc, yc)
            rc = 0
        except RuntimeError:
            rc = 1
        assert rc == int(d["rect_rc_%d" % k])
        if rc == 0:
            assert np.array_equal(m, d["rect_mask_%d" % k])
        k += 1


def test_initialisers_match_reference():
    for name, args, kind in [("init_droplet_40x36_s1", (1.0, 0.9, 0.05, 0.04, 20.5, 18.0, 9.5), "droplet"),
                             ("dry_rect_48x32_s60", (2.0, 2.0, 0.0, 0.0, 20.0, 14.0, 9.0, 6.0), "rect")]:
        case = gc.load(name)
        s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"])
        if kind == "droplet":
            s.init_droplet(*args)
        else:
            assert s.init_rectangle(*args) == 0
        for c in range(2):
            assert np.array_equal(s.get_f(c), case["f0"][c])


@pytest.mark.skipif(not po.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_matches_compiled_reference_live(seed):
    rng = np.random.default_rng(seed)
    nx, ny = int(rng.integers(5, 40)), int(rng.integers(5, 40))
    mask = (rng.random((ny, nx)) > 0.25).astype(np.int32)
    tau = list(rng.uniform(0.55, 1.6, 2))
    gf = list(rng.uniform(0.5, 2.0, 2))
    gs = list(rng.uniform(-0.5, 0.5, 2))
    f = [rng.uniform(0, 0.3, 9 * nx * ny) * np.tile(mask.ravel(), 9) for _ in range(2)]
    forces = [rng.uniform(-1e-4, 1e-4, 9), rng.uniform(-1e-4, 1e-4, 9)]
    sys = [po.System(mask, tau, gf, gs, kind=k) for k in ("oracle", "ref")]
    for s in sys:
        for c in range(2):
            s.set_f(c, f[c])
        s.solid_surface_force()
        s.run(25, forces[0], forces[1])
    for nm, fld in FIELD.items():
        for c in range(2):
            assert np.array_equal(sys[0].get(c, fld), sys[1].get(c, fld)), (nm, c)
    sys1 = [po.System(mask, tau[:1], kind=k) for k in ("oracle", "ref")]
    for s in sys1:
        s.set_f(0, f[0])
        s.run(25, forces[0])
    for fld in (po.F, po.FEQ, po.RHO, po.UX, po.UY):
        assert np.array_equal(sys1[0].get(0, fld), sys1[1].get(0, fld))


def test_feq_at_rest_is_weighted_density():
    # tests/fluid/two_fluids_test_main.cpp:236-267: feq(u = 0) = w_i * rho
    nx, ny = 7, 5
    mask = np.ones((ny, nx), np.int32)
    s = po.System(mask, [1.0, 1.0], [0.0, 0.0], [0.0, 0.0])
    rho = np.linspace(0.1, 2.0, nx * ny)
    s.set(0, po.RHO, rho)
    s.lib.oracle_feq_mc(s.h, 0)
    feq = s.get(0, po.FEQ).reshape(9, -1)
    for i in range(9):
        assert np.array_equal(feq[i], wl.W[i] * rho)


def test_feq_moments():
    rng = np.random.default_rng(3)
    n = 200
    mask = np.ones((1, n), np.int32)
    s = po.System(mask, [1.0, 1.0])
    rho = rng.uniform(0.05, 2.0, n)
    ux, uy = rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n)
    s.set(0, po.RHO, rho)
    s.set(0, po.UEQX, ux)
    s.set(0, po.UEQY, uy)
    s.lib.oracle_feq_mc(s.h, 0)
    feq = s.get(0, po.FEQ).reshape(9, -1)
    assert np.allclose(feq.sum(0), rho, rtol=1e-14, atol=0)
    assert np.allclose((feq * wl.CX[:, None]).sum(0), rho * ux, rtol=1e-12, atol=1e-16)
    assert np.allclose((feq * wl.CY[:, None]).sum(0), rho * uy, rtol=1e-12, atol=1e-16)


def test_stream_open_periodic_is_cyclic_shift():
    rng = np.random.default_rng(4)
    nx, ny = 9, 6
    mask = np.ones((ny, nx), np.int32)
    s = po.System(mask, [1.0])
    f = rng.random(9 * nx * ny)
    s.set_f(0, f)
    s.lib.oracle_stream(s.h)
    g = s.get_f(0).reshape(9, ny, nx)
    f = f.reshape(9, ny, nx)
    for i in range(9):
        assert np.array_equal(g[i], np.roll(np.roll(f[i], wl.CX[i], axis=1), wl.CY[i], axis=0))


def test_stream_is_permutation_and_bounces_at_wall():
    rng = np.random.default_rng(5)
    nx, ny = 11, 8
    mask = wl.walls_x(nx, ny)
    mask[4, 5] = 0
    s = po.System(mask, [1.0])
    f = rng.random((9, ny, nx)) * mask[None]
    s.set_f(0, f.ravel())
    s.lib.oracle_stream(s.h)
    g = s.get_f(0).reshape(9, ny, nx)
    fl = mask[None].repeat(9, 0) == 1
    assert np.array_equal(np.sort(g[fl]), np.sort(f[fl]))
    assert np.all(g[:, mask == 0] == 0)
    # at the first fluid row above the bottom wall, the downward populations
    # (4 S, 7 SW, 8 SE) come back as 2 N, 5 NE, 6 NW of the same site
    for x in range(nx):
        if mask[2, x] == 0 or x in (4, 5, 6):
            continue
        assert g[2, 1, x] == f[4, 1, x]
        assert g[5, 1, x] == f[7, 1, x]
        assert g[6, 1, x] == f[8, 1, x]


def test_repulsion_uniform_and_hand_stencil():
    nx, ny = 6, 5
    mask = np.ones((ny, nx), np.int32)
    s = po.System(mask, [1.0, 1.0], [0.9, 1.1], [0.0, 0.0])
    s.set(0, po.RHO, np.full(nx * ny, 0.7))
    s.set(1, po.RHO, np.full(nx * ny, 1.3))
    s.lib.oracle_repulsion(s.h)
    for c in range(2):
        assert np.all(s.get(c, po.FRX) == 0) and np.all(s.get(c, po.FRY) == 0)
    # one site with a single bright neighbour to the east: F_A = -G_A psi_A (1/9) psi_B(east)
    rb = np.zeros(nx * ny)
    rb[2 * nx + 3] = 2.0
    ra = np.full(nx * ny, 0.5)
    s.set(0, po.RHO, ra)
    s.set(1, po.RHO, rb)
    s.lib.oracle_repulsion(s.h)
    fx = s.get(0, po.FRX)
    assert fx[2 * nx + 2] == (1.0 / 9 * 2.0) * (-0.9 * 0.5)
    assert fx[2 * nx + 4] == (-(1.0 / 9) * 2.0) * (-0.9 * 0.5)


def test_flat_wall_adhesion():
    nx, ny = 8, 6
    mask = wl.walls_x(nx, ny)
    gs = [0.3, -0.2]
    s = po.System(mask, [1.0, 1.0], [0.0, 0.0], gs)
    s.solid_surface_force()
    for c in range(2):
        fsy = s.get(c, po.FSY).reshape(ny, nx)
        fsx = s.get(c, po.FSX).reshape(ny, nx)
        # S_y = -(1/9 + 1/36 + 1/36) = -1/6 next to the bottom wall, +1/6 under the top one
        assert np.allclose(fsy[1], (-1.0 * gs[c]) * (-1.0 / 6), rtol=1e-15)
        assert np.allclose(fsy[ny - 2], (-1.0 * gs[c]) * (1.0 / 6), rtol=1e-15)
        assert np.all(fsx == 0)
        assert np.all(fsy[2:ny - 2] == 0)


def test_label_swap_is_exact():
    c = wl.droplet(32, 8, seed=5)
    f = c["f"]
    a = po.System(c["mask"], [1.0, 1.0], [0.9, 0.9], [0.0, 0.0])
    b = po.System(c["mask"], [1.0, 1.0], [0.9, 0.9], [0.0, 0.0])
    a.set_f(0, f[0]); a.set_f(1, f[1])
    b.set_f(0, f[1]); b.set_f(1, f[0])
    for s in (a, b):
        s.solid_surface_force()
        s.run(40)
    assert np.array_equal(a.get_f(0), b.get_f(1))
    assert np.array_equal(a.get_f(1), b.get_f(0))


def test_translation_equivariance():
    c = wl.droplet(24, 6, seed=9)
    f = [x.reshape(9, 24, 24) for x in c["f"]]
    sh = (5, 7)
    a = po.System(c["mask"], [1.0, 0.8], [1.0, 1.0], [0, 0])
    b = po.System(c["mask"], [1.0, 0.8], [1.0, 1.0], [0, 0])
    for k in range(2):
        a.set_f(k, f[k].ravel())
        b.set_f(k, np.roll(f[k], sh, axis=(1, 2)).ravel())
    for s in (a, b):
        s.solid_surface_force()
        s.run(30)
    for k in range(2):
        ga = a.get_f(k).reshape(9, 24, 24)
        gb = b.get_f(k).reshape(9, 24, 24)
        assert np.array_equal(np.roll(ga, sh, axis=(1, 2)), gb)


def test_mass_conservation_and_momentum_gain():
    nx, ny = 16, 12
    mask = np.ones((ny, nx), np.int32)
    s = po.System(mask, [0.9])
    s.set_f(0, wl.f_from_rho(np.full((ny, nx), 1.0), mask))
    g = 1e-4
    m0 = s.moments(0)
    steps = 50
    s.run(steps, wl.body_force(g))
    m1 = s.moments(0)
    assert abs(m1[0] - m0[0]) < 1e-12
    assert abs(m1[1] - g * nx * ny * steps) < 1e-10
    assert abs(m1[2]) < 1e-12


def test_poiseuille_profile_closed_form():
    # a narrow channel run to steady state against the parabolic profile
    nx, ny, g, tau = 4, 20, 1e-6, 1.0
    c = wl.poiseuille(nx, ny, tau=tau, g=g)
    s = po.System(c["mask"], [tau])
    s.set_f(0, c["f"][0])
    s.run(8000, c["force"][0])
    s.lib.oracle_density(s.h, 0)
    s.lib.oracle_velocities(s.h, 0)
    ux = s.get(0, po.UX).reshape(ny, nx)[:, 0]
    want = wl.poiseuille_profile(ny, g, tau)
    inner = slice(1, ny - 1)
    assert np.max(np.abs(ux[inner] - want[inner])) < 2e-2 * want.max()


# ---- droplet observables (restatement of the reference's post-processing
#      scripts, oracle/observables.py), pinned by closed forms --------------

from oracle import observables as ob


def test_observables_pressure_of_uniform_states():
    """p = (rho_b + rho_d)/3 + G rho_b rho_d / 3 on uniform states with exact
    values: 3/3 = 1; (1 + 1)/3 + 3*1*1/3 = 5/3; and the Laplace difference of
    two such states."""
    assert ob.pressure(np.array(3.0), np.array(0.0), 0.9) == 1.0
    assert abs(ob.pressure(np.array(1.0), np.array(1.0), 3.0) - 5.0 / 3.0) < 1e-15
    n = 16
    rho_b = np.full((n, n), 1.5)
    rho_d = np.full((n, n), 0.25)
    rho_b[4, 5], rho_d[4, 5] = 3.0, 0.0      # p = 1 at (y 4, x 5)
    rho_b[9, 9], rho_d[9, 9] = 1.0, 1.0      # p = 5/3 at (y 9, x 9) with G = 3
    dp, _ = ob.laplace(rho_d, rho_b, 3.0, (9, 9), (4, 5), 1.9, 0.06)
    assert abs(dp - 2.0 / 3.0) < 1e-15


def test_observables_radius_of_a_known_set():
    """The Laplace set (rho_d < rho_b, rho_disp < rho_b <= rho_bulk) placed on
    a ring of known extents: rows 7..31 and columns 12..40 give
    R = 0.5 * ((31 - 7)/2 + (40 - 12)/2) = 13; sites failing any one of the
    three conditions do not count."""
    ny, nx = 48, 64
    rho_b = np.full((ny, nx), 2.0)           # bulk: above rho_bulk, not in the set
    rho_d = np.full((ny, nx), 0.01)
    y, x = np.mgrid[0:ny, 0:nx]
    ring = ((y == 7) | (y == 31)) & (x >= 12) & (x <= 40) | ((x == 12) | (x == 40)) & (y >= 7) & (y <= 31)
    rho_b[ring], rho_d[ring] = 1.0, 0.5
    rho_b[20, 2], rho_d[20, 2] = 1.0, 1.5    # rho_d > rho_b
    rho_b[2, 20], rho_d[2, 20] = 0.05, 0.01  # rho_b below rho_disp
    rho_b[45, 60], rho_d[45, 60] = 1.95, 0.5  # rho_b above rho_bulk
    _, R = ob.laplace(rho_d, rho_b, 0.9, (20, 26), (2, 2), 1.9, 0.06)
    assert R == 13.0


@pytest.mark.parametrize("h,half_base", [(5, 5), (3, 9), (8, 4), (6, 6)])
def test_observables_contact_angle_of_a_constructed_cap(h, half_base):
    """An interface whose lowest row spans x = a..b and the next row
    a+1..b-1, top row h above: the script extrapolates the base to
    L = b - a + 1 and the height to H = h + 0.5, and the circle through them
    meets the base at theta = 2 atan(2H / L) (the cap angle of elementary
    geometry), 90 degrees for L = 2H."""
    ny, nx = 40, 64
    rho_b = np.full((ny, nx), 2.0)
    rho_d = np.full((ny, nx), 0.01)
    y0, a = 3, 20
    b = a + 2 * half_base
    iface = np.zeros((ny, nx), bool)
    iface[y0, a:b + 1] = True
    iface[y0 + 1, a + 1:b] = True
    iface[y0 + h, (a + b) // 2] = True
    rho_b[iface], rho_d[iface] = 0.5, 1.0
    theta, L, H, R = ob.contact_angle(rho_d, rho_b, 2.0, 0.06)
    assert L == b - a + 1 and H == h + 0.5
    assert abs(R - (4 * H * H + L * L) / (8 * H)) < 1e-12
    assert abs(theta - np.degrees(2 * np.arctan(2 * H / L))) < 1e-9
    if L == 2 * H:
        assert theta == 90.0


def test_domain_length_statistic():
    """The domain-size statistic of the spinodal tests: for a single Fourier
    mode phi = cos(2 pi k x / n) the structure factor sits at |k| = 2 pi k / n,
    so L = n / k exactly, along x and along y; two modes of equal power give
    the harmonic mean of their wavelengths weighted as the first moment."""
    from tests.gpu_helpers import domain_length
    n = 256
    x = np.arange(n)
    for k in (4, 16):
        phi = np.cos(2 * np.pi * k * x / n)[None, :].repeat(n, 0)
        assert abs(domain_length(phi) - n / k) < 1e-9
        assert abs(domain_length(phi.T) - n / k) < 1e-9
    phi = (np.cos(2 * np.pi * 4 * x / n) + np.cos(2 * np.pi * 12 * x / n))[None, :].repeat(n, 0)
    assert abs(domain_length(phi) - n / 8.0) < 1e-9
