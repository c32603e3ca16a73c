"""Writes the golden vectors in tests/golden/*.npz by running the compiled
reference (oracle/_ref/libscref.so, built from the reference's own sources by
`make -C oracle ref`) through its public API.

Run in a checkout where the reference sources exist:
    make -C oracle ref && python tests/golden/make_golden.py

Each fixture holds the inputs (mask, tau, couplings, forces, initial
populations or initial densities) and the reference's outputs as raw float64
(exact bits).  The cases are small so that the whole set stays a few MB.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from sc_lbm_amd import workloads as wl  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def run_case(name, mask, tau, gf, gs, f0, forces, steps, keep_inter=True, init=None):
    s = po.System(mask, tau, gf, gs, kind="ref")
    nc = len(tau)
    if init is not None:
        kind, args = init
        if kind == "droplet":
            s.init_droplet(*args)
        else:
            rc = s.init_rectangle(*args)
            assert rc == 0
        f0 = [s.get_f(c) for c in range(nc)]
    else:
        for c in range(nc):
            s.set_f(c, f0[c])
    if nc == 2:
        s.solid_surface_force()
    fa = forces[0]
    fb = forces[1] if len(forces) > 1 else None
    s.run(steps, fa, fb)
    out = dict(mask=mask.astype(np.int8), tau=np.array(tau), g_fluid=np.array(gf),
               g_solid=np.array(gs), steps=np.int64(steps))
    for c in range(nc):
        out["f0_%d" % c] = f0[c]
        out["f_%d" % c] = s.get_f(c)
        if forces[c] is not None if c < len(forces) else False:
            out["force_%d" % c] = np.asarray(forces[c])
        if keep_inter:
            names = ["feq", "rho", "ux", "uy"]
            fields = [po.FEQ, po.RHO, po.UX, po.UY]
            if nc == 2:
                names += ["ueqx", "ueqy", "frx", "fry", "fsx", "fsy"]
                fields += [po.UEQX, po.UEQY, po.FRX, po.FRY, po.FSX, po.FSY]
            for nm, fld in zip(names, fields):
                out["%s_%d" % (nm, c)] = s.get(c, fld)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)
    print("wrote", name, {k: v.shape for k, v in out.items() if hasattr(v, "shape")})


def main():
    assert po.ref_available(), "build the reference first: make -C oracle ref"

    # 1. Single-component Poiseuille channel, 200x50, x walls, tau 1, rho0 1,
    #    g = 1e-6 (the 200x50 channel case), 200 steps.
    c = wl.poiseuille(200, 50, tau=1.0, rho0=1.0, g=1e-6)
    run_case("poiseuille_200x50_s200", c["mask"], c["tau"], [0.0], [0.0], c["f"], [c["force"][0]], 200)

    # 2. Single component, tau 0.6 (omega = 1/tau is inexact), a cylinder,
    #    y walls, body force along y, 300 steps on 60x40.
    rc, m = po.ref_geometry(60, 40, 1, "y", [(11, 30, 20)])
    assert rc == 0
    rho = 1.0 + 0.05 * wl.uniform01(7, 60 * 40).reshape(40, 60)
    run_case("channel_cyl_60x40_s300", m, [0.6], [0.0], [0.0], [wl.f_from_rho(rho, m)],
             [wl.body_force(2e-5, "y")], 300)

    # 3. Two-component droplet, 64x64 periodic, r = 16, 1.0 / 0.06 mirrored,
    #    +-1% seeded noise, G = 0.9, 300 steps.
    c = wl.droplet(64, 16, seed=11)
    run_case("droplet_64_s300", c["mask"], c["tau"], c["g_fluid"], c["g_solid"], c["f"], [None, None], 300)

    # 4. Two-component porous slice, 64x48, x walls, three disks, adhesion
    #    +0.1 / -0.1, G = 0.9, tau 1.0 / 0.7, body force 1e-5 on A only,
    #    A = 1.0 / B = 0.06 in the left 30% and reversed elsewhere, 400 steps.
    rc, m = po.ref_geometry(64, 48, 1, "x", [(13, 16, 20), (9, 40, 12), (17, 45, 33)])
    assert rc == 0
    x = np.arange(64)[None, :].repeat(48, 0)
    left = x < int(0.3 * 64)
    ra = np.where(left, 1.0, 0.06)
    rb = np.where(left, 0.06, 1.0)
    run_case("porous_64x48_s400", m, [1.0, 0.7], [0.9, 0.9], [0.1, -0.1],
             [wl.f_from_rho(ra, m), wl.f_from_rho(rb, m)], [wl.body_force(1e-5), None], 400)

    # 5. Two components with zero dissolved densities (the zero-density
    #    branch of the equilibrium velocity stays active for tens of steps),
    #    y walls, rectangle initialiser (initialize_fluid_rectangle),
    #    tau 1.3 / 0.8, adhesion +-0.1 and a force on both components,
    #    60 steps on 48x32.
    rc, m = po.ref_geometry(48, 32, 1, "y", [])
    assert rc == 0
    run_case("dry_rect_48x32_s60", m, [1.3, 0.8], [0.9, 0.9], [0.1, -0.1], None,
             [wl.body_force(-3e-5, "y"), wl.body_force(-3e-5, "y")], 60,
             init=("rect", (2.0, 2.0, 0.0, 0.0, 20.0, 14.0, 9.0, 6.0)))

    # 6. initialize_droplet on a geometry with a disk (solid sites skipped),
    #    one step so the initial state and one full update are both pinned.
    rc, m = po.ref_geometry(40, 36, 0, "x", [(9, 12, 10)])
    assert rc == 0
    run_case("init_droplet_40x36_s1", m, [1.0, 1.0], [0.9, 0.9], [0.0, 0.0], None,
             [None, None], 1, init=("droplet", (1.0, 0.9, 0.05, 0.04, 20.5, 18.0, 9.5)))

    # 7. Geometry node sets: add_walls and add_circle (odd/even diameters,
    #    odd/even lattice sizes) and add_rectangle, plus the bounds errors.
    geo = {}
    cases = [
        (33, 27, 1, "x", [(5, 10, 10), (6, 20, 13), (1, 3, 3), (2, 30, 24)]),
        (40, 40, 2, "y", [(11, 20, 20), (12, 8, 30), (3, 31, 5)]),
        (64, 48, 1, "x", [(13, 16, 20), (9, 40, 12), (17, 45, 33)]),
        (20, 20, 0, "x", [(20, 10, 10)]),   # upper bound violated -> error 1
        (20, 20, 0, "x", [(9, 3, 10)]),     # lower bound violated -> error 1
        (8, 8, 9, "x", []),                 # wall thicker than the domain -> error 1
    ]
    for k, (nx, ny, dh, where, circ) in enumerate(cases):
        rc, mm = po.ref_geometry(nx, ny, dh, where, circ)
        geo["rc_%d" % k] = np.int64(rc)
        geo["mask_%d" % k] = mm.astype(np.int8)
        geo["spec_%d" % k] = np.array([nx, ny, dh, 1 if where == "x" else 0] +
                                      [v for cc in circ for v in cc], dtype=np.int64)
    rects = [(30, 20, 7, 5, 10, 8), (30, 20, 6, 4, 20, 10), (30, 20, 1, 1, 0, 0), (30, 20, 9, 9, 3, 10)]
    for k, (nx, ny, lx, ly, xc, yc) in enumerate(rects):
        out = np.zeros(nx * ny, dtype=np.int32)
        rc = po.ref_lib().ref_add_rectangle(nx, ny, lx, ly, xc, yc,
                                            out.ctypes.data_as(po._ip))
        geo["rect_rc_%d" % k] = np.int64(rc)
        geo["rect_mask_%d" % k] = out.reshape(ny, nx).astype(np.int8)
        geo["rect_spec_%d" % k] = np.array([nx, ny, lx, ly, xc, yc], dtype=np.int64)
    np.savez_compressed(os.path.join(OUT, "geometry_cases.npz"), **geo)
    print("wrote geometry_cases")


# Object and array builders (kind 0 add_ellipse, 1 add_square, 2 add_array by
# gaps, 3 add_array by numbers): (nx, ny, kind, type, props, bounds, u, v).
# The array cases include the four array drivers' own calls
# (benchmarks/single_phase_one_component_flows/x_flow_array_ellipses.cpp,
# x_flow_array_staggered_circles.cpp, y_flow_array_squares.cpp,
# y_flow_staggered_array_rectangles.cpp).
OBJECT_CASES = [
    (30, 20, 0, "", (7, 5, 10, 8), (0, 0, 0, 0), 0, 0),
    (30, 20, 0, "", (8, 6, 12, 9), (0, 0, 0, 0), 0, 0),
    (30, 20, 0, "", (11, 3, 15, 10), (0, 0, 0, 0), 0, 0),
    (30, 20, 0, "", (9, 9, 3, 10), (0, 0, 0, 0), 0, 0),       # lower bound violated
    (30, 20, 1, "", (5, 10, 8), (0, 0, 0, 0), 0, 0),
    (30, 20, 1, "", (6, 20, 10), (0, 0, 0, 0), 0, 0),
    (30, 20, 1, "", (7, 27, 10), (0, 0, 0, 0), 0, 0),         # upper bound violated
    (200, 100, 2, "circle", (5, 10, 15), (5, 190, 3, 75), 10, 25),
    (200, 100, 2, "circle", (7, 17, 29), (10, 182, 10, 100), 15, 25),
    (100, 200, 3, "square", (5, 15, 10), (3, 70, 5, 190), 5, 20),
    (200, 100, 3, "ellipse", (9, 5, 10, 10), (3, 190, 3, 95), 9, 6),
    (100, 200, 2, "rectangle", (4, 7, 10, 10), (5, 90, 5, 190), 6, 9),
    (100, 200, 3, "Rectangle", (5, 3, 12, 20), (6, 95, 10, 150), 6, 4),
    (60, 40, 2, "ELLIPSE", (6, 4, 5, 5), (1, 55, 1, 35), 2, 3),
    (60, 40, 2, "hexagon", (5, 5, 5), (1, 55, 1, 35), 2, 2),   # invalid type
    (60, 40, 2, "circle", (5, 5, 5), (1, 80, 1, 35), 2, 2),    # past the lattice in x: wraps rows
    (60, 40, 2, "circle", (5, 5, 5), (1, 55, 1, 60), 2, 2),    # past the lattice in y: out_of_range
]


def geometry_objects():
    geo = {}
    for k, (nx, ny, kind, typ, props, bounds, u, v) in enumerate(OBJECT_CASES):
        rc, mm = po.ref_geometry_object(nx, ny, kind, props, typ, bounds, u, v)
        geo["rc_%d" % k] = np.int64(rc)
        geo["mask_%d" % k] = mm.astype(np.int8)
        geo["spec_%d" % k] = np.array([nx, ny, kind, u, v] + list(bounds) + list(props), dtype=np.int64)
        geo["type_%d" % k] = np.array(typ)
    np.savez_compressed(os.path.join(OUT, "geometry_objects.npz"), **geo)
    print("wrote geometry_objects", {k: int(geo["rc_%d" % k]) for k in range(len(OBJECT_CASES))})


if __name__ == "__main__":
    if sys.argv[1:] == ["objects"]:
        assert po.ref_available(), "build the reference first: make -C oracle ref"
        geometry_objects()
    else:
        main()
        geometry_objects()
