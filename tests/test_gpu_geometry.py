"""The geometry pipeline on the device: add_walls / add_circle node sets
rasterised, packed into the kernels' bit mask and counted on the GPU
(sc_lbm_amd/csrc/geom.h), against the host sampler and rasteriser of the
served jobs (scl_job_mask) and against the compiled reference's own masks
(tests/golden/geometry_cases.npz: Geometry::add_walls, add_circle,
src/geometry.cpp:34-67, 132-175; src/geom_object/ellipse.cpp:12-33)."""
import os

import numpy as np
import pytest

from sc_lbm_amd import native as nat
from sc_lbm_amd import serve
from sc_lbm_amd import workloads as wl

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module", autouse=True)
def _init():
    nat.init(0)


def pack(mask):
    """ny x nx ints as the kernels' bit mask (independent of the library)."""
    ny, nx = mask.shape
    wpr = (nx + 63) // 64
    padded = np.zeros((ny, wpr * 64), dtype=np.uint8)
    padded[:, :nx] = mask != 0
    return np.packbits(padded, axis=1, bitorder="little").view(np.uint64).reshape(ny, wpr)


@pytest.mark.parametrize("nx,ny,por,seed,walls,rmin,rmax", [
    (130, 77, 0.7, 3, True, 8, 20),
    (200, 64, 0.85, 11, False, 4, 9),
    (64, 48, 1.0, 0, True, 8, 20),
    (257, 129, 0.9, 77, True, 2, 6),
    (1000, 600, 0.6, 12345, False, 8, 20),
    (2048, 2048, 0.6, 2025, True, 8, 20),
    (2048, 2048, 0.95, 7, False, 1, 3),
    (2048, 1024, 0.75, 99, True, 30, 40),
])
def test_device_mask_is_the_host_mask(nx, ny, por, seed, walls, rmin, rmax):
    spec = dict(nx=nx, ny=ny, steps=1, porosity=por, rmin=rmin, rmax=rmax, walls=walls, geom_seed=seed)
    print("case", spec, flush=True)
    host = serve.job_mask(**spec)
    bits, fluid, (t_sample, t_dev) = serve.job_mask_bits(**spec)
    assert np.array_equal(bits, pack(host))
    assert np.array_equal(serve.unpack_bits(bits, nx), host)
    assert fluid == int(host.sum())
    assert t_sample >= 0.0 and t_dev > 0.0
    if por < 1.0 and nx * ny <= 300000:   # the numpy sampler (workloads) draws the same disks
        ref, disks = wl.porous_mask(nx, ny, por, seed, rmin, rmax, walls=walls)
        assert np.array_equal(host, ref)
        # and the same circles through the reference-shaped entry point
        b2, f2, _ = serve.mask_raster(nx, ny, disks, wall_thickness=1 if walls else 0, walls_along=1 if walls else 0)
        assert np.array_equal(b2, bits) and f2 == fluid


def test_caller_mask_is_packed():
    rng = np.random.default_rng(4)
    m = (rng.random((33, 70)) > 0.4).astype(np.int32)
    bits, fluid, _ = serve.job_mask_bits(nx=70, ny=33, steps=1, mask=m)
    assert np.array_equal(bits, pack(m)) and fluid == int(m.sum())


def test_reference_geometry_cases():
    """Every add_walls + add_circle case of the compiled reference: the same
    node sets, and an error wherever the reference threw."""
    d = np.load(os.path.join(GOLD, "geometry_cases.npz"))
    k = 0
    while "spec_%d" % k in d:
        spec = d["spec_%d" % k]
        nx, ny, dh, wx = [int(v) for v in spec[:4]]
        circ = spec[4:].reshape(-1, 3)
        print("geometry case", k, nx, ny, dh, wx, circ.tolist(), flush=True)
        try:
            bits, fluid, _ = serve.mask_raster(nx, ny, circ, wall_thickness=dh, walls_along=(1 if wx else 2) if dh > 0 else 0)
            rc = 0
        except nat.SclError:
            rc = 1
        assert rc == int(d["rc_%d" % k])
        if rc == 0:
            want = d["mask_%d" % k].astype(np.int32)
            assert np.array_equal(serve.unpack_bits(bits, nx), want)
            assert fluid == int(want.sum())
        k += 1
    assert k >= 6


def test_large_and_overlapping_circles():
    """Circles that overlap each other and the walls, even and odd sizes, and
    one as large as the lattice: the host rasteriser of the oracle gives the
    same node sets."""
    from oracle import pyoracle as po
    nx, ny = 301, 257
    circ = [(257, 150, 128), (40, 30, 30), (41, 50, 40), (1, 0, 0), (2, 299, 255), (101, 250, 200)]
    m = np.ones((ny, nx), np.int32)
    m = po.add_walls(m, 3, "y")
    for (dd, xc, yc) in circ:
        m = po.add_circle(m, dd, xc, yc)
    bits, fluid, _ = serve.mask_raster(nx, ny, circ, wall_thickness=3, walls_along=2)
    assert np.array_equal(serve.unpack_bits(bits, nx), m)
    assert fluid == int(m.sum())


def test_served_jobs_run_on_device_masks():
    """A served porous job (its mask rasterised on the device when the
    server builds it) gives the bits of the same job with the host mask
    passed in explicitly."""
    job = dict(nx=192, ny=96, steps=40, porosity=0.75, rmin=4, rmax=9, walls=True, geom_seed=31, ic=serve.IC_DROPLET,
               rho_hi=1.0, rho_lo=0.06, radius=20, noise=0.01, ic_seed=5, tau=(1.0, 0.8), g_fluid=(0.9, 0.9),
               deliver=serve.DELIVER_F)
    m = serve.job_mask(**job)
    out = []
    with serve.Server(precision=nat.FP32_DEV, budget_bytes=1 << 30, max_jobs=8) as srv:
        for spec in (job, dict(job, mask=m, porosity=1.0)):
            i = srv.submit(**spec)
            assert srv.wait(i, 60000)["state"] == "done"
            out.append([srv.result(i, serve.F, c) for c in range(2)])
    for c in range(2):
        assert np.array_equal(out[0][c], out[1][c])
