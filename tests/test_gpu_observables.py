"""Droplet observables on the device (scl_laplace_observables,
scl_contact_angle) against the restatement of the reference's
post-processing scripts (oracle/observables.py), evaluated on the densities
of the same populations downloaded from the device."""
import numpy as np
import pytest

from oracle import observables as ob
from sc_lbm_amd import native as nat
from sc_lbm_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _densities(fl, ny, nx):
    return [ob.density(f.get_f(), ny, nx) for f in fl]


def _fields(mask, rho_d, rho_b, prec=nat.FP64, G=0.9, gs=(0.0, 0.0)):
    g = nat.Geometry(mask)
    fl = [nat.Field(g, 1.0, prec, 1.0, g_solid=gs[c], g_fluid=G) for c in range(2)]
    fl[0].set_f(wl.f_from_rho(rho_d, mask))
    fl[1].set_f(wl.f_from_rho(rho_b, mask))
    return g, fl


def test_laplace_observables_of_uploaded_fields():
    """A ring of known extents and two pressure sites, uploaded as
    populations rho/9: the device result equals the restatement's on the
    densities of those populations, bit for bit."""
    ny, nx = 48, 80
    m = wl.open_mask(nx, ny)
    rho_b = np.full((ny, nx), 2.0)
    rho_d = np.full((ny, nx), 0.01)
    y, x = np.mgrid[0:ny, 0:nx]
    ring = ((x - 40) ** 2 + (y - 24) ** 2 <= 15 ** 2) & ((x - 40) ** 2 + (y - 24) ** 2 > 12 ** 2)
    rho_b[ring], rho_d[ring] = 1.0, 0.5
    inside = (x - 40) ** 2 + (y - 24) ** 2 <= 12 ** 2
    rho_b[inside], rho_d[inside] = 0.03, 1.97
    g, fl = _fields(m, rho_d, rho_b)
    got = nat.laplace_observables(g, fl[0], fl[1], 0.9, (24, 40), (2, 2), 1.9, 0.06)
    rd, rb = _densities(fl, ny, nx)
    want = ob.laplace(rd, rb, 0.9, (24, 40), (2, 2), 1.9, 0.06)
    assert got == want
    assert want[1] == 15.0
    for f in fl:
        f.close()
    g.close()


@pytest.mark.parametrize("prec", [nat.FP64, nat.FP32_DEV])
def test_laplace_observables_of_a_relaxing_droplet(prec):
    """The Laplace-law driver in small (laplace_law.cpp: a square droplet of
    component 2 in component 1, G = 0.9, densities 2.0 / 0.06), 400 steps:
    device observables equal the restatement's on the downloaded state."""
    n, side = 96, 16
    m = wl.open_mask(n, n)
    y, x = np.mgrid[0:n, 0:n]
    sq = (np.abs(x - n // 2) < side) & (np.abs(y - n // 2) < side)
    rho_d = np.where(sq, 2.0, 0.06)
    rho_b = np.where(sq, 0.06, 2.0)
    g, fl = _fields(m, rho_d, rho_b, prec)
    nat.solid_surface_force(g, fl[0], fl[1])
    nat.step2(g, fl[0], fl[1], None, None, 400)
    got = nat.laplace_observables(g, fl[0], fl[1], 0.9, (n // 2, n // 2), (4, 4), 1.9, 0.06)
    rd, rb = _densities(fl, n, n)
    want = ob.laplace(rd, rb, 0.9, (n // 2, n // 2), (4, 4), 1.9, 0.06)
    assert got == want
    assert 0 < want[1] < n // 2   # (after 400 steps the droplet still oscillates: no sign for the jump)
    for f in fl:
        f.close()
    g.close()


@pytest.mark.parametrize("prec", [nat.FP64, nat.FP32_DEV])
def test_contact_angle_of_a_sessile_droplet(prec):
    """The contact-angle driver in small (contact_angles.cpp: a droplet on
    the lower wall of a channel, adhesion +-0.2): after 600 steps the
    device's (theta, L, H, R) equal the restatement's on the downloaded
    state (theta to 1e-12: atan of the C library against numpy's)."""
    nx, ny = 128, 64
    m = wl.walls_x(nx, ny)
    y, x = np.mgrid[0:ny, 0:nx]
    drop = (np.abs(x - nx // 2) < 16) & (y >= 1) & (y < 20)
    rho_d = np.where(drop, 2.0, 0.06)
    rho_b = np.where(drop, 0.06, 2.0)
    g, fl = _fields(m, rho_d, rho_b, prec, gs=(0.2, -0.2))
    nat.solid_surface_force(g, fl[0], fl[1])
    nat.step2(g, fl[0], fl[1], None, None, 600)
    got = nat.contact_angle(g, fl[0], fl[1], 2.0, 0.06)
    rd, rb = _densities(fl, ny, nx)
    want = ob.contact_angle(rd, rb, 2.0, 0.06)
    assert got[1:] == want[1:]
    assert abs(got[0] - want[0]) <= 1e-12 * abs(want[0])
    assert 0 < want[0] < 180
    for f in fl:
        f.close()
    g.close()


def test_observables_errors():
    """Sites outside the lattice and an empty set are range errors."""
    n = 32
    m = wl.open_mask(n, n)
    g, fl = _fields(m, np.full((n, n), 0.06), np.full((n, n), 2.0))
    with pytest.raises(nat.SclError) as e:
        nat.laplace_observables(g, fl[0], fl[1], 0.9, (n, 0), (0, 0), 1.9, 0.06)
    assert e.value.code == 3
    with pytest.raises(nat.SclError) as e:
        nat.laplace_observables(g, fl[0], fl[1], 0.9, (1, 1), (0, 0), 1.9, 0.06)   # rho_b above rho_bulk
    assert e.value.code == 3
    with pytest.raises(nat.SclError) as e:
        nat.contact_angle(g, fl[0], fl[1], 2.0, 0.06)
    assert e.value.code == 3
    for f in fl:
        f.close()
    g.close()
