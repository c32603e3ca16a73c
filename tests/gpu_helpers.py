"""Shared helpers for the GPU tests: run a case through libsclbm.so."""
import numpy as np

from sc_lbm_amd import native as nat

NAMES = {"f": nat.F, "feq": nat.FEQ, "rho": nat.RHO, "ux": nat.UX, "uy": nat.UY, "ueqx": nat.UEQX,
         "ueqy": nat.UEQY, "frx": nat.FRX, "fry": nat.FRY, "fsx": nat.FSX, "fsy": nat.FSY}


def make(case, precision=nat.FP64, rho0=(1.0, 1.0), tile=None):
    g = nat.Geometry(case["mask"])
    if tile:
        g.set_tile(tile[0], tile[1])
        if len(tile) > 2:
            g.set_lookahead(tile[2])
    nc = len(case["tau"])
    fl = []
    for c in range(nc):
        f = nat.Field(g, case["tau"][c], precision, rho0[c] if rho0 else 1.0,
                      g_solid=case["g_solid"][c] if nc == 2 else 0.0,
                      g_fluid=case["g_fluid"][c] if nc == 2 else 0.0)
        f.set_f(case["f0"][c] if "f0" in case else case["f"][c])
        fl.append(f)
    if nc == 2:
        nat.solid_surface_force(g, fl[0], fl[1])
    return g, fl


def run_fused(case, steps=None, precision=nat.FP64, rho0=(1.0, 1.0), tile=None, chunks=None):
    g, fl = make(case, precision, rho0, tile)
    steps = case["steps"] if steps is None else steps
    forces = case.get("force", [None, None])
    parts = chunks or [steps]
    for n in parts:
        if len(fl) == 1:
            nat.step1(g, fl[0], forces[0], n)
        else:
            nat.step2(g, fl[0], fl[1], forces[0], forces[1], n)
    return g, fl


def run_perop(case, steps=None):
    """The reference's per-method sequence, one C-ABI call per method."""
    g, fl = make(case)
    steps = case["steps"] if steps is None else steps
    forces = case.get("force", [None, None])
    for _ in range(steps):
        if len(fl) == 1:
            nat.collide1(g, fl[0])
            if forces[0] is not None:
                nat.volume_force(g, fl[0], forces[0])
            nat.stream1(g, fl[0])
        else:
            nat.compute_density(fl[0])
            nat.compute_density(fl[1])
            nat.fluid_repulsion(g, fl[0], fl[1])
            nat.equilibrium_velocities(g, fl[0], fl[1])
            nat.collide2(fl[0], fl[1])
            for c in range(2):
                if forces[c] is not None:
                    nat.volume_force(g, fl[c], forces[c])
            nat.stream2(g, fl[0], fl[1])
    return g, fl


def max_rel(a, b):
    """The reference tests' metric: |a-b| / max(1, |a|, |b|)."""
    d = np.abs(a - b) / np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))
    return float(np.max(d)) if d.size else 0.0


def domain_length(phi):
    """Domain size of a phase field: 2 pi / k1, k1 the first moment of the
    radially averaged structure factor of phi - <phi> (phi real, square)."""
    n = phi.shape[0]
    s = np.abs(np.fft.rfft2(phi - phi.mean())) ** 2
    ky = np.fft.fftfreq(n) * 2 * np.pi
    kx = np.fft.rfftfreq(n) * 2 * np.pi
    kk = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    w = np.full(kx.size, 2.0)          # rfft columns stand for two of the full plane
    w[0] = 1.0
    if n % 2 == 0:
        w[-1] = 1.0
    s[0, 0] = 0.0
    return 2 * np.pi * (w * s).sum() / (w * kk * s).sum()
