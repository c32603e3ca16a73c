"""Restatement of the reference's droplet post-processing scripts in numpy.

ORACLE / TEST INFRASTRUCTURE: only tests/ import this module, and only as
the checker of the device entry points scl_laplace_observables and
scl_contact_angle.

The reference evaluates these observables in Python from the density files
its drivers write; the functions below take the same two density arrays
(rows = y, columns = x, as np.loadtxt reads the files) and follow the
scripts' arithmetic line by line.  They are pinned in tests/test_oracle.py
by closed forms (a digital disk's radius, the pressure of uniform states, a
lens whose angle is fixed by construction).
"""
import numpy as np


def pressure(rho_b, rho_d, G12):
    """Total pressure, laplace_law/compute_surface_tension.py:63."""
    return 1.0 / 3.0 * (rho_b + rho_d) + G12 / 3.0 * rho_b * rho_d


def laplace(rho_d, rho_b, G12, p_in, p_out, rho_bulk, rho_disp):
    """(pressure difference, radius) as compute_surface_tension.py:63-70."""
    p_tot = pressure(rho_b, rho_d, G12)
    dp = p_tot[p_in[0], p_in[1]] - p_tot[p_out[0], p_out[1]]
    select_indices = np.where((rho_d - rho_b < 0) & (rho_b <= rho_bulk) & (rho_b > rho_disp))
    Rx = (np.amax(select_indices[0]) - np.amin(select_indices[0])) / 2.0
    Ry = (np.amax(select_indices[1]) - np.amin(select_indices[1])) / 2.0
    return float(dp), float(0.5 * (Rx + Ry))


def contact_angle(rho_d, rho_b, rho_bulk, rho_disp):
    """(theta in degrees, base length L, height H, radius R) as
    contact_angles/compute_contact_angles.py:46-100."""
    interface = np.where((rho_b - rho_d <= 0) & (rho_d <= rho_bulk) & (rho_b > rho_disp))
    ind_left_min = np.argmin(interface[0])
    y_left_min = interface[0][ind_left_min]
    x_left_min = interface[1][np.argwhere(interface[0] == y_left_min)]
    y_above_left_min = y_left_min + 1
    x_above_left_min = interface[1][np.argwhere(interface[0] == y_above_left_min)]
    xL1 = np.amin(x_above_left_min)
    xL2 = np.amin(x_left_min)
    xR1 = np.amax(x_above_left_min)
    xR2 = np.amax(x_left_min)
    yL1 = np.amin(y_above_left_min)
    yL2 = np.amin(y_left_min)
    yR1 = np.amax(y_above_left_min)
    yR2 = np.amax(y_left_min)
    yL = np.amin(interface[0]) - 0.5
    yR = np.amin(interface[0]) - 0.5
    if xL1 == xL2:
        xL1 = xL1 - 1
    if xR1 == xR2:
        xR1 = xR1 - 1
    slopeL = (yL1 - yL2) / (xL1 - xL2)
    slopeR = (yR1 - yR2) / (xR1 - xR2)
    interL = yL2 - slopeL * xL2
    interR = yR2 - slopeR * xR2
    xL = 1 / slopeL * (yL - interL)
    xR = 1 / slopeR * (yR - interR)
    L = xR - xL
    H = np.amax(interface[0]) - np.amin(interface[0]) + 0.5
    R = (4 * H * H + L * L) / (8 * H)
    theta = np.degrees(np.arctan([L / (2 * (R - H))]))
    if theta[0] < 0:
        theta[0] += 180
    return float(theta[0]), float(L), float(H), float(R)


def density(f, ny, nx):
    """Density of populations f[9 * ny * nx] (reference layout) in the
    reference's summation order, ((0 + f0) + f1) + ... (src/fluid.cpp:175-183)."""
    f = np.asarray(f, dtype=np.float64).reshape(9, ny, nx)
    r = np.zeros((ny, nx))
    for i in range(9):
        r = r + f[i]
    return r
