"""Synthetic workloads: masks and initial populations for the standard cases.

Everything here is deterministic from an integer seed (SplitMix64), so the
same case can be rebuilt bit for bit on any machine.  Conventions follow the
reference drivers:
  * walls "x" are solid rows y = 0 and y = ny-1 (src/geometry.cpp:53-62);
  * disks are placed with add_circle semantics (src/geom_object/ellipse.cpp:12-33);
  * initial populations are rho/9 in every direction on fluid sites and 0 on
    solid sites (src/fluid.cpp:96-107, src/lbm.cpp:33-46);
  * a body force g along x is the per-direction vector (g/6)*c_x
    (benchmarks/single_phase_one_component_flows/laminar_flow_x_dir.cpp:21-24).
"""
import numpy as np

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)

_M64 = (1 << 64) - 1


def splitmix64(seed, n):
    """n successive SplitMix64 outputs (uint64) from the given seed."""
    with np.errstate(over="ignore"):
        s = (np.uint64(seed & _M64) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n):
    """n doubles in [0, 1): the top 53 bits of SplitMix64 outputs."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


class Stream:
    """Sequential uniform draws from one seed (used by the samplers)."""

    def __init__(self, seed, block=4096):
        self.seed = seed
        self.block = block
        self.k = 0
        self.buf = np.empty(0)
        self.pos = 0

    def next(self):
        if self.pos >= self.buf.size:
            z = splitmix64(self.seed + self.k * 0x632BE59BD9B4E019, self.block)
            self.buf = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
            self.k += 1
            self.pos = 0
        v = self.buf[self.pos]
        self.pos += 1
        return float(v)


def body_force(g, axis="x"):
    c = CX if axis == "x" else CY
    beta = 1.0 / 6.0
    return np.array([float(v) * (g * beta) for v in c])


def open_mask(nx, ny):
    return np.ones((ny, nx), dtype=np.int32)


def walls_x(nx, ny, dh=1):
    m = open_mask(nx, ny)
    m[:dh, :] = 0
    m[ny - dh:, :] = 0
    return m


def circle_nodes(d, xc, yc):
    """Node list of add_circle(d, xc, yc): even d shrinks by one, integer
    test ((x-xc)^2*L^2)/4 + ((y-yc)^2*L^2)/4 <= L^4/16 on the box xc+-L/2."""
    L = d - 1 if d % 2 == 0 else d
    h = L // 2
    xs = np.arange(xc - h, xc + h + 1)
    ys = np.arange(yc - h, yc + h + 1)
    X, Y = np.meshgrid(xs, ys)
    lhs = ((X - xc) ** 2 * L * L) // 4 + ((Y - yc) ** 2 * L * L) // 4
    keep = lhs <= (L * L * L * L) // 16
    return X[keep], Y[keep]


def add_circle(mask, d, xc, yc):
    ny, nx = mask.shape
    if xc - d // 2 < 0 or yc - d // 2 < 0 or xc + d // 2 >= nx or yc + d // 2 >= ny:
        raise RuntimeError("Circle bounds do not fit the domain.")
    X, Y = circle_nodes(d, xc, yc)
    mask[Y, X] = 0
    return mask


def porous_mask(nx, ny, porosity, seed, rmin=8, rmax=20, walls=True, max_tries=200000):
    """Non-overlapping solid disks with integer radii uniform in [rmin, rmax]
    (diameter 2r+1), centres uniform over the positions add_circle accepts,
    placed by rejection sampling until the fluid fraction (wall rows
    included) is at most `porosity`.  Returns (mask, disks)."""
    m = walls_x(nx, ny) if walls else open_mask(nx, ny)
    st = Stream(seed)
    disks = []
    y_lo = 1 if walls else 0
    y_hi = ny - 2 if walls else ny - 1
    tries = 0
    while m.mean() > porosity and tries < max_tries:
        tries += 1
        r = rmin + int(st.next() * (rmax - rmin + 1))
        r = min(r, rmax)
        d = 2 * r + 1
        lo_x, hi_x = r, nx - 1 - r
        lo_y, hi_y = y_lo + r, y_hi - r
        if hi_x < lo_x or hi_y < lo_y:
            continue
        xc = lo_x + int(st.next() * (hi_x - lo_x + 1))
        yc = lo_y + int(st.next() * (hi_y - lo_y + 1))
        ok = True
        for (d2, x2, y2) in disks:
            r2 = d2 // 2
            if (xc - x2) ** 2 + (yc - y2) ** 2 <= (r + r2 + 1) ** 2:
                ok = False
                break
        if not ok:
            continue
        add_circle(m, d, xc, yc)
        disks.append((d, xc, yc))
    return m, disks


def f_from_rho(rho, mask):
    """Populations rho/9 in all nine directions on fluid sites, 0 on solids."""
    r = np.where(mask == 1, rho, 0.0).ravel() / 9
    return np.tile(r, 9)


def noisy(rho, seed, amp=0.01):
    """rho * (1 + U[-amp, amp)) per site from a seeded stream."""
    u = uniform01(seed, rho.size).reshape(rho.shape)
    return rho * (1.0 + (2.0 * u - 1.0) * amp)


# --------------------------------------------------------------------------
# The standard cases.  Each returns a dict with the mask, per-component tau,
# couplings, forces and initial populations (FP64, reference layout).
# --------------------------------------------------------------------------

def poiseuille(nx=200, ny=50, tau=1.0, rho0=1.0, g=1e-6):
    m = walls_x(nx, ny)
    return dict(name="poiseuille_%dx%d" % (nx, ny), mask=m, tau=[tau], g_fluid=[0.0], g_solid=[0.0],
                f=[f_from_rho(np.full((ny, nx), rho0), m)], force=[body_force(g), None])


def droplet(n=256, radius=40, rho_in=1.0, rho_out=0.06, G=0.9, tau=1.0, seed=12345, noise=0.01):
    m = open_mask(n, n)
    y, x = np.mgrid[0:n, 0:n]
    c = n // 2
    inside = (x - c) ** 2 + (y - c) ** 2 <= radius * radius
    ra = np.where(inside, rho_in, rho_out)
    rb = np.where(inside, rho_out, rho_in)
    ra = noisy(ra, seed, noise)
    rb = noisy(rb, seed + 1, noise)
    return dict(name="droplet_%d" % n, mask=m, tau=[tau, tau], g_fluid=[G, G], g_solid=[0.0, 0.0],
                f=[f_from_rho(ra, m), f_from_rho(rb, m)], force=[None, None])


def porous(nx=1024, ny=512, porosity=0.70, seed=2024, G=0.9, gads=0.1, tau=1.0, g=1e-5,
           rmin=8, rmax=20):
    m, disks = porous_mask(nx, ny, porosity, seed, rmin, rmax)
    x = np.arange(nx)[None, :].repeat(ny, 0)
    left = x < int(0.3 * nx)
    ra = np.where(left, 1.0, 0.06)
    rb = np.where(left, 0.06, 1.0)
    return dict(name="porous_%dx%d" % (nx, ny), mask=m, tau=[tau, tau], g_fluid=[G, G],
                g_solid=[gads, -gads], f=[f_from_rho(ra, m), f_from_rho(rb, m)],
                force=[body_force(g), None], disks=disks)


def spinodal(n=4096, G=1.2, tau=1.0, seed=777, noise=0.01):
    m = open_mask(n, n)
    ra = noisy(np.ones((n, n)), seed, noise)
    rb = noisy(np.ones((n, n)), seed + 1, noise)
    return dict(name="spinodal_%d" % n, mask=m, tau=[tau, tau], g_fluid=[G, G], g_solid=[0.0, 0.0],
                f=[f_from_rho(ra, m), f_from_rho(rb, m)], force=[None, None])


def poiseuille_profile(ny, g, tau, rho=1.0):
    """Steady channel profile u_x(y) between link bounce-back walls at rows 0
    and ny-1 (walls half-way between nodes): h = (ny-2)/2 fluid half-width,
    u(y) = g/(2 rho nu) (h^2 - (y - yc)^2), nu = (tau - 1/2)/3
    (scripts/benchmark_utils.py:8-52)."""
    nu = (tau - 0.5) / 3.0
    h = (ny - 2) / 2.0
    yc = (ny - 1) / 2.0
    y = np.arange(ny, dtype=np.float64)
    u = g / (2.0 * rho * nu) * (h * h - (y - yc) ** 2)
    u[0] = u[-1] = 0.0
    return u


# --------------------------------------------------------------------------
# Served jobs (the serving layer generates these on the device; the same
# expressions evaluated here give the inputs of the checks)
# --------------------------------------------------------------------------

def served_density(nx, ny, ic=0, rho_hi=1.0, rho_lo=1.0, radius=0.0, noise=0.0, seed=0):
    """Initial densities (rho_A, rho_B) of a generated served job: ic 0
    uniform, 1 a disk of `radius` about (nx//2, ny//2), 2 rows y < ny//2;
    A takes rho_hi in the region and rho_lo elsewhere, B the mirror; then
    rho * (1 + U[-noise, noise)) with seed (A) and seed + 1 (B)."""
    y, x = np.mgrid[0:ny, 0:nx]
    if ic == 1:
        dx, dy = x - nx // 2, y - ny // 2
        a_hi = (dx * dx + dy * dy).astype(np.float64) <= radius * radius
    elif ic == 2:
        a_hi = y < ny // 2
    else:
        a_hi = np.ones((ny, nx), dtype=bool)
    ra = np.where(a_hi, rho_hi, rho_lo).astype(np.float64)
    rb = np.where(a_hi, rho_lo, rho_hi).astype(np.float64)
    if noise != 0.0:
        ra = noisy(ra, seed, noise)
        rb = noisy(rb, seed + 1, noise)
    return ra, rb


def served_case(nx, ny, porosity=1.0, rmin=8, rmax=20, walls=False, geom_seed=0, ic=0, rho_hi=1.0, rho_lo=1.0,
                radius=0.0, noise=0.0, ic_seed=0, tau=(1.0, 1.0), g_fluid=(0.0, 0.0), g_solid=(0.0, 0.0),
                force=(None, None), **_):
    """Mask and initial populations of a generated served job, in the case
    dict form of the standard cases above."""
    if porosity < 1.0:
        m, _d = porous_mask(nx, ny, porosity, geom_seed, rmin, rmax, walls=walls)
    else:
        m = walls_x(nx, ny) if walls else open_mask(nx, ny)
    ra, rb = served_density(nx, ny, ic, rho_hi, rho_lo, radius, noise, ic_seed)
    return dict(name="served_%dx%d" % (nx, ny), mask=m, tau=list(tau), g_fluid=list(g_fluid),
                g_solid=list(g_solid), f=[f_from_rho(ra, m), f_from_rho(rb, m)], force=list(force))


SERVED_SIDES = (128, 256, 512, 1024, 2048)


def served_stream(n_jobs=2000, seed=2025, steps=(2000, 50000), step_scale=1.0):
    """The served job mix: lattice sides drawn independently from
    {128, 256, 512, 1024, 2048}; seeded disk obstacles (radii 8..20, no
    walls) to a porosity uniform in [0.6, 1.0]; G uniform in [0.8, 1.1] and
    tau uniform in [0.7, 1.3] (both components); a droplet (radius a quarter
    of the shorter side) or stratified initial state, densities 1.0 / 0.06
    with 1% noise; a step count uniform in `steps`, times step_scale.
    Returns keyword dicts for serve.make_job (without release times)."""
    st = Stream(seed)
    jobs = []
    for k in range(n_jobs):
        nx = SERVED_SIDES[min(4, int(st.next() * 5))]
        ny = SERVED_SIDES[min(4, int(st.next() * 5))]
        porosity = 0.6 + 0.4 * st.next()
        G = 0.8 + 0.3 * st.next()
        tau = 0.7 + 0.6 * st.next()
        ic = 1 if st.next() < 0.5 else 2
        n = steps[0] + int(st.next() * (steps[1] - steps[0] + 1))
        jobs.append(dict(nx=nx, ny=ny, steps=max(1, int(round(n * step_scale))), porosity=porosity, rmin=8, rmax=20,
                         geom_seed=seed * 1000003 + 2 * k, ic=ic, rho_hi=1.0, rho_lo=0.06,
                         radius=min(nx, ny) / 4.0, noise=0.01, ic_seed=seed * 1000003 + 2 * k + 1,
                         tau=(tau, tau), g_fluid=(G, G)))
    return jobs


def poisson_arrivals(n, rate_per_s, seed=7):
    """Arrival times in ms of a Poisson process of the given rate."""
    u = uniform01(seed, n)
    gaps = -np.log1p(-u) / rate_per_s * 1e3
    return np.cumsum(gaps)
