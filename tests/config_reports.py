"""Per-configuration reports on the GPU (writes a markdown summary).

    python tests/config_reports.py [--out FILE] [--only c1,c2,c3,c5]

(default output: profiles/configs_report.md)

For each standard configuration this runs the library on the GPU and states
the quantities that characterise the configuration:
  C1  Poiseuille 200x50, FP64, 10000 steps: full population field against
      the oracle (bit for bit), u_x(y) against the analytic profile;
  C2  static droplet 256^2, FP64 and FP32 deviations, 20000 steps: FP64
      against the oracle after 2000 steps (bit for bit), interface radius
      and Laplace pressure jump for every form, FP32 / FP16 against FP64;
      the specified coupling (G 0.9 at total density 1.06) lies below the
      two-component critical coupling, so the same droplet is also run at
      G 2.2, where it persists (Laplace law);
  C3  porous 1024x512, FP64 and FP32 deviations, 10000 steps: the oracle
      after 200 steps (bit for bit), FP32 against FP64 (density, velocity);
  C5  spinodal 4096^2, FP64, FP32 and FP16 deviations, 10000 steps: max
      density error of each form against FP64 and the domain-size growth
      curve (first moment of the structure factor) every 500 steps, beside
      FP64 started from populations perturbed by 1e-12 (relative): the
      spread that phase separation alone gives any two runs.
The oracle is the checker here, never the thing measured.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from sc_lbm_amd import native as nat  # noqa: E402
from sc_lbm_amd import workloads as wl  # noqa: E402

FORMS = {"fp64": nat.FP64, "fp32dev": nat.FP32_DEV, "fp16dev": nat.FP16_DEV}
# SHA-256 of the compiled reference's final populations for the long cases
# (tests/golden/make_config_digests.py)
DIGESTS = os.path.join(ROOT, "tests", "golden", "config_digests.json")


def ref_digest_line(name, f, steps):
    """Whether the FP64 populations equal the compiled reference's after the
    case's full step count (by SHA-256 of their float64 bytes)."""
    if not os.path.exists(DIGESTS) or name not in json.load(open(DIGESTS)):
        return "- (no reference digest for %s)" % name
    want = json.load(open(DIGESTS))[name]
    got = [hashlib.sha256(np.ascontiguousarray(x, dtype="<f8").tobytes()).hexdigest() for x in f]
    return ("- FP64 populations after %d steps equal to the compiled reference's bit for bit (SHA-256 of both "
            "components): **%s**" % (steps, got == want["f_sha256"] and want["steps"] == steps))


def log(msg):
    print(msg, flush=True)


class Run:
    """A case on the GPU in one storage form (two components or one)."""

    def __init__(self, case, form, rho0=1.0):
        self.case = case
        self.g = nat.Geometry(case["mask"])
        self.nc = len(case["tau"])
        prec = FORMS[form]
        self.prec = prec
        self.fl = []
        for c in range(self.nc):
            f = nat.Field(self.g, case["tau"][c], prec, rho0, g_solid=case["g_solid"][c] if self.nc == 2 else 0.0,
                          g_fluid=case["g_fluid"][c] if self.nc == 2 else 0.0)
            f.set_f(case["f"][c])
            self.fl.append(f)
        if self.nc == 2:
            nat.solid_surface_force(self.g, self.fl[0], self.fl[1])
        self.force = case.get("force", [None, None])

    def step(self, n):
        if self.nc == 1:
            nat.step1(self.g, self.fl[0], self.force[0], n)
        else:
            nat.step2(self.g, self.fl[0], self.fl[1], self.force[0], self.force[1], n)

    def f(self, c):
        return self.fl[c].get_f()

    def rho(self, c):
        if self.prec == nat.FP64:      # on the device (per-operator calls take FP64 fields)
            nat.compute_density(self.fl[c])
            return self.fl[c].download(nat.RHO).reshape(self.case["mask"].shape)
        return self.f(c).reshape(9, -1).sum(axis=0).reshape(self.case["mask"].shape)

    def close(self):
        for f in self.fl:
            f.close()
        self.g.close()


def perturbed(case, rel, seed=99):
    """The case with every population multiplied by 1 + rel * U[-1, 1)."""
    rng = np.random.default_rng(seed)
    out = dict(case)
    out["f"] = [f * (1.0 + rel * (2.0 * rng.random(f.size) - 1.0)) for f in case["f"]]
    return out


PERTS = (1e-12, 6e-8)   # 6e-8: unit roundoff of FP32


def oracle(case, steps):
    s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"])
    for c in range(len(case["tau"])):
        s.set_f(c, case["f"][c])
    if len(case["tau"]) == 2:
        s.solid_surface_force()
    forces = case.get("force", [None, None])
    s.run(steps, forces[0], forces[1] if len(forces) > 1 else None)
    return [s.get_f(c) for c in range(len(case["tau"]))]


def velocity(f, mask):
    p = f.reshape(9, -1)
    rho = p.sum(axis=0)
    jx = p[1] - p[3] + p[5] - p[6] - p[7] + p[8]
    jy = p[2] - p[4] + p[5] + p[6] - p[7] - p[8]
    fl = mask.ravel() == 1
    ux = np.where(fl, jx / np.where(fl, rho, 1.0), 0.0)
    uy = np.where(fl, jy / np.where(fl, rho, 1.0), 0.0)
    return ux.reshape(mask.shape), uy.reshape(mask.shape)


def c1(rep):
    case = wl.poiseuille(200, 50, tau=1.0, rho0=1.0, g=1e-6)
    t = time.time()
    r = Run(case, "fp64")
    r.step(10000)
    f = r.f(0)
    r.close()
    ref = oracle(case, 10000)[0]
    ux, _ = velocity(f, case["mask"])
    prof = ux.mean(axis=1)
    ana = wl.poiseuille_profile(50, 1e-6, 1.0)
    inner = slice(1, 49)
    dev = np.max(np.abs(prof[inner] - ana[inner])) / np.max(ana)
    rep.append("## C1 Poiseuille 200x50, FP64, 10000 steps\n")
    rep.append("- populations equal to the oracle's bit for bit: **%s** (max |diff| %.3g)" %
               (np.array_equal(f, ref), float(np.max(np.abs(f - ref)))))
    rep.append("- u_x(y) vs analytic g/(2 rho nu)(h^2 - (y - y_c)^2): max |u - u_ana| / u_max = %.3g "
               "(u_max %.6g analytic, %.6g measured; 10^4 steps are 0.72 of the diffusion time H^2/nu, so "
               "the slowest mode has not fully decayed)" % (dev, float(ana.max()), float(prof.max())))
    rep.append("- wall time %.1f s (GPU run + oracle)\n" % (time.time() - t))


def droplet_stats(ra, rb, G):
    inside = ra > rb
    R = np.sqrt(inside.sum() / np.pi)
    p = (ra + rb) / 3.0 + G * ra * rb / 3.0
    n = ra.shape[0]
    c = n // 2
    y, x = np.mgrid[0:n, 0:n]
    d = np.sqrt((x - c) ** 2 + (y - c) ** 2)
    p_in = p[d < 0.5 * R].mean()
    p_out = p[d > min(R + 0.35 * (n / 2 - R) + 10, 0.48 * n)].mean()
    return R, p_in - p_out


def c2(rep):
    case = wl.droplet(n=256, radius=40, rho_in=1.0, rho_out=0.06, G=0.9, tau=1.0, seed=12345, noise=0.01)
    rep.append("## C2 static droplet 256^2, G 0.9, tau 1, 20000 steps\n")
    rep.append("At G 0.9 and total density rho_A + rho_B = 1.06 the coupling is below the critical value of the "
               "two-component model (G rho_tot about 2), so the components mix and no interface remains "
               "(radius 0); the oracle's populations match the FP64 run bit for bit, so the reference behaves "
               "alike.  The droplet at G 2.2 below persists.\n")
    r = Run(case, "fp64")
    r.step(2000)
    f2000 = [r.f(c) for c in range(2)]
    ref = oracle(case, 2000)
    same = all(np.array_equal(f2000[c], ref[c]) for c in range(2))
    rep.append("- FP64 populations after 2000 steps equal to the oracle's bit for bit: **%s**" % same)
    r.step(18000)
    rep.append(ref_digest_line("c2", [r.f(c) for c in range(2)], 20000))
    rho64 = [r.rho(c) for c in range(2)]
    r.close()
    R64, dp64 = droplet_stats(rho64[0], rho64[1], 0.9)
    rep.append("- FP64: radius %.4f, Laplace pressure jump %.6g (sigma = R * dp = %.6g)" % (R64, dp64, R64 * dp64))
    for form, rho0 in (("fp32dev", 0.5), ("fp16dev", 0.5)):
        r = Run(case, form, rho0=rho0)
        r.step(20000)
        rho = [r.rho(c) for c in range(2)]
        r.close()
        R, dp = droplet_stats(rho[0], rho[1], 0.9)
        err = max(float(np.max(np.abs(rho[c] - rho64[c]))) for c in range(2))
        rep.append("- %s (rho0 %.1f): radius %.4f, dp %.6g, max |rho - rho_FP64| %.3g" % (form, rho0, R, dp, err))
    rep.append("")
    rep.append("### Same droplet at G 2.2 (above the critical coupling), 20000 steps\n")
    hot = wl.droplet(n=256, radius=40, rho_in=1.0, rho_out=0.06, G=2.2, tau=1.0, seed=12345, noise=0.01)
    r = Run(hot, "fp64")
    r.step(2000)
    f2000 = [r.f(c) for c in range(2)]
    ref = oracle(hot, 2000)
    same = all(np.array_equal(f2000[c], ref[c]) for c in range(2))
    rep.append("- FP64 populations after 2000 steps equal to the oracle's bit for bit: **%s**" % same)
    r.step(18000)
    f64 = [r.f(c) for c in range(2)]
    rep.append(ref_digest_line("c2_hot", f64, 20000))
    u64 = [velocity(f64[c], hot["mask"]) for c in range(2)]
    rho64 = [r.rho(c) for c in range(2)]
    r.close()
    R64, dp64 = droplet_stats(rho64[0], rho64[1], 2.2)
    rep.append("- FP64: radius %.4f, Laplace pressure jump %.6g (sigma = R * dp = %.6g)" % (R64, dp64, R64 * dp64))
    for rel in PERTS:
        r = Run(perturbed(hot, rel), "fp64")
        r.step(20000)
        rho = [r.rho(c) for c in range(2)]
        r.close()
        R, dp = droplet_stats(rho[0], rho[1], 2.2)
        err = max(float(np.max(np.abs(rho[c] - rho64[c]))) for c in range(2))
        rep.append("- FP64 from populations perturbed by %.0e (noise floor): radius %.4f, dp %.6g, "
                   "max |rho - rho_FP64| %.3g" % (rel, R, dp, err))
    for form, rho0 in (("fp32dev", 0.5), ("fp16dev", 0.5)):
        r = Run(hot, form, rho0=rho0)
        r.step(20000)
        fx = [r.f(c) for c in range(2)]
        rho = [r.rho(c) for c in range(2)]
        r.close()
        R, dp = droplet_stats(rho[0], rho[1], 2.2)
        err = max(float(np.max(np.abs(rho[c] - rho64[c]))) for c in range(2))
        du = max(float(np.max(np.abs(velocity(fx[c], hot["mask"])[k] - u64[c][k]))) for c in range(2) for k in range(2))
        rep.append("- %s (rho0 %.1f): radius %.4f, dp %.6g, max |rho - rho_FP64| %.3g, max |u - u_FP64| %.3g"
                   % (form, rho0, R, dp, err, du))
    rep.append("")


def c3(rep):
    case = wl.porous(nx=1024, ny=512, porosity=0.70, seed=2024, G=0.9, gads=0.1, tau=1.0, g=1e-5)
    por = float(case["mask"].mean())
    rep.append("## C3 porous 1024x512 (porosity %.4f incl. wall rows, %d disks), 10000 steps\n"
               % (por, len(case["disks"])))
    r = Run(case, "fp64")
    r.step(200)
    f200 = [r.f(c) for c in range(2)]
    r.close()
    ref = oracle(case, 200)
    same = all(np.array_equal(f200[c], ref[c]) for c in range(2))
    rep.append("- FP64 populations after 200 steps equal to the oracle's bit for bit: **%s**" % same)
    out = {}
    # FP32 deviations about rho0 = 1 and about the midpoint of the two phases'
    # densities (the storage spacing scales with |f - w rho0|)
    runs = [("fp64", case, "fp64", 1.0), ("fp32dev (rho0 1.0)", case, "fp32dev", 1.0),
            ("fp32dev (rho0 0.53)", case, "fp32dev", 0.53)] + \
           [("fp64 perturbed %.0e" % rel, perturbed(case, rel), "fp64", 1.0) for rel in PERTS]
    n_fluid = int(case["mask"].sum())
    for name, cs, form, rho0 in runs:
        r = Run(cs, form, rho0=rho0)
        r.step(2)   # bootstrap and shape selection outside the timing
        nat.synchronize()
        t = time.perf_counter()
        r.step(9998)
        nat.synchronize()
        dt = time.perf_counter() - t
        out[name] = [r.f(c) for c in range(2)]
        r.close()
        if "perturbed" not in name:
            rep.append("- %s throughput: %.0f MLUPS (all sites), %.0f MLUPS on fluid sites"
                       % (name, case["mask"].size * 9998 / dt / 1e6, n_fluid * 9998 / dt / 1e6))
    rep.append(ref_digest_line("c3", out["fp64"], 10000))
    m = case["mask"]
    for name, _, _, _ in runs[1:]:
        drho = max(float(np.max(np.abs(out["fp64"][c].reshape(9, -1).sum(0) - out[name][c].reshape(9, -1).sum(0))))
                   for c in range(2))
        du = 0.0
        for c in range(2):
            a = velocity(out["fp64"][c], m)
            b = velocity(out[name][c], m)
            du = max(du, float(np.max(np.abs(a[0] - b[0]))), float(np.max(np.abs(a[1] - b[1]))))
        rep.append("- %s vs FP64 after 10000 steps: max |d rho| %.3g, max |d u| %.3g" % (name, drho, du))
    rep.append("")


def domain_length(phi):
    """2 pi / k1 with k1 the first moment of the radially averaged structure
    factor of phi - <phi>."""
    n = phi.shape[0]
    s = np.abs(np.fft.rfft2(phi - phi.mean())) ** 2     # half plane (phi is real)
    ky = np.fft.fftfreq(n) * 2 * np.pi
    kx = np.fft.rfftfreq(n) * 2 * np.pi
    kk = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    w = np.full(kx.size, 2.0)                            # columns standing for two of the full plane
    w[0] = 1.0
    if n % 2 == 0:
        w[-1] = 1.0
    s[0, 0] = 0.0
    return 2 * np.pi * (w * s).sum() / (w * kk * s).sum()


def c5(rep, n=4096, steps=10000, every=500):
    case = wl.spinodal(n, G=1.2, tau=1.0, seed=777, noise=0.01)
    rep.append("## C5 spinodal %d^2, G 1.2, tau 1, %d steps\n" % (n, steps))
    names = ("fp64", "fp64 perturbed 1e-12", "fp64 perturbed 6e-08", "fp32dev", "fp16dev")
    runs = {"fp64": Run(case, "fp64"), "fp64 perturbed 1e-12": Run(perturbed(case, 1e-12), "fp64"),
            "fp64 perturbed 6e-08": Run(perturbed(case, 6e-8), "fp64"),
            "fp32dev": Run(case, "fp32dev"), "fp16dev": Run(case, "fp16dev")}
    curves = {k: [] for k in names}
    errs = {k: [] for k in names[1:]}
    t = time.time()
    for k in range(steps // every):
        rho = {}
        for name in names:
            runs[name].step(every)
            rho[name] = [runs[name].rho(c) for c in range(2)]
            curves[name].append(domain_length(rho[name][0] - rho[name][1]))
        for name in names[1:]:
            errs[name].append(max(float(np.max(np.abs(rho[name][c] - rho["fp64"][c]))) for c in range(2)))
        log("  C5 step %d: L %s; max |d rho| vs FP64 %s (%.0f s)"
            % ((k + 1) * every, " ".join("%.4f" % curves[x][-1] for x in names),
               " ".join("%.2e" % errs[x][-1] for x in names[1:]), time.time() - t))
    ufin = {name: [velocity(runs[name].f(c), case["mask"]) for c in range(2)] for name in names}
    du = {name: max(float(np.max(np.abs(ufin[name][c][k] - ufin["fp64"][c][k]))) for c in range(2) for k in range(2))
          for name in names[1:]}
    for r in runs.values():
        r.close()
    rep.append("Max |u - u_FP64| over both components after %d steps: " % steps +
               "; ".join("%s %.3g" % (name, du[name]) for name in names[1:]) + ".\n")
    rep.append("Max |rho - rho_FP64| over both components after %d steps: FP64 from populations perturbed by "
               "1e-12: %.3g; by 6e-8 (the unit roundoff of FP32): %.3g; FP32 deviations: %.3g; FP16 deviations: "
               "%.3g.  Spinodal decomposition amplifies differences in the initial state by about 1e8 in the "
               "first 500 steps (the linear regime); an FP64 run started within FP32's roundoff of the same "
               "state therefore ends as far from FP64 as the FP32 run does, and pointwise agreement after 10^4 "
               "steps exists only between bit-identical runs.  The table gives the error over time and the "
               "domain size L, the statistical observable.\n"
               % (steps, errs[names[1]][-1], errs[names[2]][-1], errs["fp32dev"][-1], errs["fp16dev"][-1]))
    rep.append("| step | L FP64 | L FP64 +1e-12 | L FP64 +6e-8 | L FP32 dev | L FP16 dev | err +1e-12 | err +6e-8 "
               "| err FP32 dev | err FP16 dev |")
    rep.append("|---|---|---|---|---|---|---|---|---|---|")
    for k in range(steps // every):
        rep.append("| %d | %.4f | %.4f | %.4f | %.4f | %.4f | %.2e | %.2e | %.2e | %.2e |"
                   % ((k + 1) * every, curves["fp64"][k], curves[names[1]][k], curves[names[2]][k],
                      curves["fp32dev"][k], curves["fp16dev"][k], errs[names[1]][k], errs[names[2]][k],
                      errs["fp32dev"][k], errs["fp16dev"][k]))
    rep.append("")


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--out", default=os.path.join(ROOT, "profiles", "configs_report.md"))
    p.add_argument("--only", default="c1,c2,c3,c5")
    a = p.parse_args()
    nat.init(0)
    rep = ["# Configuration reports (GPU: libsclbm.so on MI355X; checker: the oracle)", "",
           "Command: `python tests/config_reports.py --only %s`" % a.only, ""]
    for name in a.only.split(","):
        t = time.time()
        log("running %s" % name)
        {"c1": c1, "c2": c2, "c3": c3, "c5": c5}[name](rep)
        log("  %s done in %.0f s" % (name, time.time() - t))
        os.makedirs(os.path.dirname(a.out), exist_ok=True)
        open(a.out, "w").write("\n".join(rep) + "\n")
    print("\n".join(rep))


if __name__ == "__main__":
    main()
