"""Serial CPU time of the two-component Shan-Chen step (the CPU baseline that
bench.py reports beside the GPU figures).

ORACLE / TEST INFRASTRUCTURE: bench.py's cpu_baseline leg runs this, in a
child process pinned to one core (the reference is serial), on the same
workload as the GPU headline: the 4096^2 periodic spinodal lattice (rho_A =
rho_B = 1 +- 1% seeded noise, tau 1, G 1.2).  kind "oracle" times the
clean-room restatement (oracle/sc_oracle.cpp, g++ -O3 -ffp-contract=off);
kind "ref" times the reference's own sources compiled by oracle/Makefile
(oracle/_ref), where they were built.  One untimed step, then `steps` timed
steps, each with a wall clock; MLUPS = sites x steps / seconds / 1e6.

    python oracle/cpu_baseline.py [--n 4096] [--steps 3] [--kind oracle|ref] [--core K]
prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def measure(n, steps, kind, core=None):
    if core is not None:
        os.sched_setaffinity(0, {core})
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as po
    from sc_lbm_amd import workloads as wl
    t0 = time.perf_counter()
    case = wl.spinodal(n, G=1.2, seed=777)
    s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"], kind=kind)
    for c in range(2):
        s.set_f(c, case["f"][c])
    del case
    s.solid_surface_force()
    s.run(1)
    setup = time.perf_counter() - t0
    per = []
    for _ in range(steps):
        t = time.perf_counter()
        s.run(1)
        per.append(time.perf_counter() - t)
    s.close()
    total = sum(per)
    return dict(n=n, steps=steps, kind=kind, seconds=total, per_step_s=per, setup_s=setup,
                mlups=n * n * steps / total / 1e6,
                core=sorted(os.sched_getaffinity(0)))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_child(n=4096, steps=3, kind="oracle", core=None, timeout=900):
    """Runs measure() in a child process (so the pinning applies to it only)
    and returns its result."""
    import subprocess
    if core is None:
        core = sorted(os.sched_getaffinity(0))[0]
    cmd = [sys.executable, os.path.abspath(__file__), "--n", str(n), "--steps", str(steps), "--kind", kind,
           "--core", str(core)]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError("cpu baseline failed: %s" % p.stderr[-2000:])
    return json.loads(p.stdout.strip().splitlines()[-1])


if __name__ == "__main__":
    a = argparse.ArgumentParser()
    a.add_argument("--n", type=int, default=4096)
    a.add_argument("--steps", type=int, default=3)
    a.add_argument("--kind", default="oracle", choices=["oracle", "ref"])
    a.add_argument("--core", type=int, default=None)
    args = a.parse_args()
    print(json.dumps(measure(args.n, args.steps, args.kind, args.core)))
