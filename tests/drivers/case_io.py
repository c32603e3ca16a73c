"""Case files for tests/drivers/sc_driver (binary, little-endian) and the
reader of its output."""
import os
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "bin", "sc_driver")
MAGIC = 0x424C4353

OUT_1 = ["f", "rho", "ux", "uy", "feq"]
OUT_2 = OUT_1 + ["ueqx", "ueqy", "frx", "fry", "fsx", "fsy"]


def write_case(path, case, steps, mode=0):
    mask = np.ascontiguousarray(case["mask"], dtype=np.int32)
    ny, nx = mask.shape
    nc = len(case["tau"])
    forces = list(case.get("force", [None, None])) + [None, None]
    hdr = [MAGIC, nx, ny, nc, steps, mode, int(forces[0] is not None), int(forces[1] is not None)]

    def two(v):
        v = list(v) + [0.0] * (2 - len(v))
        return np.array(v[:2], dtype=np.float64)

    with open(path, "wb") as fh:
        fh.write(struct.pack("<8i", *hdr))
        fh.write(two(case["tau"]).tobytes())
        fh.write(two(case.get("g_fluid", [0.0, 0.0])).tobytes())
        fh.write(two(case.get("g_solid", [0.0, 0.0])).tobytes())
        for k in range(2):
            fv = forces[k] if forces[k] is not None else np.zeros(9)
            fh.write(np.ascontiguousarray(fv, dtype=np.float64).tobytes())
        fh.write(mask.tobytes())
        f0 = case["f0"] if "f0" in case else case["f"]
        for c in range(nc):
            fh.write(np.ascontiguousarray(f0[c], dtype=np.float64).tobytes())


def read_output(path, ncomp):
    names = OUT_1 if ncomp == 1 else OUT_2
    out = {}
    with open(path, "rb") as fh:
        for c in range(ncomp):
            for nm in names:
                (n,) = struct.unpack("<Q", fh.read(8))
                out[(nm, c)] = np.frombuffer(fh.read(8 * n), dtype=np.float64).copy()
    return out


def run(case, steps, mode, workdir):
    cpath = os.path.join(workdir, "case.bin")
    opath = os.path.join(workdir, "out.bin")
    write_case(cpath, case, steps, mode)
    p = subprocess.run([BIN, cpath, opath], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    if mode == 3:
        return p.returncode, p.stdout
    if p.returncode != 0:
        raise RuntimeError("sc_driver failed (%d): %s" % (p.returncode, p.stdout[-2000:]))
    return read_output(opath, len(case["tau"]))


def build():
    """Compile the driver against include/ and libsclbm.so (rpath to the
    in-tree library)."""
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    lib = os.path.join(ROOT, "sc_lbm_amd")
    cmd = ["g++", "-std=c++14", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "sc_driver.cpp"), "-o", BIN, "-L" + lib, "-lsclbm",
           "-Wl,-rpath,$ORIGIN/../../../sc_lbm_amd"]
    subprocess.check_call(cmd)
    build_dropin_host()
    build_dropin_bench()
    build_serve_cpu()
    return BIN


DROPIN_BENCH = os.path.join(HERE, "bin", "dropin_bench")


def build_dropin_bench():
    """The reference's phase-separation loop on the drop-in classes, timed."""
    os.makedirs(os.path.dirname(DROPIN_BENCH), exist_ok=True)
    lib = os.path.join(ROOT, "sc_lbm_amd")
    cmd = ["g++", "-std=c++14", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "dropin_bench.cpp"), "-o", DROPIN_BENCH, "-L" + lib, "-lsclbm",
           "-Wl,-rpath,$ORIGIN/../../../sc_lbm_amd"]
    subprocess.check_call(cmd)
    return DROPIN_BENCH


def run_dropin_bench(workdir, n, steps, precision, read_every, out=True):
    """Runs tests/drivers/bin/dropin_bench; returns (its JSON line as a
    dict, [f_A, f_B] or None)."""
    import json
    path = os.path.join(workdir, "dropin_bench.bin")
    args = [DROPIN_BENCH, str(n), str(steps), str(precision), str(read_every)] + ([path] if out else [])
    p = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600, cwd=workdir)
    if p.returncode != 0:
        raise RuntimeError("dropin_bench failed (%d): %s" % (p.returncode, p.stdout[-2000:]))
    res = json.loads(p.stdout.strip().splitlines()[-1])
    fs = None
    if out:
        raw = np.fromfile(path, dtype=np.float64)
        fs = [raw[:raw.size // 2], raw[raw.size // 2:]]
    return res, fs


DROPIN_HOST = os.path.join(HERE, "bin", "dropin_host")


def build_dropin_host():
    """The builder / initialiser / text-I/O driver of the drop-in classes."""
    os.makedirs(os.path.dirname(DROPIN_HOST), exist_ok=True)
    lib = os.path.join(ROOT, "sc_lbm_amd")
    cmd = ["g++", "-std=c++11", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "dropin_host.cpp"), "-o", DROPIN_HOST, "-L" + lib, "-lsclbm",
           "-Wl,-rpath,$ORIGIN/../../../sc_lbm_amd"]
    subprocess.check_call(cmd)
    return DROPIN_HOST


def run_dropin_host(workdir, *cmds):
    """Runs tests/drivers/bin/dropin_host with the given command tokens;
    returns (status, mask, [f_A, f_B] or [], stdout)."""
    out = os.path.join(workdir, "dropin_host.bin")
    args = [DROPIN_HOST, out] + [repr(float(c)) if isinstance(c, float) else str(c) for c in cmds]
    p = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600,
                       cwd=workdir)
    if p.returncode != 0:
        raise RuntimeError("dropin_host failed (%d): %s" % (p.returncode, p.stdout[-2000:]))
    with open(out, "rb") as fh:
        status, nx, ny = struct.unpack("<3i", fh.read(12))
        mask = np.frombuffer(fh.read(4 * nx * ny), dtype=np.int32).reshape(ny, nx).copy()
        (nf,) = struct.unpack("<i", fh.read(4))
        fs = [np.frombuffer(fh.read(8 * 9 * nx * ny), dtype=np.float64).copy() for _ in range(nf)]
    return status, mask, fs, p.stdout


SERVE_CPU = os.path.join(HERE, "bin", "libserve_cpu.so")


def build_serve_cpu():
    """The serving scheduler with the oracle as its backend (CPU tests)."""
    os.makedirs(os.path.dirname(SERVE_CPU), exist_ok=True)
    cmd = ["g++", "-std=c++14", "-O2", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-Wall",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "sc_lbm_amd", "host"),
           os.path.join(HERE, "serve_cpu.cpp"), os.path.join(ROOT, "sc_lbm_amd", "host", "serve_core.cpp"),
           os.path.join(ROOT, "oracle", "sc_oracle.cpp"), "-o", SERVE_CPU]
    subprocess.check_call(cmd)
    return SERVE_CPU
