"""Builds libsclbm.so (hipcc, gfx950) in-tree.

Flags: -ffp-contract=off keeps every FP64 operation unfused so the plain
form reproduces the reference bit for bit; packed FP32 instructions are
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout)
        raise RuntimeError("build step failed: %s" % " ".join(cmd))
    return p.stdout


def build_library():
    out = os.path.join(HERE, "libsclbm.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, "sclbm.hip")]
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + srcs)
    os.replace(tmp, out)
    return out


def build_oracle():
    od = os.path.join(ROOT, "oracle")
    _run(["make", "-C", od, "liboracle.so"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-C", od, "ref"])


if __name__ == "__main__":
    print(build_library())
    build_oracle()
