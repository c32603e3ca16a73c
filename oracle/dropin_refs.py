"""Compiles the reference's own drivers and unit tests, unchanged and where
they lie under the reference checkout, against this repository's drop-in
headers (include/) and libsclbm.so.

TEST INFRASTRUCTURE: only tests/ and __graft_entry__.build() use it.  The
reference sources are read in place and never copied; what comes out are
executables under oracle/_ref/dropin/ (git-ignored), linked to the in-tree
libsclbm.so.

The reference's drivers reach its headers by relative paths
("../../include/lbm.h", "../../../include/lbm.h", "../include/lbm.h",
"../common/test_utils.h").  The compiler's -I- switch stops a quoted include
from being looked up beside the file that includes it, so those paths resolve
against the -I directories instead: a temporary shadow tree whose include/ is
this repository's include/ and whose common/ is the reference's test
utilities (tests/common).  Each driver is first compiled against the
reference's own headers; one that does not compile there either
(benchmarks/two_phase_two_components_flows/droplet_test.cpp names a
../common/test_utils.h that does not exist beside it) is reported as such.
"""
import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = os.environ.get("SCLBM_REFERENCE", "/root/reference")
OUT = os.path.join(HERE, "_ref", "dropin")
LIBDIR = os.path.join(ROOT, "sc_lbm_amd")

# unit-test programs of the reference: (name, main source, helper sources)
TESTS = [
    ("fluid_test_main", "tests/fluid/fluid_test_main.cpp", []),
    ("two_fluids_test_main", "tests/fluid/two_fluids_test_main.cpp", []),
    ("rng_tests", "tests/rng/rng_tests.cpp", []),
    ("add_objects_test", "tests/geometry/add_objects_test.cpp", []),
    ("make_arrays_test", "tests/geometry/make_arrays_test.cpp", []),
    ("misc_geom_tests", "tests/geometry/misc_geom_tests.cpp", []),
    ("create_geometries", "tests/lbm/create_geometries.cpp", []),
    ("no_force_tests", "tests/lbm/no_force_tests.cpp", ["tests/lbm/lbm_tests.cpp"]),
    ("uni_force_tests", "tests/lbm/uni_force_tests.cpp", ["tests/lbm/lbm_tests.cpp"]),
    ("multi_force_tests", "tests/lbm/multi_force_tests.cpp", ["tests/lbm/lbm_tests.cpp"]),
    ("droplet_tests", "tests/lbm/droplet_tests.cpp", ["tests/lbm/lbm_tests.cpp"]),
    ("restart_tests", "tests/lbm/restart_tests.cpp", ["tests/lbm/lbm_tests.cpp"]),
]


def available():
    return os.path.isdir(os.path.join(REF, "benchmarks")) and os.path.isdir(os.path.join(REF, "tests"))


def benchmark_drivers():
    out = []
    for dp, _, files in os.walk(os.path.join(REF, "benchmarks")):
        out += [os.path.join(dp, f) for f in files if f.endswith(".cpp")]
    return sorted(out)


def _shadow():
    d = tempfile.mkdtemp(prefix="sclbm_shadow_")
    os.makedirs(os.path.join(d, "d1", "d2", "d3"))
    os.symlink(os.path.join(ROOT, "include"), os.path.join(d, "include"))
    os.symlink(os.path.join(REF, "tests", "common"), os.path.join(d, "d1", "common"))
    return d


def _dropin_flags(shadow):
    return ["-I-", "-I" + os.path.join(shadow, "d1", "d2", "d3"), "-I" + os.path.join(shadow, "d1", "d2"),
            "-I" + os.path.join(shadow, "d1"), "-I" + os.path.join(shadow, "include"),
            "-I" + os.path.join(shadow, "include", "io_operations")]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return p.returncode, p.stdout


def compile_against_reference(src, extra_dirs=()):
    """Syntax check of a driver against the reference's own headers."""
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-I" + os.path.join(REF, "src")]
    cmd += ["-I" + d for d in extra_dirs]
    return _run(cmd + [src])


def compile_dropin(srcs, exe, shadow, extra_dirs=()):
    """Compile and link sources against include/ and libsclbm.so."""
    cmd = ["g++", "-std=c++11", "-O1", "-w"] + _dropin_flags(shadow)
    cmd += ["-I" + d for d in extra_dirs]
    rpath = "$ORIGIN/" + os.path.relpath(LIBDIR, os.path.dirname(os.path.abspath(exe)))
    cmd += srcs + ["-o", exe, "-L" + LIBDIR, "-lsclbm", "-Wl,-rpath," + rpath]
    return _run(cmd)


def build(out_dir=OUT, keep_benchmarks=False, jobs=8):
    """Compile every benchmark driver and unit-test program.  Returns
    {name: (status, log)} with status "ok", "fail" or "skip" (does not compile
    against the reference itself).  Unit-test executables go to out_dir;
    benchmark executables too when keep_benchmarks, else to a scratch
    directory (they run for hours and are only compiled)."""
    if not available():
        raise FileNotFoundError(REF)
    os.makedirs(out_dir, exist_ok=True)
    shadow = _shadow()
    scratch = tempfile.mkdtemp(prefix="sclbm_drv_")
    tasks = []
    for src in benchmark_drivers():
        name = "bench_" + os.path.splitext(os.path.basename(src))[0]
        tasks.append((name, [src], [], out_dir if keep_benchmarks else scratch))
    common = os.path.join(REF, "tests", "common")
    for name, main, helpers in TESTS:
        srcs = [os.path.join(REF, main)] + [os.path.join(REF, h) for h in helpers]
        if name != "create_geometries":
            srcs.append(os.path.join(common, "test_utils.cpp"))
        tasks.append((name, srcs, [os.path.dirname(srcs[0]), common], out_dir))

    def one(t):
        name, srcs, dirs, odir = t
        rc, log = compile_against_reference(srcs[0], dirs)
        if rc != 0:
            return name, ("skip", log)
        rc, log = compile_dropin(srcs, os.path.join(odir, name), shadow, dirs)
        return name, ("ok" if rc == 0 else "fail", log)

    try:
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            res = dict(pool.map(one, tasks))
    finally:
        shutil.rmtree(shadow, ignore_errors=True)
        shutil.rmtree(scratch, ignore_errors=True)
    return res


if __name__ == "__main__":
    for k, (st, log) in sorted(build().items()):
        print("%-40s %s" % (k, st))
        if st == "fail":
            print(log[-2000:])
