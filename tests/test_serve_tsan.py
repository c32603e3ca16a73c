"""Race detection for the serving scheduler: concurrent clients
(tests/drivers/serve_stress.cpp) against serve_core.cpp with the CPU backend,
built with ThreadSanitizer (host compiler only)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(HERE, "drivers", "bin", "serve_stress_tsan")


def build():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = ["g++", "-std=c++14", "-O1", "-g", "-pthread", "-fsanitize=thread", "-ffp-contract=off",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "sc_lbm_amd", "host"),
           os.path.join(HERE, "drivers", "serve_stress.cpp"), os.path.join(HERE, "drivers", "serve_cpu.cpp"),
           os.path.join(ROOT, "sc_lbm_amd", "host", "serve_core.cpp"), os.path.join(ROOT, "oracle", "sc_oracle.cpp"),
           "-o", BIN]
    subprocess.check_call(cmd)


@pytest.mark.slow
def test_scheduler_has_no_data_races():
    build()
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1")
    p = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env, timeout=300)
    assert "ThreadSanitizer" not in p.stderr, p.stderr[-4000:]
    assert p.returncode == 0, p.stdout + p.stderr[-2000:]
