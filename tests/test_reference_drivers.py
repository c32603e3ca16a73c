"""The reference's own programs compile unchanged against the drop-in: every
benchmark driver (benchmarks/**/*.cpp) and every unit-test program
(tests/fluid, tests/geometry, tests/lbm, tests/rng) of the reference, compiled
where it lies under the reference checkout with the drop-in headers
(include/) taking the place of the reference's, and linked to libsclbm.so
(oracle/dropin_refs.py).  The RNG test, which needs no GPU, also runs.  The
Fluid test programs run on the GPU in tests/test_gpu_dropin.py.  Skipped
where the reference checkout is absent."""
import os
import subprocess

import pytest

from oracle import dropin_refs

pytestmark = pytest.mark.skipif(not dropin_refs.available(), reason="reference checkout not present")

# the drivers a reader of the reference would run first
NAMED = ["bench_laminar_flow_x_dir", "bench_phase_separation", "bench_laplace_law", "bench_contact_angles",
         "bench_droplet_in_a_channel", "bench_x_flow_array_ellipses", "fluid_test_main", "two_fluids_test_main"]


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    out = tmp_path_factory.mktemp("dropin")
    return out, dropin_refs.build(out_dir=str(out))


def test_every_reference_program_compiles_against_the_dropin(built):
    _, res = built
    failed = {k: v[1][-1500:] for k, v in res.items() if v[0] == "fail"}
    assert not failed, failed
    for k in NAMED:
        assert res[k][0] == "ok", k
    # only programs that do not compile against the reference itself are skipped
    skipped = sorted(k for k, v in res.items() if v[0] == "skip")
    assert skipped == ["bench_droplet_test"], skipped
    assert len([k for k in res if k.startswith("bench_")]) == 13


def test_reference_rng_test_passes(built, tmp_path):
    out, _ = built
    p = subprocess.run([os.path.join(str(out), "rng_tests")], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120, cwd=str(tmp_path))
    assert p.returncode == 0
    assert "test passed" in p.stdout and "failed" not in p.stdout, p.stdout
