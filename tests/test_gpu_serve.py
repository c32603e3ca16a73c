"""GPU tests of the serving layer (scl_server_*): served FP64 jobs are bit
for bit the oracle's run of the same case; every job's result is the same
alone, in a mix of jobs of other sizes admitted while it runs, under
memory back-pressure, with other tile shapes, and on the single-lattice
path; the watchdog isolates a failing job; FP32 deviations stay within
1e-5 of FP64 on a served job."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from sc_lbm_amd import native as nat
from sc_lbm_amd import serve
from sc_lbm_amd import workloads as wl

pytestmark = pytest.mark.gpu

TILE = dict(tile_cols=256, tile_rows=32, lookahead=1)
DROP = dict(ic=serve.IC_DROPLET, rho_hi=1.0, rho_lo=0.06, noise=0.01, tau=(1.0, 0.8), g_fluid=(0.9, 0.9))


@pytest.fixture(scope="module", autouse=True)
def _init():
    nat.init(0)
    yield


def gpu_server(precision="fp32dev", **cfg):
    base = dict(budget_bytes=2 << 30, round_steps=16, precision=nat.PRECISION[precision], **TILE)
    base.update(cfg)
    return serve.Server(**base)


def oracle_run(case, steps):
    s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"])
    for c in range(2):
        s.set_f(c, case["f"][c])
    s.solid_surface_force()
    s.run(steps, case["force"][0], case["force"][1])
    return [s.get_f(c) for c in range(2)]


def dens_ref(f):
    p = f.reshape(9, -1)
    r = np.zeros(p.shape[1])
    for i in range(9):
        r = r + p[i]
    return r


def footprint(nx, ny, es, stage):
    """Device bytes of one resident job (the backend's block layout: mask,
    2 x 2 population buffers, equilibrium-velocity planes (4 in FP64, 4 x 4
    rotating in the deviation forms), 2 density planes in FP64, in FP64 the
    zero-density rule's records (a 32-bit word per site and 2 more per
    component), optional FP64 staging; each region 256-byte aligned)."""
    a = lambda n: (n + 255) // 256 * 256
    pitch = (nx + 63) // 64 * 64
    plane = pitch * ny
    nueq = 1 if es == 8 else 4
    total = a(ny * pitch // 64 * 8) + 4 * a(9 * plane * es) + (4 * nueq + 2) * a(plane * 8)
    if es == 8:
        total += 2 * a((plane + 2) * 4)
    return total + (a(9 * plane * 8) if stage else 0)


def served_rho0(job):
    """The server's automatic deviation reference density of a generated
    initial state: the midpoint of rho_hi and rho_lo, rounded (half away from
    zero) to a multiple of 2^-10."""
    return math.floor(0.5 * (job["rho_hi"] + job["rho_lo"]) * 1024.0 + 0.5) / 1024.0


def run_alone(job, precision="fp32dev", **cfg):
    with gpu_server(precision, **cfg) as srv:
        jid = srv.submit(**job)
        st = srv.wait(jid, 60000)
        assert st["state"] == "done", st
        assert st["bytes"] == footprint(job["nx"], job["ny"], {"fp64": 8, "fp32dev": 4, "fp16dev": 2}[precision],
                                        "f" in job or job.get("deliver", 0) & serve.DELIVER_F)
        return [srv.result(jid, serve.F, c) for c in range(2)]


POROUS = dict(nx=150, ny=96, steps=300, porosity=0.8, rmin=4, rmax=9, walls=True, geom_seed=3, radius=20,
              ic_seed=21, g_solid=(0.1, -0.1), force=(wl.body_force(1e-5), None), **DROP)


def test_fp64_served_job_is_the_oracle_run():
    """Generated mask (host sampler) and initial state (device kernel), then
    300 steps in rounds of 16 with other jobs resident: bit-exact."""
    others = [dict(nx=64 + 32 * k, ny=80, steps=100 + 50 * k, ic_seed=k, radius=10, **DROP) for k in range(3)]
    with gpu_server("fp64") as srv:
        ids = [srv.submit(**o) for o in others]
        jid = srv.submit(deliver=serve.DELIVER_F | serve.DELIVER_RHO, **POROUS)
        st = srv.wait(jid, 60000)
        assert st["state"] == "done" and st["steps_done"] == 300
        got = [srv.result(jid, serve.F, c) for c in range(2)]
        rho = [srv.result(jid, serve.RHO, c) for c in range(2)]
        for i in ids:
            assert srv.wait(i, 60000)["state"] == "done"
    ref = oracle_run(wl.served_case(**POROUS), 300)
    for c in range(2):
        assert np.array_equal(got[c].reshape(-1), ref[c])
        assert np.array_equal(rho[c].reshape(-1), dens_ref(ref[c]))


def test_fp64_job_with_caller_arrays_is_the_oracle_run():
    case = wl.porous(nx=128, ny=72, porosity=0.8, seed=9, rmin=4, rmax=8)
    job = dict(nx=128, ny=72, steps=200, mask=case["mask"], f=case["f"], tau=case["tau"], g_fluid=case["g_fluid"],
               g_solid=case["g_solid"], force=case["force"], deliver=serve.DELIVER_F)
    got = run_alone(job, "fp64", round_steps=32)
    ref = oracle_run(case, 200)
    for c in range(2):
        assert np.array_equal(got[c].reshape(-1), ref[c])


@pytest.mark.parametrize("precision", ["fp32dev", "fp16dev"])
def test_alone_equals_mix_and_single_lattice_path(precision):
    """Continuous batching: the target runs alone, then in a mix whose other
    jobs are of other sizes and arrive while it runs; both results and the
    single-lattice fused path agree bit for bit."""
    target = dict(deliver=serve.DELIVER_F, **POROUS)
    alone = run_alone(target, precision)
    mix = [dict(nx=300, ny=200, steps=500, porosity=0.9, geom_seed=5, ic=serve.IC_STRATIFIED, rho_hi=1.0,
                rho_lo=0.1, noise=0.01, ic_seed=8, tau=(0.9, 1.2), g_fluid=(1.0, 1.0)),
           dict(nx=64, ny=64, steps=40, ic_seed=1, radius=12, **DROP),
           dict(nx=517, ny=33, steps=120, ic_seed=2, radius=10, release_ms=5.0, **DROP),
           dict(nx=128, ny=256, steps=260, porosity=0.75, rmin=5, rmax=12, geom_seed=7, ic_seed=3, radius=30,
                release_ms=20.0, **DROP)]
    with gpu_server(precision) as srv:
        t0 = srv.stats()["clock_ms"]
        ids = [srv.submit(**dict(m, release_ms=t0 + m.get("release_ms", 0.0))) for m in mix[:2]]
        jid = srv.submit(**dict(target, release_ms=t0 + 2.0))
        ids += [srv.submit(**dict(m, release_ms=t0 + m["release_ms"])) for m in mix[2:]]
        st = srv.wait(jid, 60000)
        assert st["state"] == "done"
        mixed = [srv.result(jid, serve.F, c) for c in range(2)]
        for i in ids:
            assert srv.wait(i, 60000)["state"] == "done"
        stats = srv.stats()
    assert stats["max_resident"] >= 3
    for c in range(2):
        assert np.array_equal(alone[c], mixed[c])
    # the single-lattice path from the same initial state
    case = wl.served_case(**POROUS)
    g = nat.Geometry(case["mask"])
    fl = [nat.Field(g, case["tau"][c], nat.PRECISION[precision], served_rho0(POROUS), g_solid=case["g_solid"][c],
                    g_fluid=case["g_fluid"][c]) for c in range(2)]
    for c in range(2):
        fl[c].set_f(case["f"][c])
    nat.solid_surface_force(g, fl[0], fl[1])
    nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], POROUS["steps"])
    for c in range(2):
        assert np.array_equal(fl[c].get_f().reshape(-1), alone[c].reshape(-1))


def test_tile_shape_and_round_length_do_not_change_results():
    job = dict(deliver=serve.DELIVER_F, **dict(POROUS, steps=77))
    a = run_alone(job, "fp32dev", tile_cols=128, tile_rows=16, lookahead=2, round_steps=5)
    b = run_alone(job, "fp32dev", tile_cols=256, tile_rows=64, lookahead=1, round_steps=64)
    for c in range(2):
        assert np.array_equal(a[c], b[c])


def test_memory_back_pressure_keeps_results():
    spec = [dict(nx=256, ny=256, steps=60 + 10 * k, ic_seed=40 + k, radius=50, deliver=serve.DELIVER_F, **DROP)
            for k in range(5)]
    per_job = footprint(256, 256, 4, stage=True)
    with gpu_server("fp32dev", budget_bytes=2 * per_job + 4096) as srv:
        ids = [srv.submit(**s) for s in spec]
        sts = [srv.wait(i, 60000) for i in ids]
        stats = srv.stats()
        mixed = [[srv.result(i, serve.F, c) for c in range(2)] for i in ids]
    assert all(s["state"] == "done" for s in sts)
    assert stats["max_resident"] <= 2 and stats["done"] == 5
    for k in (0, 4):
        alone = run_alone(spec[k])
        for c in range(2):
            assert np.array_equal(alone[c], mixed[k][c])


def test_watchdog_fails_only_the_bad_job():
    case = wl.droplet(n=64, radius=12, G=0.9, seed=4)
    bad = [case["f"][0].copy(), case["f"][1]]
    bad[0][5 * 64 * 64 + 700] = np.inf
    with gpu_server("fp32dev") as srv:
        good = [srv.submit(nx=96, ny=64, steps=50, ic_seed=k, radius=12, **DROP) for k in range(3)]
        jbad = srv.submit(nx=64, ny=64, steps=50, f=bad, tau=case["tau"], g_fluid=case["g_fluid"])
        st = srv.wait(jbad, 60000)
        assert st["state"] == "failed" and "non-finite mass" in st["error"], st
        assert all(srv.wait(g, 60000)["state"] == "done" for g in good)
        assert srv.stats()["arena_used"] == 0


def test_fp32dev_served_job_within_1e5_of_fp64():
    job = dict(nx=256, ny=256, steps=1000, ic_seed=5, radius=60, deliver=serve.DELIVER_RHO, **DROP)
    with gpu_server("fp64") as s64, gpu_server("fp32dev") as s32:
        a, b = s64.submit(**job), s32.submit(**job)
        assert s64.wait(a, 60000)["state"] == "done" and s32.wait(b, 60000)["state"] == "done"
        err = max(float(np.max(np.abs(s64.result(a, serve.RHO, c) - s32.result(b, serve.RHO, c))))
                  for c in range(2))
    assert err <= 1e-5, err


def test_autotuned_server_runs_and_reports_its_shape():
    with serve.Server(precision=nat.FP32_DEV, budget_bytes=1 << 30) as srv:
        st = srv.stats()
        assert st["tile_cols"] in (128, 256) and st["tile_rows"] in (16, 32, 64, 128)
        assert st["lookahead"] in (1, 2) and st["workgroups"] >= 256
        jid = srv.submit(nx=200, ny=120, steps=33, ic_seed=9, radius=30, deliver=serve.DELIVER_F, **DROP)
        assert srv.wait(jid, 60000)["state"] == "done"
        tuned = [srv.result(jid, serve.F, c) for c in range(2)]
        st = srv.stats()
        assert st["job_steps"] == 33 and st["lattice_updates"] == 200 * 120 * 33
    fixed = run_alone(dict(nx=200, ny=120, steps=33, ic_seed=9, radius=30, deliver=serve.DELIVER_F, **DROP))
    for c in range(2):
        assert np.array_equal(tuned[c], fixed[c])
    # with the width timed, a caller's rows and lookahead are kept
    with serve.Server(precision=nat.FP32_DEV, budget_bytes=1 << 30, tile_rows=24, lookahead=2) as srv:
        st = srv.stats()
        assert st["tile_cols"] in (128, 256) and st["tile_rows"] == 24 and st["lookahead"] == 2
    with pytest.raises(nat.SclError):
        serve.Server(precision=nat.FP32_DEV, budget_bytes=1 << 30, lookahead=3)


@pytest.mark.parametrize("form", ["fp32dev", "fp16dev"])
@pytest.mark.parametrize("k", [2, 3])
def test_two_steps_per_launch_in_the_server_keep_results(form, k):
    """Deviation forms: jobs advance k steps per launch of the persistent
    kernel; the steps past the last whole group of k in a round (rounds of
    16 and 5 steps) take single steps.  The result is the same bits with one
    step per launch, in a mix, and on the single-lattice path."""
    wide = dict(POROUS, nx=300, ny=120, steps=151, geom_seed=4, ic_seed=6, deliver=serve.DELIVER_F)
    two = run_alone(wide, form, round_steps=16, steps_per_launch=k)
    one = run_alone(wide, form, round_steps=16, steps_per_launch=1)
    odd = run_alone(wide, form, round_steps=5, band_rows=17, steps_per_launch=k)
    for c in range(2):
        assert np.array_equal(two[c], one[c])
        assert np.array_equal(two[c], odd[c])
    with gpu_server(form, steps_per_launch=k) as srv:
        assert srv.stats()["steps_per_launch"] == k
        t0 = srv.stats()["clock_ms"]
        others = [dict(nx=257, ny=64, steps=90, ic_seed=3, radius=12, **DROP),
                  dict(nx=128, ny=200, steps=61, ic_seed=4, radius=20, **DROP),
                  dict(nx=512, ny=40, steps=33, ic_seed=5, radius=10, release_ms=t0 + 3.0, **DROP)]
        ids = [srv.submit(**o) for o in others[:2]]
        jid = srv.submit(**dict(wide, release_ms=t0 + 1.0))
        ids.append(srv.submit(**others[2]))
        assert srv.wait(jid, 60000)["state"] == "done"
        mixed = [srv.result(jid, serve.F, c) for c in range(2)]
        for i in ids:
            assert srv.wait(i, 60000)["state"] == "done"
    for c in range(2):
        assert np.array_equal(two[c], mixed[c])
    case = wl.served_case(**{k: v for k, v in wide.items() if k not in ("steps", "deliver")})
    g = nat.Geometry(case["mask"])
    fl = [nat.Field(g, case["tau"][c], nat.PRECISION[form], served_rho0(wide), g_solid=case["g_solid"][c],
                    g_fluid=case["g_fluid"][c]) for c in range(2)]
    for c in range(2):
        fl[c].set_f(case["f"][c])
    nat.solid_surface_force(g, fl[0], fl[1])
    nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], wide["steps"])
    for c in range(2):
        assert np.array_equal(fl[c].get_f().reshape(-1), two[c].reshape(-1))


@pytest.mark.parametrize("form", ["fp32dev", "fp16dev"])
@pytest.mark.parametrize("nx", [8, 37, 128, 192, 193, 224, 225, 255])
@pytest.mark.parametrize("k", [2, 3])
def test_narrow_jobs_in_two_step_strips(form, nx, k):
    """A job narrower than a k-step strip (two steps: 224 output columns in
    FP32, 250 in FP16; three: 192, 246) takes one strip, or a full one and a
    narrow one, whose lanes past the job's columns and halos idle: the same
    bits as one step per launch and as the single-lattice path's single
    steps."""
    job = dict(POROUS, nx=nx, ny=96, steps=41, rmin=2, rmax=5, radius=min(nx, 96) / 4.0, geom_seed=5,
               ic_seed=9, deliver=serve.DELIVER_F)
    two = run_alone(job, form, round_steps=16, steps_per_launch=k)
    one = run_alone(job, form, round_steps=16, steps_per_launch=1)
    for c in range(2):
        assert np.array_equal(two[c], one[c])
    case = wl.served_case(**{k: v for k, v in job.items() if k not in ("steps", "deliver")})
    g = nat.Geometry(case["mask"])
    g.set_step_fusion(1)
    fl = [nat.Field(g, case["tau"][c], nat.PRECISION[form], served_rho0(job), g_solid=case["g_solid"][c],
                    g_fluid=case["g_fluid"][c]) for c in range(2)]
    for c in range(2):
        fl[c].set_f(case["f"][c])
    nat.solid_surface_force(g, fl[0], fl[1])
    nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], job["steps"])
    for c in range(2):
        assert np.array_equal(fl[c].get_f().reshape(-1), two[c].reshape(-1))


def test_three_step_server_takes_two_steps_for_a_two_step_round():
    """With three steps per launch, a job whose last round has two steps left
    (34 steps in rounds of 16) takes them as one two-step launch, beside
    three-step jobs of the same round: the same bits as one step per
    launch."""
    short = dict(POROUS, nx=260, ny=90, steps=34, geom_seed=7, ic_seed=2, deliver=serve.DELIVER_F)
    longer = dict(POROUS, nx=300, ny=100, steps=47, geom_seed=8, ic_seed=3, deliver=serve.DELIVER_F)
    got, want = [], []
    for k, out in ((3, got), (1, want)):
        with gpu_server("fp32dev", round_steps=16, steps_per_launch=k) as srv:
            ids = [srv.submit(**short), srv.submit(**longer)]
            for i in ids:
                assert srv.wait(i, 60000)["state"] == "done"
            out.extend(srv.result(i, serve.F, c) for i in ids for c in range(2))
            if k == 3:
                st = srv.stats()
                # multi-step launches: 34 = (15 + 1) + (15 + 1) + 2 steps and
                # 47 = (15 + 1) + (15 + 1) + 15, singles in brackets' tails
                assert st["pair_updates"] == 32 * 260 * 90 + 45 * 300 * 100
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_reference_density_of_the_deviation_forms():
    """A served job stores f_i - w_i rho0 with rho0 the job's (automatic: the
    midpoint of its phases' densities): the same bits as the single-lattice
    path with that rho0; an explicit rho0 = 1 gives other bits, the same
    physics within the FP32 tolerance, and its own single-lattice twin."""
    job = dict(POROUS, steps=120, deliver=serve.DELIVER_F)
    auto = run_alone(job, "fp32dev")
    one = run_alone(dict(job, rho0=(1.0, 1.0)), "fp32dev")
    assert served_rho0(job) == 543.0 / 1024.0
    assert not np.array_equal(auto[0], one[0])
    for c in range(2):
        ra, r1 = auto[c].reshape(9, -1).sum(0), one[c].reshape(9, -1).sum(0)
        assert np.max(np.abs(ra - r1)) < 1e-5
    case = wl.served_case(**job)
    for rho0, want in ((served_rho0(job), auto), (1.0, one)):
        g = nat.Geometry(case["mask"])
        fl = [nat.Field(g, case["tau"][c], nat.FP32_DEV, rho0, g_solid=case["g_solid"][c],
                        g_fluid=case["g_fluid"][c]) for c in range(2)]
        for c in range(2):
            fl[c].set_f(case["f"][c])
        nat.solid_surface_force(g, fl[0], fl[1])
        nat.step2(g, fl[0], fl[1], case["force"][0], case["force"][1], job["steps"])
        for c in range(2):
            assert np.array_equal(fl[c].get_f().reshape(-1), want[c].reshape(-1))


@pytest.mark.parametrize("form", ["fp64", "fp32dev", "fp16dev"])
def test_suspended_job_resumes_bit_for_bit(form):
    """Pre-emption on the device: a wide porous job (two-step tiles in the
    deviation forms) suspended after some rounds leaves the device with its
    exact state, another job runs in its memory, and after resume it
    finishes with the same bits as the job run uninterrupted."""
    import time
    job = dict(POROUS, nx=300, ny=120, steps=161, geom_seed=4, ic_seed=6, deliver=serve.DELIVER_F)
    alone = run_alone(job, form, round_steps=8)
    with gpu_server(form, round_steps=8, max_jobs=1) as srv:
        jid = srv.submit(**job)
        t_end = time.time() + 30
        while srv.status(jid)["steps_done"] < 24 and time.time() < t_end:
            time.sleep(0.001)
        srv.suspend(jid)
        while srv.status(jid)["state"] != "suspended" and time.time() < t_end:
            time.sleep(0.001)
        st = srv.status(jid)
        assert st["state"] == "suspended" and 24 <= st["steps_done"] < job["steps"], st
        other = srv.submit(nx=257, ny=64, steps=40, ic_seed=3, radius=12, **DROP)
        assert srv.wait(other, 60000)["state"] == "done"
        srv.resume(jid)
        assert srv.wait(jid, 60000)["state"] == "done"
        got = [srv.result(jid, serve.F, c) for c in range(2)]
    for c in range(2):
        assert np.array_equal(got[c], alone[c])


def test_launches_without_graphs_give_the_same_bits():
    """use_graphs = 0 (the launches of each round issued one by one, for
    profilers) runs the same kernels in the same order: same populations."""
    job = dict(POROUS, nx=300, steps=77, deliver=serve.DELIVER_F)   # wide: two-step tiles too
    want = run_alone(job, "fp32dev", round_steps=16)
    got = run_alone(job, "fp32dev", round_steps=16, use_graphs=0)
    for c in range(2):
        assert np.array_equal(got[c], want[c])
