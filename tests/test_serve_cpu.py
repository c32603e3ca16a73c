"""CPU tests of the serving layer's host logic: the job generator of
libsclbm.so against its Python restatement, and the scheduler (the same
serve_core.cpp libsclbm.so links) driven by an oracle backend
(tests/drivers/serve_cpu.cpp): results, admission by memory budget,
back-pressure, release times, cancellation, time-outs, failure isolation
and the mass watchdog, all forced from the host."""
import ctypes
import os
import time

import numpy as np
import pytest

from oracle import pyoracle as po
from sc_lbm_amd import serve
from sc_lbm_amd import workloads as wl

HERE = os.path.dirname(os.path.abspath(__file__))
SERVE_CPU = os.path.join(HERE, "drivers", "bin", "libserve_cpu.so")

DROP = dict(ic=serve.IC_DROPLET, rho_hi=1.0, rho_lo=0.06, noise=0.01, tau=(1.0, 0.8), g_fluid=(0.9, 0.9))


@pytest.fixture(scope="module")
def cpulib():
    import sys
    sys.path.insert(0, os.path.join(HERE, "drivers"))
    import case_io
    if not os.path.exists(SERVE_CPU) or os.path.getmtime(SERVE_CPU) < max(
            os.path.getmtime(p) for p in [os.path.join(HERE, "drivers", "serve_cpu.cpp"),
                                          os.path.join(HERE, "..", "sc_lbm_amd", "host", "serve_core.cpp"),
                                          os.path.join(HERE, "..", "sc_lbm_amd", "host", "serve_core.h")]):
        case_io.build_serve_cpu()
    lib = ctypes.CDLL(SERVE_CPU)
    lib.sct_set_round_delay.argtypes = [ctypes.c_int]
    lib.sct_fail_setup_for_seed.argtypes = [ctypes.c_ulonglong]
    yield lib
    lib.sct_set_round_delay(0)
    lib.sct_fail_setup_for_seed(2 ** 64 - 1)


def cpu_server(lib, **cfg):
    base = dict(budget_bytes=64 << 20, round_steps=8, precision=0)
    base.update(cfg)
    return serve.Server(lib=lib, create="sct_server_create", **base)


def oracle_run(case, steps):
    s = po.System(case["mask"], case["tau"], case["g_fluid"], case["g_solid"])
    for c in range(2):
        s.set_f(c, case["f"][c])
    s.solid_surface_force()
    s.run(steps, case["force"][0], case["force"][1])
    return [s.get_f(c) for c in range(2)]


def dens_ref(f):
    """Density in the reference order ((0 + f0) + f1) + ... (src/fluid.cpp:177-182)."""
    p = f.reshape(9, -1)
    r = np.zeros(p.shape[1])
    for i in range(9):
        r = r + p[i]
    return r


# ---------------------------------------------------------------------------
# generator
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nx,ny,por,seed,walls,rmin,rmax", [
    (120, 80, 0.7, 3, True, 8, 20),
    (200, 64, 0.85, 11, False, 4, 9),
    (96, 96, 0.6, 5, False, 8, 20),
    (64, 48, 1.0, 0, True, 8, 20),
    (257, 129, 0.9, 77, True, 2, 6),
])
def test_job_mask_matches_workloads_sampler(nx, ny, por, seed, walls, rmin, rmax):
    m = serve.job_mask(nx=nx, ny=ny, steps=1, porosity=por, rmin=rmin, rmax=rmax, walls=walls, geom_seed=seed)
    if por < 1.0:
        ref, disks = wl.porous_mask(nx, ny, por, seed, rmin, rmax, walls=walls)
        assert len(disks) > 0
    else:
        ref = wl.walls_x(nx, ny) if walls else wl.open_mask(nx, ny)
    assert np.array_equal(m, ref)
    assert m.mean() <= por or por == 1.0


def test_job_mask_passes_explicit_mask_through():
    rng = np.random.default_rng(0)
    mask = (rng.random((30, 50)) > 0.3).astype(np.int32) * 7   # any nonzero is fluid
    m = serve.job_mask(nx=50, ny=30, steps=1, mask=mask)
    assert np.array_equal(m, (mask != 0).astype(np.int32))


def test_job_mask_rejects_bad_specs():
    with pytest.raises(serve.ServerError):
        serve.job_mask(nx=0, ny=10, steps=1)
    with pytest.raises(serve.ServerError):
        serve.job_mask(nx=10, ny=10, steps=0)
    with pytest.raises(serve.ServerError):
        serve.job_mask(nx=10, ny=10, steps=1, porosity=0.5, rmin=5, rmax=3)


# ---------------------------------------------------------------------------
# scheduler with the oracle backend
# ---------------------------------------------------------------------------
def test_results_equal_the_oracle_run_of_the_same_case(cpulib):
    jobs = [dict(nx=48, ny=40, steps=29, porosity=0.85, rmin=3, rmax=6, geom_seed=1, walls=True, radius=9,
                 ic_seed=10, g_solid=(0.1, -0.1), force=(wl.body_force(1e-5), None), **DROP),
            dict(nx=64, ny=32, steps=8, ic=serve.IC_STRATIFIED, rho_hi=1.0, rho_lo=0.1, noise=0.02, ic_seed=3,
                 tau=(0.9, 1.1), g_fluid=(1.0, 1.0)),
            dict(nx=33, ny=35, steps=1, ic=serve.IC_UNIFORM, rho_hi=0.7, rho_lo=0.3, noise=0.0,
                 g_fluid=(0.5, 0.5))]
    with cpu_server(cpulib, round_steps=4) as srv:
        ids = [srv.submit(deliver=serve.DELIVER_F | serve.DELIVER_RHO, **j) for j in jobs]
        for jid, j in zip(ids, jobs):
            st = srv.wait(jid, 20000)
            assert st["state"] == "done" and st["steps_done"] == j["steps"]
            ref = oracle_run(wl.served_case(**j), j["steps"])
            for c in range(2):
                f = srv.result(jid, serve.F, c)
                assert np.array_equal(f.reshape(-1), ref[c])
                rho = srv.result(jid, serve.RHO, c)
                assert np.array_equal(rho.reshape(-1), dens_ref(ref[c]))


def test_explicit_populations_and_mask(cpulib):
    case = wl.droplet(n=40, radius=9, G=0.9, seed=5)
    with cpu_server(cpulib) as srv:
        jid = srv.submit(nx=40, ny=40, steps=13, mask=case["mask"], f=case["f"], tau=case["tau"],
                         g_fluid=case["g_fluid"], deliver=serve.DELIVER_F)
        assert srv.wait(jid, 20000)["state"] == "done"
        ref = oracle_run(case, 13)
        for c in range(2):
            assert np.array_equal(srv.result(jid, serve.F, c).reshape(-1), ref[c])


def test_budget_admission_and_alone_vs_mix(cpulib):
    """An arena that holds two of these jobs: never more than two resident,
    every job completes, and each result equals the same job run alone."""
    spec = [dict(nx=40, ny=40, steps=24 + 3 * k, ic_seed=50 + k, radius=8, **DROP) for k in range(6)]
    per_job = 40 * 40 * 192 + 4096
    with cpu_server(cpulib, budget_bytes=2 * per_job + 8192) as srv:
        ids = [srv.submit(deliver=serve.DELIVER_F, **s) for s in spec]
        sts = [srv.wait(i, 30000) for i in ids]
        st = srv.stats()
        mixed = [[srv.result(i, serve.F, c) for c in range(2)] for i in ids]
    assert all(s["state"] == "done" for s in sts)
    assert st["max_resident"] == 2 and st["done"] == 6
    assert st["arena_peak"] <= st["arena_bytes"]
    for k, s in enumerate(spec):
        with cpu_server(cpulib) as srv:
            jid = srv.submit(deliver=serve.DELIVER_F, **s)
            srv.wait(jid, 30000)
            for c in range(2):
                assert np.array_equal(srv.result(jid, serve.F, c), mixed[k][c])


def test_oversized_and_invalid_jobs_are_rejected(cpulib):
    with cpu_server(cpulib, budget_bytes=1 << 20) as srv:
        with pytest.raises(IndexError):        # SCL_ERR_RANGE: can never fit the arena
            srv.submit(nx=200, ny=200, steps=5)
        with pytest.raises(serve.ServerError):
            srv.submit(nx=10, ny=10, steps=5, tau=(0.0, 1.0))
        with pytest.raises(serve.ServerError):
            srv.submit(nx=10, ny=10, steps=5, f=[np.zeros(900), None])
        assert srv.stats()["rejected"] == 3


def test_queue_full_is_back_pressure(cpulib):
    cpulib.sct_set_round_delay(50)
    try:
        with cpu_server(cpulib, max_pending=2, max_jobs=1) as srv:
            a = srv.submit(nx=16, ny=16, steps=64)      # resident, slow rounds
            time.sleep(0.1)
            srv.submit(nx=16, ny=16, steps=1)
            srv.submit(nx=16, ny=16, steps=1)
            with pytest.raises(serve.Busy):
                srv.submit(nx=16, ny=16, steps=1)
            assert srv.stats()["rejected_busy"] == 1
            srv.cancel(a)
            srv.drain(20000)
            assert srv.stats()["done"] == 2
    finally:
        cpulib.sct_set_round_delay(0)


def test_release_time_is_respected(cpulib):
    with cpu_server(cpulib) as srv:
        t0 = srv.stats()["clock_ms"]
        jid = srv.submit(nx=16, ny=16, steps=3, release_ms=t0 + 150.0)
        st = srv.wait(jid, 10000)
        assert st["state"] == "done" and st["admit_ms"] >= t0 + 150.0


def test_cancel_queued_and_resident_jobs(cpulib):
    cpulib.sct_set_round_delay(20)
    try:
        with cpu_server(cpulib, max_jobs=1) as srv:
            a = srv.submit(nx=16, ny=16, steps=10000)
            b = srv.submit(nx=16, ny=16, steps=5)
            time.sleep(0.1)
            assert srv.status(a)["state"] == "running" and srv.status(b)["state"] == "queued"
            srv.cancel(b)
            assert srv.status(b)["state"] == "cancelled"      # queued: dropped at once
            srv.cancel(a)
            st = srv.wait(a, 5000)
            assert st["state"] == "cancelled" and 0 < st["steps_done"] < 10000
            assert srv.stats()["arena_used"] == 0
    finally:
        cpulib.sct_set_round_delay(0)


def test_timeouts_queued_and_resident(cpulib):
    cpulib.sct_set_round_delay(20)
    try:
        with cpu_server(cpulib, max_jobs=1) as srv:
            a = srv.submit(nx=16, ny=16, steps=100000, timeout_ms=150)
            b = srv.submit(nx=16, ny=16, steps=5, timeout_ms=50)
            sa, sb = srv.wait(a, 5000), srv.wait(b, 5000)
            assert sa["state"] == "timed_out" and sb["state"] == "timed_out"
            assert sb["admit_ms"] < 0 and sa["steps_done"] > 0
            c = srv.submit(nx=16, ny=16, steps=5)
            assert srv.wait(c, 5000)["state"] == "done"
    finally:
        cpulib.sct_set_round_delay(0)


def test_nan_job_fails_alone(cpulib):
    case = wl.droplet(n=24, radius=5, G=0.9, seed=1)
    bad = [case["f"][0].copy(), case["f"][1]]
    bad[0][3 * 24 * 24 + 100] = np.nan
    with cpu_server(cpulib) as srv:
        good = [srv.submit(nx=24, ny=24, steps=20, ic_seed=k, radius=5, **DROP) for k in range(3)]
        jbad = srv.submit(nx=24, ny=24, steps=20, f=bad, tau=case["tau"], g_fluid=case["g_fluid"])
        st = srv.wait(jbad, 10000)
        assert st["state"] == "failed" and "non-finite mass" in st["error"]
        assert all(srv.wait(g, 10000)["state"] == "done" for g in good)
        assert srv.stats()["failed"] == 1


def test_setup_failure_is_isolated(cpulib):
    cpulib.sct_fail_setup_for_seed(4242)
    try:
        with cpu_server(cpulib) as srv:
            a = srv.submit(nx=20, ny=20, steps=5, ic_seed=4242)
            b = srv.submit(nx=20, ny=20, steps=5, ic_seed=1)
            sa = srv.wait(a, 5000)
            assert sa["state"] == "failed" and "injected setup failure" in sa["error"]
            assert srv.wait(b, 5000)["state"] == "done"
            assert srv.stats()["arena_used"] == 0
    finally:
        cpulib.sct_fail_setup_for_seed(2 ** 64 - 1)


def test_large_job_is_not_starved(cpulib):
    """A job too large for the free arena is passed over by at most
    max_bypass later jobs; then admission waits for it."""
    small = 16 * 16 * 192 + 4096
    big = 48 * 48 * 192 + 4096
    cpulib.sct_set_round_delay(10)
    try:
        with cpu_server(cpulib, budget_bytes=big + small, max_bypass=2, round_steps=1) as srv:
            holder = srv.submit(nx=16, ny=16, steps=6)           # keeps the arena from holding the big job
            time.sleep(0.05)
            jbig = srv.submit(nx=48, ny=48, steps=2)
            smalls = [srv.submit(nx=16, ny=16, steps=6) for _ in range(5)]
            sts = {i: srv.wait(i, 20000) for i in [holder, jbig] + smalls}
            assert all(s["state"] == "done" for s in sts.values())
            big_admit = sts[jbig]["admit_ms"]
            before = [i for i in smalls if sts[i]["admit_ms"] < big_admit]
            assert len(before) <= 2
    finally:
        cpulib.sct_set_round_delay(0)


def test_destroy_with_live_jobs_returns_promptly(cpulib):
    cpulib.sct_set_round_delay(20)
    try:
        srv = cpu_server(cpulib, max_jobs=1)
        srv.submit(nx=16, ny=16, steps=100000)
        srv.submit(nx=16, ny=16, steps=100000)
        time.sleep(0.05)
        t0 = time.perf_counter()
        srv.close()   # resident and queued jobs end cancelled; the loop joins
        assert time.perf_counter() - t0 < 2.0
    finally:
        cpulib.sct_set_round_delay(0)


def test_wait_timeout_and_unknown_ids(cpulib):
    cpulib.sct_set_round_delay(20)
    try:
        with cpu_server(cpulib) as srv:
            a = srv.submit(nx=16, ny=16, steps=100000)
            with pytest.raises(serve.WaitTimeout):
                srv.wait(a, 30)
            with pytest.raises(serve.ServerError):
                srv.status(999)
            with pytest.raises(serve.ServerError):
                srv.result(a, serve.RHO, 0)             # not finished
            with pytest.raises(serve.ServerError):
                srv.release(a)                           # still live
            srv.cancel(a)
            srv.wait(a, 5000)
            srv.release(a)
            with pytest.raises(serve.ServerError):
                srv.status(a)
    finally:
        cpulib.sct_set_round_delay(0)


def test_fluid_updates_count_fluid_sites_only(cpulib):
    """The statistics count lattice updates of fluid sites beside all sites:
    a generated porous job and a job with a caller mask."""
    gen = dict(nx=48, ny=40, steps=20, porosity=0.8, rmin=3, rmax=6, walls=True, geom_seed=5, ic_seed=1, radius=8,
               **DROP)
    case = wl.porous(nx=40, ny=24, porosity=0.8, seed=3, rmin=3, rmax=5)
    caller = dict(nx=40, ny=24, steps=12, mask=case["mask"], f=case["f"], tau=case["tau"], g_fluid=case["g_fluid"])
    with cpu_server(cpulib) as srv:
        ids = [srv.submit(**gen), srv.submit(**caller)]
        for i in ids:
            assert srv.wait(i, 30000)["state"] == "done"
        st = srv.stats()
    fluid_gen = int(serve.job_mask(**gen).sum())
    fluid_caller = int(case["mask"].sum())
    assert st["lattice_updates"] == 48 * 40 * 20 + 40 * 24 * 12
    assert st["fluid_updates"] == fluid_gen * 20 + fluid_caller * 12
    assert 0 < st["fluid_updates"] < st["lattice_updates"]


def _wait_state(srv, jid, state, timeout_s=10.0):
    t_end = time.time() + timeout_s
    while time.time() < t_end:
        st = srv.status(jid)
        if st["state"] == state:
            return st
        time.sleep(0.005)
    raise AssertionError("job %d did not reach %s: %s" % (jid, state, srv.status(jid)))


def test_suspend_and_resume_continue_bit_for_bit(cpulib):
    """Pre-emption: a resident job suspended mid-run leaves the device (its
    memory freed, another job admitted in its place), waits suspended, and
    after resume finishes with the populations of the uninterrupted oracle
    run.  A queued job is set aside at once; suspending twice is harmless;
    finished jobs refuse; a suspended job can be cancelled; resume of a job
    whose suspension is still pending just keeps it running."""
    job = dict(nx=48, ny=40, steps=60, porosity=0.85, rmin=3, rmax=6, geom_seed=1, walls=True, radius=9,
               ic_seed=10, g_solid=(0.1, -0.1), force=(wl.body_force(1e-5), None), **DROP)
    ref = oracle_run(wl.served_case(**job), job["steps"])
    cpulib.sct_set_round_delay(15)
    try:
        with cpu_server(cpulib, round_steps=4, max_jobs=1) as srv:
            a = srv.submit(deliver=serve.DELIVER_F, **job)
            b = srv.submit(nx=16, ny=16, steps=8)
            q = srv.submit(nx=16, ny=16, steps=8)
            srv.suspend(q)                                   # queued: set aside at once
            assert srv.status(q)["state"] == "suspended"
            _wait_state(srv, a, "running")
            while srv.status(a)["steps_done"] < 8:
                time.sleep(0.005)
            srv.suspend(a)
            srv.suspend(a)
            st = _wait_state(srv, a, "suspended")
            assert 8 <= st["steps_done"] < job["steps"]
            assert srv.wait(b, 10000)["state"] == "done"      # admitted in a's place
            assert srv.status(a)["state"] == "suspended" and srv.stats()["arena_used"] == 0
            srv.resume(a)
            srv.resume(q)
            assert srv.wait(a, 20000)["state"] == "done"
            assert srv.wait(q, 20000)["state"] == "done"
            for c in range(2):
                assert np.array_equal(srv.result(a, serve.F, c).reshape(-1), ref[c])
            with pytest.raises(serve.ServerError):
                srv.suspend(a)                               # finished
            with pytest.raises(serve.ServerError):
                srv.resume(b)                                # not suspended
            c_ = srv.submit(nx=16, ny=16, steps=100000)
            _wait_state(srv, c_, "running")
            srv.suspend(c_)
            srv.resume(c_)                                   # still pending: keeps running
            time.sleep(0.1)
            assert srv.status(c_)["state"] == "running"
            srv.suspend(c_)
            _wait_state(srv, c_, "suspended")
            srv.cancel(c_)
            assert srv.status(c_)["state"] == "cancelled"
    finally:
        cpulib.sct_set_round_delay(0)
