"""Benchmark of the fused D2Q9 two-component Shan-Chen step on one MI355X.

Workload (one "step" = every site of the lattice advanced by one reference
time step): the 4096 x 4096 fully periodic two-component lattice (spinodal
initial condition, rho_A = rho_B = 1 +- 1% seeded noise, tau = 1, G = 1.2),
synthetic and resident in HBM before the timed region.  The headline value
is the FP32-deviation form; the FP64 form (bit-exact with the reference) and
the FP16-deviation storage form are timed beside it.  The deviation forms
advance three (k_stepk) or two (k_step2x) steps per launch, as the
autotuner measures fastest, FP64 one (k_step).  The roofline object is for
the dominant kernel: `achieved` = the algorithmic bytes of the lattice
updates it performs (every population of both components read once and
written once per update, 144 B per site and step in FP32) over its
HIP-event time, i.e. kernel MLUPS x 144 B; a launch that fuses K steps
performs K updates per site while moving the populations once, so this
exceeds the one-step HBM roofline.  `launch_pass_gbs` (one pass over the
populations per launch) and `hbm_moved_gbs` (measured traffic per launch)
give the bytes it actually moves.  `effective_pct_hbm` is the bench's MLUPS
x those bytes per update over the peak.

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
BYTES_PER_LU = {"fp64": 288, "fp32dev": 144, "fp16dev": 72}   # 2 * 9 * 2 comps * bytes


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--n", type=int, default=4096, help="lattice side")
    p.add_argument("--forms", default="fp32dev,fp64,fp16dev")
    p.add_argument("--cpu-steps", type=int, default=3, help="timed steps of the CPU baseline (about 2.7 s each)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--tile", default="", help="cols,rows[,lookahead] override (default: autotuned)")
    p.add_argument("--fusion", default="", help="steps per launch[,band rows] override (default: autotuned)")
    p.add_argument("--no-serve", action="store_true", help="skip the served job stream")
    p.add_argument("--serve-jobs", type=int, default=2000)
    p.add_argument("--serve-step-scale", type=float, default=0.025,
                   help="multiplies the per-job step counts drawn from [2000, 50000]")
    p.add_argument("--serve-load", type=float, default=0.8, help="offered load / measured server capacity")
    p.add_argument("--serve-budget-gib", type=float, default=64.0)
    p.add_argument("--serve-sample", type=int, default=20, help="every k-th job is rerun in FP64 (5%%)")
    p.add_argument("--serve-part", default="", help="I/P: serve only the I-th of P equal slices of the stream "
                   "(jobs keep their stream numbers; per-job latencies are added to the output)")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_form(nat, wl, case, form, steps, warmup, tile, fusion=""):
    g = nat.Geometry(case["mask"])
    prec = nat.PRECISION[form]
    fl = []
    for c in range(2):
        f = nat.Field(g, case["tau"][c], prec, 1.0, g_solid=0.0, g_fluid=case["g_fluid"][c])
        f.set_f(case["f"][c])
        fl.append(f)
    nat.solid_surface_force(g, fl[0], fl[1])
    # The launch shape is chosen by measurement (on scratch fields) after the
    # uploads, right before the warmup: the ~0.3 s of GPU work it takes keeps
    # the clocks up.  With the autotuner first, the GPU idled for the ~0.7 s
    # the host needs to convert and upload the 4096^2 populations, and the
    # first ~100 ms of launches after that ran at lower clocks (k_stepk 0.75
    # ms against 0.585 ms per launch, whatever the data).
    if tile:
        v = [int(t) for t in tile.split(",")]
        g.set_tile(v[0], v[1])
        if len(v) > 2:
            g.set_lookahead(v[2])
    else:
        g.autotune(2, prec)
    if fusion:
        v = [int(t) for t in fusion.split(",")]
        g.set_step_fusion(v[0], v[1] if len(v) > 1 else 0)
    nat.step2(g, fl[0], fl[1], None, None, warmup)
    nat.synchronize()
    nat.timing(1)                       # record every launch of the timed region
    t0 = time.perf_counter()
    nat.step2(g, fl[0], fl[1], None, None, steps)
    nat.synchronize()
    dt = time.perf_counter() - t0
    avg_ms, cnt = nat.timing(0)         # average per time step over the recorded launches
    nat.timing(-1)
    mass = [f.moments()[0] for f in fl]
    tile_used = g.tile(2, prec)
    spl, band = g.step_fusion(prec)
    for f in fl:
        f.close()
    g.close()
    n_sites = case["mask"].size
    mlups = n_sites * steps / dt / 1e6
    launch_ms = avg_ms * steps / cnt if cnt else 0.0
    # launch_pass_gbs: the algorithmic bytes of one launch, one read and one
    # write of every population, whatever the steps it fuses (one: k_step,
    # two: k_step2x, three: k_stepk), over the launch time; effective_gbs:
    # the lattice updates the kernel performs x BYTES_PER_LU over its time
    # (kernel MLUPS x bytes per update), what one step per launch would move
    return dict(form=form, seconds=dt, ms_per_step=dt * 1e3 / steps, mlups=mlups,
                bytes_per_launch=n_sites * BYTES_PER_LU[form],
                kernel={1: "k_step", 2: "k_step2d" if form == "fp64" else "k_step2x", 3: "k_stepk"}[spl],
                steps_per_launch=spl, band_rows=band,
                kernel_avg_ms=launch_ms, kernel_step_ms=avg_ms, kernel_launches=cnt, tile=list(tile_used),
                effective_gbs=n_sites * BYTES_PER_LU[form] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None,
                launch_pass_gbs=n_sites * BYTES_PER_LU[form] / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else None,
                effective_pct_hbm=100.0 * mlups * 1e6 * BYTES_PER_LU[form] / (HBM_PEAK_GBS * 1e9), mass=mass)


def served_stream(nat, a):
    """The served job stream in FP32 deviations: capacity probe, Poisson
    arrivals at `serve_load` of it, p50/p99 completion latency, aggregate
    MLUPS, and the max density error of a 5% sample against an FP64 rerun."""
    from sc_lbm_amd import serve
    from sc_lbm_amd import workloads as wl
    jobs = wl.served_stream(a.serve_jobs, seed=2025, step_scale=a.serve_step_scale)
    first_job = 0
    if a.serve_part:
        part, parts = (int(x) for x in a.serve_part.split("/"))
        first_job = part * len(jobs) // parts
        jobs = jobs[first_job:(part + 1) * len(jobs) // parts]
    cfg = dict(precision=nat.FP32_DEV, budget_bytes=int(a.serve_budget_gib * (1 << 30)), max_jobs=256,
               max_pending=max(4096, 2 * a.serve_jobs))
    with serve.Server(**cfg) as srv:
        st0 = srv.stats()
        log("  server tile %d x %d, lookahead %d, %d workgroups per step"
            % (st0["tile_cols"], st0["tile_rows"], st0["lookahead"], st0["workgroups"]))
        # capacity: 64 jobs of the same mix (another seed) submitted together
        probe = wl.served_stream(64, seed=77, step_scale=a.serve_step_scale)
        ids = [srv.submit(**p) for p in probe]
        for i in ids:
            srv.wait(i)
        st1 = srv.stats()
        cap = (st1["lattice_updates"] - st0["lattice_updates"]) / (st1["clock_ms"] - st0["clock_ms"]) / 1e3
        for i in ids:
            srv.release(i)
        lus = np.array([j["nx"] * j["ny"] * j["steps"] for j in jobs], dtype=np.float64)
        rate = a.serve_load * cap * 1e6 / lus.mean()
        arr = wl.poisson_arrivals(len(jobs), rate, seed=11)
        log("  capacity %.0f MLUPS; %d jobs at %.1f/s (offered load %.2f), mean %.3g LU/job"
            % (cap, len(jobs), rate, a.serve_load, lus.mean()))
        t0 = srv.stats()["clock_ms"] + 50.0
        sample = set(k for k in range(len(jobs)) if (first_job + k) % a.serve_sample == 0)
        ids = []
        for k, j in enumerate(jobs):
            ids.append(srv.submit(release_ms=t0 + arr[k], deliver=serve.DELIVER_RHO if k in sample else 0, **j))
        sts = []
        for k, i in enumerate(ids):
            sts.append(srv.wait(i))
            if (k + 1) % 100 == 0:
                log("  %d jobs finished" % (k + 1))
        rho32 = {k: [srv.result(ids[k], serve.RHO, c) for c in range(2)]
                 for k in sorted(sample) if sts[k]["state"] == "done"}
        st2 = srv.stats()
    done = [k for k, s in enumerate(sts) if s["state"] == "done"]
    lat = np.array([sts[k]["done_ms"] - sts[k]["release_ms"] for k in done])
    first = min(s["release_ms"] for s in sts)
    last = max(sts[k]["done_ms"] for k in done)
    lu_done = float(lus[done].sum())
    # FP64 rerun of the sample
    errs, worst, errs_by_job = [], [], []
    with serve.Server(precision=nat.FP64, budget_bytes=int(a.serve_budget_gib * (1 << 30)), max_jobs=256) as s64:
        ids64 = {k: s64.submit(deliver=serve.DELIVER_RHO, **jobs[k]) for k in rho32}
        for k, i in ids64.items():
            st = s64.wait(i)
            if st["state"] != "done":
                continue
            e = max(float(np.max(np.abs(s64.result(i, serve.RHO, c) - rho32[k][c]))) for c in range(2))
            errs.append(e)
            errs_by_job.append([first_job + k, e])
            j = jobs[k]
            worst.append(dict(err=e, job=first_job + k, nx=j["nx"], ny=j["ny"], steps=j["steps"], tau=round(j["tau"][0], 4),
                              G=round(j["g_fluid"][0], 4), porosity=round(j["porosity"], 4),
                              ic="droplet" if j["ic"] == serve.IC_DROPLET else "stratified",
                              rho_min_fp64=min(float(s64.result(i, serve.RHO, c).min()) for c in range(2))))
    worst.sort(key=lambda w: -w["err"])
    failed_jobs = [dict(job=first_job + k, nx=jobs[k]["nx"], ny=jobs[k]["ny"], steps=jobs[k]["steps"],
                        tau=round(jobs[k]["tau"][0], 4), G=round(jobs[k]["g_fluid"][0], 4),
                        porosity=round(jobs[k]["porosity"], 4), error=sts[k].get("error", ""),
                        steps_done=sts[k].get("steps_done"))
                   for k, st in enumerate(sts) if st["state"] == "failed"]
    scope = ("the served stream as specified (steps U[2000, 50000])" if a.serve_step_scale == 1.0 else
             "a shortened served stream: the specified steps U[2000, 50000] x %g; not the specified "
             "configuration, whose full run is recorded in profiles/served_stream_full.md" % a.serve_step_scale)
    extra = {}
    if a.serve_part:
        extra = dict(part=a.serve_part, first_job=first_job, sample_errs=sorted(errs_by_job),
                     latency_ms=[round(float(s["done_ms"] - s["release_ms"]), 1) if s["state"] == "done" else None
                                 for s in sts],
                     lattice_updates=lu_done, first_release_ms=first, last_done_ms=last)
    return dict(scope=scope, jobs=len(jobs), done=len(done), failed=sum(s["state"] == "failed" for s in sts),
                failed_jobs=failed_jobs,
                step_scale=a.serve_step_scale, steps_range=[int(round(2000 * a.serve_step_scale)),
                                                            int(round(50000 * a.serve_step_scale))],
                offered_load=a.serve_load, capacity_mlups=cap, arrival_rate_per_s=rate,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()),
                aggregate_mlups=lu_done / (last - first) / 1e3,
                # fluid sites only (solids count in the dense figure above)
                aggregate_fluid_mlups=(st2["fluid_updates"] - st1["fluid_updates"]) / (last - first) / 1e3,
                fluid_fraction=(st2["fluid_updates"] - st1["fluid_updates"]) /
                max(1, st2["lattice_updates"] - st1["lattice_updates"]),
                busy_mlups=st2["lattice_updates"] / st2["busy_ms"] / 1e3,
                device_mlups=st2["lattice_updates"] / st2["device_ms"] / 1e3 if st2["device_ms"] else None,
                busy_ms=st2["busy_ms"], device_ms=st2["device_ms"], setup_ms=st2["setup_ms"],
                finish_ms=st2["finish_ms"], rounds=st2["rounds"],
                makespan_s=(last - first) / 1e3, max_resident=st2["max_resident"],
                arena_peak_gib=st2["arena_peak"] / (1 << 30),
                tile=[st2["tile_cols"], st2["tile_rows"], st2["lookahead"]], round_steps=st2["round_steps"],
                sample_jobs=len(errs), sample_max_abs_rho_err=max(errs) if errs else None,
                sample_median_abs_rho_err=float(np.median(errs)) if errs else None,
                sample_frac_within_1e5=float(np.mean(np.array(errs) <= 1e-5)) if errs else None,
                sample_worst=worst[:5], **extra)


def cpu_baseline(n, steps):
    """Serial CPU time of the same two-component step on the same 4096^2
    lattice: the oracle restatement (oracle/sc_oracle.cpp, the reference's
    operation order), one thread pinned to one core in a child process, one
    untimed step and `steps` timed ones.  The compiled reference's own step
    runs at 0.44x the oracle's on one core at 4096^2 (measured in the build
    container: 2.79 against 6.34 MLUPS)."""
    from oracle import cpu_baseline as cb
    r = cb.run_child(n=n, steps=steps, kind="oracle")
    return dict(value=r["mlups"], unit="MLUPS", cores=1, kind="port",
                sample="two-component Shan-Chen step, %dx%d periodic spinodal (the GPU workload), %d timed steps "
                       "in %.1f s (%s s each) after one untimed step; 1 thread pinned to core %s of %d (%s)"
                       % (n, n, steps, r["seconds"], ", ".join("%.2f" % t for t in r["per_step_s"]),
                          r["core"][0], os.cpu_count() or 0, cb.cpu_model()))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d refused)" % a.gpus)
        sys.exit(2)
    from sc_lbm_amd import native as nat
    from sc_lbm_amd import workloads as wl
    nat.init(0)
    log("building the %dx%d two-component case" % (a.n, a.n))
    case = wl.spinodal(a.n, G=1.2, seed=777)
    res = {}
    for form in a.forms.split(","):
        log("timing %s: %d warmup + %d steps" % (form, a.warmup, a.steps))
        res[form] = time_form(nat, wl, case, form, a.steps, a.warmup, a.tile, a.fusion)
        log("  %s: %.1f MLUPS, %s %.4f ms per launch (%d steps), %.0f GB/s per launch pass (%.0f GB/s per step)"
            % (form, res[form]["mlups"], res[form]["kernel"], res[form]["kernel_avg_ms"],
               res[form]["steps_per_launch"], res[form]["launch_pass_gbs"] or 0.0, res[form]["effective_gbs"] or 0.0))
    head = res.get("fp32dev") or next(iter(res.values()))
    traffic = valu_busy = tsrc = wg_per_cu = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s_%s_%d.json" % (head["kernel"], head["form"], a.n))
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        traffic = tj.get("bytes_per_launch")
        valu_busy = tj.get("valu_busy_pct")
        wg_per_cu = tj.get("workgroups_per_cu")
        tsrc = os.path.relpath(tpath, ROOT)
    achieved = head["launch_pass_gbs"]
    if valu_busy is not None and achieved and achieved < 0.7 * HBM_PEAK_GBS:
        limiter = ("instruction issue and latency, not HBM bandwidth: VALU busy %.1f%%%s, HBM at %.0f%% of peak"
                   % (valu_busy, ", %d workgroups per CU" % wg_per_cu if wg_per_cu else "",
                      100.0 * achieved / HBM_PEAK_GBS))
    else:
        limiter = "HBM bandwidth"
    out = {
        "metric": "MLUPS + %HBM peak (FP32-dev vs FP64), 2-comp SC; p50/p99 served-job latency",
        "value": head["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head["form"] == "fp32dev" else ("f64" if head["form"] == "fp64" else "f16-store/f32"),
        "data": "synthetic",
        "config": {"workload": "two-component Shan-Chen %dx%d periodic spinodal (rho 1 +- 1%%, tau 1, G 1.2), "
                               "%s fused step" % (a.n, a.n, head["form"]),
                   "lattice": [a.n, a.n], "components": 2, "form": head["form"],
                   "tile": head["tile"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "bytes_per_launch": head["bytes_per_launch"],
                     "kernel": head["kernel"], "steps_per_launch": head["steps_per_launch"],
                     "limiter": limiter, "valu_busy_pct": valu_busy, "traffic_source": tsrc,
                     "per_step_effective_gbs": head["effective_gbs"],
                     "per_step_effective_frac": head["effective_gbs"] / HBM_PEAK_GBS if head["effective_gbs"] else None,
                     "note": "achieved: the algorithmic bytes of one launch of the kernel (every population of both "
                             "components read once and written once, %d B per site, whatever the %d steps it "
                             "fuses) over its HIP-event time; traffic: FETCH_SIZE + WRITE_SIZE per launch from the "
                             "profile in traffic_source, calibrated on a copy kernel; per_step_effective: the "
                             "kernel's lattice updates x %d B over its time, the bandwidth a one-step-per-launch "
                             "kernel would need for the same MLUPS"
                             % (BYTES_PER_LU[head["form"]], head["steps_per_launch"], BYTES_PER_LU[head["form"]])},
        "forms": {k: {kk: v[kk] for kk in ("mlups", "effective_pct_hbm", "kernel", "steps_per_launch",
                                           "band_rows", "kernel_avg_ms", "launch_pass_gbs", "effective_gbs",
                                           "tile")}
                  for k, v in res.items()},
    }
    if "fp64" in res and "fp32dev" in res:
        out["fp32dev_over_fp64"] = res["fp32dev"]["mlups"] / res["fp64"]["mlups"]
    if not a.no_serve:
        log("served job stream: %d jobs, step scale %g" % (a.serve_jobs, a.serve_step_scale))
        out["served"] = served_stream(nat, a)
        sv = out["served"]
        log("  p50 %.1f ms, p99 %.1f ms, aggregate %.0f MLUPS, %d failed, sample max |drho| %s"
            % (sv["p50_ms"], sv["p99_ms"], sv["aggregate_mlups"], sv["failed"], sv["sample_max_abs_rho_err"]))
    if not a.no_cpu:
        log("timing the CPU baseline (%d steps at %d^2, one core)" % (a.cpu_steps, a.n))
        out["cpu_baseline"] = cpu_baseline(a.n, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
