"""Lean drivers for profiler runs (rocprofv3 puts the program after `--`).

    python -m sc_lbm_amd.profiling step --form fp32dev --n 4096 --steps 20 [--porosity 0.8]
            [--tile 512,128,1] [--fusion 2 --band 16]
        builds the 4096^2 two-component spinodal case (optionally with seeded
        disk obstacles), runs `steps` time steps of one storage form (one or
        two per launch) and prints the mean kernel time per launch;
    python -m sc_lbm_amd.profiling copy --bytes 2147483648 --reps 10
        runs the plain copy kernel k_probe_copy (reads and writes `bytes`
        each launch) so the HBM byte counters can be calibrated;
    python -m sc_lbm_amd.profiling serve --jobs 64 --step-scale 0.025 [--form fp32dev]
        runs a batch of jobs of the served mix (submitted together) through a
        server that issues its launches one by one instead of from captured
        graphs (profilers trace graph launches poorly) and prints the
        lattice updates per second of the batch;
    python -m sc_lbm_amd.profiling geometry --n 2048 --porosity 0.6 --reps 5
        builds a served job's porous mask on the host (sampler, rasteriser,
        packing) and through the device pipeline (host sampler, device
        raster, pack and count) and prints the times of both;
    python -m sc_lbm_amd.profiling noise --jobs 860,140 [--eps 6e-8]
        for jobs of the served stream (2000 jobs, seed 2025, full step
        counts) prints the FP32-deviation run's max density error against
        FP64 and the FP64 noise floor: FP64 from the same populations
        perturbed by a relative eps against FP64 from the unperturbed ones.
"""
import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main(argv=None):
    p = argparse.ArgumentParser()
    sub = p.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("step")
    s.add_argument("--form", default="fp32dev")
    s.add_argument("--n", type=int, default=4096)
    s.add_argument("--steps", type=int, default=20)
    s.add_argument("--warmup", type=int, default=2)
    s.add_argument("--tile", default="")
    s.add_argument("--fusion", type=int, default=0, help="time steps per launch: 0 automatic, 1, 2 or 3")
    s.add_argument("--band", type=int, default=0, help="rows per band of the multi-step kernels (0 automatic)")
    s.add_argument("--porosity", type=float, default=1.0,
                   help="below 1: seeded disk obstacles (radii 8..20) to this fluid fraction")
    c = sub.add_parser("copy")
    c.add_argument("--bytes", type=float, default=2 ** 31)
    c.add_argument("--reps", type=int, default=10)
    v = sub.add_parser("serve")
    v.add_argument("--form", default="fp32dev")
    v.add_argument("--jobs", type=int, default=64)
    v.add_argument("--seed", type=int, default=77)
    v.add_argument("--step-scale", type=float, default=0.025)
    v.add_argument("--graphs", type=int, default=0)
    z = sub.add_parser("noise")
    z.add_argument("--jobs", default="860")
    z.add_argument("--eps", type=float, default=6e-8)
    q = sub.add_parser("geometry")
    q.add_argument("--n", type=int, default=2048)
    q.add_argument("--porosity", type=float, default=0.6)
    q.add_argument("--rmin", type=int, default=8)
    q.add_argument("--rmax", type=int, default=20)
    q.add_argument("--reps", type=int, default=5)
    a = p.parse_args(argv)

    from sc_lbm_amd import native as nat
    nat.init(0)
    if a.cmd == "serve":
        from sc_lbm_amd import serve
        from sc_lbm_amd import workloads as wl
        jobs = wl.served_stream(a.jobs, seed=a.seed, step_scale=a.step_scale)
        with serve.Server(precision=a.form, budget_bytes=64 << 30, max_jobs=256, use_graphs=a.graphs) as srv:
            st0 = srv.stats()
            ids = [srv.submit(**j) for j in jobs]
            sts = [srv.wait(i) for i in ids]
            st1 = srv.stats()
        lu = st1["lattice_updates"] - st0["lattice_updates"]
        ms = st1["clock_ms"] - st0["clock_ms"]
        print(json.dumps({"form": a.form, "jobs": a.jobs, "step_scale": a.step_scale, "graphs": a.graphs,
                          "done": sum(x["state"] == "done" for x in sts), "lattice_updates": lu, "ms": ms,
                          "mlups": lu / ms / 1e3, "device_mlups": lu / st1["device_ms"] / 1e3,
                          "tile": [st1["tile_cols"], st1["tile_rows"], st1["lookahead"]],
                          "steps_per_launch": st1["steps_per_launch"], "rounds": st1["rounds"],
                          "pair_updates": st1["pair_updates"] - st0["pair_updates"],
                          "sites": int(sum(j["nx"] * j["ny"] for j in jobs)),
                          "device_ms": st1["device_ms"] - st0["device_ms"]}))
        return
    if a.cmd == "noise":
        from sc_lbm_amd import serve
        from sc_lbm_amd import workloads as wl
        stream = wl.served_stream(2000, seed=2025, step_scale=1.0)
        ks = [int(k) for k in a.jobs.split(",")]
        runs = {}

        def collect(srv, ids, prec):
            for key, i in ids.items():
                st = srv.wait(i)
                runs[(prec,) + key] = ([srv.result(i, serve.RHO, c) for c in range(2)]
                                       if st["state"] == "done" else None)
                srv.release(i)
                print("precision %d job %d %s: %s" % (prec, key[0], key[1], st["state"]), file=sys.stderr,
                      flush=True)

        with serve.Server(precision=nat.FP32_DEV, budget_bytes=64 << 30, max_jobs=256) as srv:
            collect(srv, {(k, "gen"): srv.submit(deliver=serve.DELIVER_RHO, **stream[k]) for k in ks}, nat.FP32_DEV)
        with serve.Server(precision=nat.FP64, budget_bytes=64 << 30, max_jobs=256) as srv:
            for k in ks:   # one job's three runs at a time
                j = stream[k]
                m = serve.job_mask(**j)
                ra, rb = wl.served_density(j["nx"], j["ny"], j["ic"], j["rho_hi"], j["rho_lo"], j["radius"],
                                           j["noise"], j["ic_seed"])
                f0 = [wl.f_from_rho(ra, m), wl.f_from_rho(rb, m)]
                u = [wl.uniform01(1000 + 2 * k + c, f0[c].size) for c in range(2)]
                fp = [f0[c] * (1.0 + a.eps * (2.0 * u[c] - 1.0)) for c in range(2)]
                base = {kk: v for kk, v in j.items() if kk not in ("porosity", "geom_seed", "ic", "ic_seed")}
                collect(srv, {(k, "gen"): srv.submit(deliver=serve.DELIVER_RHO, **j),
                              (k, "exp"): srv.submit(deliver=serve.DELIVER_RHO, mask=m, f=f0, **base),
                              (k, "pert"): srv.submit(deliver=serve.DELIVER_RHO, mask=m, f=fp, **base)}, nat.FP64)
                g64, e64, p64 = (runs[(nat.FP64, k, t)] for t in ("gen", "exp", "pert"))
                g32 = runs[(nat.FP32_DEV, k, "gen")]
                err = lambda x, y: (max(float(np.max(np.abs(x[c] - y[c]))) for c in range(2))
                                    if x is not None and y is not None else None)
                print(json.dumps({"job": k, "nx": j["nx"], "ny": j["ny"], "steps": j["steps"], "eps": a.eps,
                                  "fp32_vs_fp64": err(g32, g64), "fp64_noise_floor": err(p64, e64),
                                  "explicit_equals_generated": bool(
                                      e64 is not None and g64 is not None and
                                      all(np.array_equal(e64[c], g64[c]) for c in range(2)))}), flush=True)
                for t in ("gen", "exp", "pert"):
                    runs.pop((nat.FP64, k, t), None)
        return
    if a.cmd == "geometry":
        import time
        from sc_lbm_amd import serve
        spec = dict(nx=a.n, ny=a.n, steps=1, porosity=a.porosity, rmin=a.rmin, rmax=a.rmax, walls=True)
        host_ms, samp_ms, dev_ms, wall_ms = [], [], [], []
        for k in range(a.reps + 1):
            t = time.perf_counter()
            m = serve.job_mask(geom_seed=k, **spec)
            host_ms.append((time.perf_counter() - t) * 1e3)
            t = time.perf_counter()
            bits, fluid, (ts, td) = serve.job_mask_bits(geom_seed=k, **spec)
            wall_ms.append((time.perf_counter() - t) * 1e3)
            samp_ms.append(ts)
            dev_ms.append(td)
            assert fluid == int(m.sum())
        med = lambda v: float(np.median(v[1:]))   # the first repetition warms up
        print(json.dumps({"n": a.n, "porosity": a.porosity, "reps": a.reps, "fluid_fraction": fluid / a.n ** 2,
                          "host_mask_ms": med(host_ms), "sampler_ms": med(samp_ms), "device_raster_ms": med(dev_ms),
                          "device_call_ms": med(wall_ms)}))
        return
    if a.cmd == "copy":
        ms = nat.probe_copy(int(a.bytes), a.reps)
        print(json.dumps({"kernel": "k_probe_copy", "bytes_read": int(a.bytes) // 16 * 16,
                          "bytes_written": int(a.bytes) // 16 * 16, "avg_ms": ms,
                          "gbs": 2 * a.bytes / (ms * 1e-3) / 1e9}))
        return
    from sc_lbm_amd import workloads as wl
    case = wl.spinodal(a.n, G=1.2, seed=777)
    if a.porosity < 1.0:
        from sc_lbm_amd import serve
        m = serve.job_mask(nx=a.n, ny=a.n, steps=1, porosity=a.porosity, rmin=8, rmax=20, geom_seed=1)
        case["mask"] = m
        case["f"] = [np.where(np.tile(m.ravel(), 9) == 1, f, 0.0) for f in case["f"]]
    g = nat.Geometry(case["mask"])
    if a.tile:
        v = [int(t) for t in a.tile.split(",")]
        g.set_tile(v[0], v[1])
        if len(v) > 2:
            g.set_lookahead(v[2])
    else:
        g.autotune(2, nat.PRECISION[a.form])
    if a.fusion or a.band:
        g.set_step_fusion(a.fusion, a.band)
    fl = [nat.Field(g, case["tau"][k], nat.PRECISION[a.form], 1.0, g_fluid=case["g_fluid"][k])
          for k in range(2)]
    for k in range(2):
        fl[k].set_f(case["f"][k])
    nat.solid_surface_force(g, fl[0], fl[1])
    nat.step2(g, fl[0], fl[1], None, None, a.warmup)
    nat.synchronize()
    nat.timing(1)
    nat.step2(g, fl[0], fl[1], None, None, a.steps)
    avg, cnt = nat.timing(0)            # per time step
    spl, band = g.step_fusion(nat.PRECISION[a.form])
    print(json.dumps({"form": a.form, "n": a.n, "porosity": a.porosity, "steps": a.steps,
                      "kernel": {1: "k_step", 2: "k_step2d" if a.form == "fp64" else "k_step2x", 3: "k_stepk"}[spl],
                      "steps_per_launch": spl, "band_rows": band,
                      "kernel_avg_ms": avg * a.steps / cnt if cnt else None, "launches": cnt,
                      "tile": list(g.tile(2, nat.PRECISION[a.form])),
                      "mlups": a.n * a.n / (avg * 1e-3) / 1e6 if avg else None}))


if __name__ == "__main__":
    main()
