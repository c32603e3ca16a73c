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
        lattice updates per second of the batch.
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
                      "kernel": {1: "k_step", 2: "k_step2x", 3: "k_stepk"}[spl], "steps_per_launch": spl, "band_rows": band,
                      "kernel_avg_ms": avg * a.steps / cnt if cnt else None, "launches": cnt,
                      "tile": list(g.tile(2, nat.PRECISION[a.form])),
                      "mlups": a.n * a.n / (avg * 1e-3) / 1e6 if avg else None}))


if __name__ == "__main__":
    main()
