#!/usr/bin/env python3
"""Benchmark of the depth-averaged MPM time step on one MI355X.

Workload (default): the slope landslide of amsc_amd.scenarios
.slope_landslide(): 1000x1000 cells (hx = 1 m), Z = 0.3 x + 8 seeded Gaussian
bumps, an elliptical release of 2.0M particles at 4 per cell, Voellmy
(phi = 15 deg, xi = 500) + Bingham (mu = 0.5, tauy = 200), rho = 2000,
adaptive dt with CFL = 0.4, merge/split every 10 steps; FP64 throughout.
A "step" is one full time step (dt, merge/split when due, P2G + nodal
update, G2P/advection/height/rheology, re-binning) with all state resident
in HBM.

Prints one JSON line: particle-steps/s over the timed steps, the HBM
roofline of the dominant kernel (algorithmic bytes per launch over its
average HIP-event duration, in a second pass with per-kernel events), and
the FP64 CPU oracle timed on a bounded sample of the same workload on the
host (1 thread).

    python bench.py [--steps K] [--warmup W] [--config slope2m|large4m|dambreak]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s (1 GPU, FP64 plain vs mixed); ensemble scen/s + p50/p99"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes (FP64): fields each kernel must read or write once.
#   p2g_nodal: reads x y Mp vpx vpy Vp hp H Fric_px Fric_py F11 F12 F22 per
#              particle (104 B); per active node reads Z, writes vL (24 B).
#   particle : reads x y hp Vp H (40 B), writes x y vpx vpy hp F11 F12 F22
#              Fric_px Fric_py (80 B) per particle; per active node reads vL
#              (16 B).
#   whole step (state counted once per step): 184 B per particle + 120 B per
#              active node.
ALGO = {"p2g_nodal": (104, 24), "particle": (120, 16)}
STEP_ALGO = (184, 120)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from amsc_amd import scenarios
    if name == "slope2m":
        d = scenarios.slope_landslide(max_steps=0)
        desc = ("slope landslide, 1000x1000 cells hx=1, Z=0.3x+8 Gaussian bumps, "
                "elliptical release, Voellmy+Bingham, adaptive CFL 0.4, "
                "merge/split every 10 steps")
    elif name == "large4m":
        d = scenarios.large_fractal()
        desc = ("large fractal slope, 2048x2048 cells hx=0.5, 25 deg mean slope, "
                "polygonal release, Voellmy+Bingham, dt=2e-3, merge/split on")
    elif name == "dambreak":
        d = scenarios.dambreak_flat(100000)
        desc = "flat dam-break, 100x100 cells hx=0.1, 1600 particles, Bingham, dt=1e-3"
    else:
        raise SystemExit(f"unknown config {name}")
    return d, desc


def cpu_baseline(d, steps, warm=2):
    from amsc_amd.data import params_from_data, particles_from_data
    from oracle import oracle
    prm = params_from_data(d)
    prm["max_steps"] = 0
    o = oracle.OracleSim(prm, particles_from_data(d), d["Z"])
    for _ in range(warm):
        o.step()
    ps = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        n = o.np
        if not o.step():
            break
        ps += n
    dt = time.perf_counter() - t0
    return ps / dt, dt, o.np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="slope2m")
    ap.add_argument("--roofline-steps", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=14)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE/"
                         "WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.arange(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch_read": n * 8,
                          "copy_bytes_per_launch_write": n * 8, "launches": 5}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, "
              f"WORLD_SIZE {world}): the MPM step has no multi-GPU path here; "
              "see DESIGN.md for how it would shard.", file=sys.stderr)
        sys.exit(2)

    from amsc_amd.simulation import Simulation

    d, desc = workload(args.config)
    np0 = len(d["x"])
    total = args.warmup + args.steps + args.roofline_steps + 5
    log(f"[bench] {args.config}: {np0} particles, grid {d['Nex']}x{d['Ney']}")
    sim = Simulation.from_data(d, capacity=3 * np0, max_steps=total)
    sim.step(args.warmup)
    log(f"[bench] warmup {args.warmup} steps done")

    ps0 = sim.particle_steps
    t0 = time.perf_counter()
    done = sim.step(args.steps)
    t1 = time.perf_counter()
    ps = sim.particle_steps - ps0
    elapsed = t1 - t0
    value = ps / elapsed
    ms_per_step = 1e3 * elapsed / max(done, 1)
    log(f"[bench] timed {done} steps: {ms_per_step:.4f} ms/step, {value:.4g} particle-steps/s")

    # second pass: per-kernel HIP events on the solver's stream
    n_act = sim.active_nodes()
    npart = sim.num_particles
    sim.set_timing(True)
    sim.reset_kernel_times()
    done_r = sim.step(args.roofline_steps)
    kt = sim.kernel_times()
    sim.set_timing(False)
    n_act2 = sim.active_nodes()
    npart2 = sim.num_particles
    n_act_m = 0.5 * (n_act + n_act2)
    np_m = 0.5 * (npart + npart2)
    dom = max(ALGO, key=lambda k: kt.get(k, (0.0, 1))[0])
    tot_ms, cnt = kt[dom]
    avg_ms = tot_ms / max(cnt, 1)
    bpp, bpn = ALGO[dom]
    algo_bytes = bpp * np_m + bpn * n_act_m
    achieved = algo_bytes / (avg_ms * 1e-3) / 1e9
    step_bytes = STEP_ALGO[0] * np_m + STEP_ALGO[1] * n_act_m
    step_gbs = step_bytes / (ms_per_step * 1e-3) / 1e9
    kernels = {k: round(v[0] / max(v[1], 1), 5) for k, v in kt.items()}
    log(f"[bench] kernels (avg ms): {kernels}")

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}_{args.config}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] CPU oracle baseline: {args.cpu_steps} steps")
        v, secs, npc = cpu_baseline(d, args.cpu_steps)
        cpu = {"value": round(v, 1), "unit": "particle-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"{args.config} workload ({np0} particles), {args.cpu_steps} "
                         f"steps after 2 warm-up steps, FP64 oracle restating "
                         f"optim_par, 1 thread, {secs:.1f} s, host "
                         f"{platform.processor() or platform.machine()}"}
        log(f"[bench] CPU: {v:.4g} particle-steps/s")

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": done, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "config": args.config, "particles": np0,
                   "grid": [int(d["Nex"]), int(d["Ney"])], "precision": "fp64",
                   "active_nodes": int(n_act_m)},
        "roofline": {"bound": "hbm", "kernel": dom,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "avg_launch_ms": round(avg_ms, 5),
                     "algorithmic_bytes_per_launch": int(algo_bytes)},
        "step_roofline": {"algorithmic_bytes_per_step": int(step_bytes),
                          "achieved": round(step_gbs, 1),
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels_avg_ms": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    sim.close()


if __name__ == "__main__":
    main()
