"""Benchmark: IISPH particle-steps/s on the 1,048,576-particle dam break, one GPU.

A step is one IISPH sub-step of the whole scene at fixed dt (keys, sort, cells,
neighbour lists, densities, the pressure solve with its per-iteration residual
read-back, pressure application and integration). Inputs are resident on the
device before the timed region; the timed region is exactly K sub-steps
bracketed by device synchronisation.

Prints one JSON line with the metric, a roofline entry for the dominant kernel
(algorithmic bytes per launch over its average launch time, measured with HIP
events on the context's stream) and the CPU baseline (the oracle, a C++
restatement of the reference's CPU path, timed on this host's cores on a
bounded sample of the same workload).

  python bench.py [--steps K] [--warmup W] [--method iisph|dfsph_p|pcisph]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per particle for one launch of each kernel class
# (FP32 SoA fields, each per-particle field read or written once per sweep,
# neighbour gathers assumed cache-resident, neighbour lists not counted).
ALGO_BYTES = {
    "keys": 8 + 8 + 4,            # x read, key write, index write
    "radix_sort": 3 * 2 * 12,     # 3 passes x (read + write) of key + index
    "cells": 12 + 16 + 16 + 8,    # sorted pair read, head/scan, gather x,v in, out
    "build": 8 + 8 + 8 + 4 + 8,   # x, v in; v out; rho; d_ii (IISPH build)
    "iisph_aii": 4 + 8 + 8 + 4 + 4 + 4,   # rho, d_ii, v in; a_ii, rho*, p out
    "iisph_sdp": 4 + 4 + 8,       # rho, p in; sum d_ij p_j out
    "iisph_beta_p": 4 + 4 + 8 + 8 + 4 + 4 + 4 + 4,  # rho,p,s,dii,aii,rho* in; p' out (+residual)
    "apply_integrate": 4 + 4 + 8 + 8 + 4 + 8 + 8,   # rho,p,x,v in; orig; x,v out
    "apply": 4 + 4 + 8 + 8,
    "dfsph_div": 8 + 4 + 4 + 4,
    "dfsph_dens": 8 + 4 + 4 + 4 + 4,
    "finalize": 8 + 8 + 4 + 8 + 8,
    "pci_init": 8 + 8 + 8 + 8 + 8 + 4,
    "pci_density": 8 + 4 + 4,
    "pci_force": 8 + 4 + 4 + 8 + 8 + 8 + 8,
    "pci_finalize": 8 + 8 + 4 + 8 + 8,
    "key_minmax": 8,
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(scene, cfg, method: str, gpu_state_after, budget_s: float = 12.0) -> dict:
    """Time the oracle (C++ restatement of the reference CPU path, OpenMP static
    chunks of 512 over host cores) on the first sub-steps of the same scene."""
    import oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_scene
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    o = oracle_scene(scene, cfg)
    o.frame_init(method)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.sub_step(method, steps * scene.dt, scene.dt)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 20:
            break
    n = len(scene.positions)
    parity = None
    if gpu_state_after is not None and steps in gpu_state_after:
        pos, vel = o.state_f32()
        gp, gv = gpu_state_after[steps]
        parity = bool(np.array_equal(pos.view(np.uint32), gp.view(np.uint32)) and
                      np.array_equal(vel.view(np.uint32), gv.view(np.uint32)))
    return {
        "value": n * steps / el,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{method} {n}-particle dam break, first {steps} sub-steps from the "
                  f"initial state ({el:.1f} s); GPU state after the same sub-steps "
                  f"bit-identical: {parity}",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--method", default="iisph", choices=["iisph", "dfsph_p", "pcisph"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are "
                                   "not part of it (see DESIGN.md for how the path would "
                                   "shard)", "n_gpus": args.gpus}))
        return 2

    from emerald_amd import lib, scenes
    from emerald_amd import sph2d_box as sb
    if lib().emsph_device_count() < 1:
        raise SystemExit("bench: no HIP device")

    method = args.method
    log("building scene (1024 x 1024 lattice dam break)")
    scene = scenes.dam_break_1m()
    cfg = scenes.config_for(scene, method)
    n = len(scene.positions)
    ctx = sb.Context(cfg, sb.Solid_state(scene.solids), n)
    ctx.set_state(sb.State(scene.positions, scene.velocities))
    ctx.frame_init(method)

    # Parity sample: the GPU state after the first k sub-steps, for the CPU leg.
    gpu_after = {}
    if not args.no_cpu_baseline:
        for k in range(1, 21):
            ctx.sub_step(method, (k - 1) * scene.dt, scene.dt)
            st = ctx.get_state()
            gpu_after[k] = (st.positions, st.velocities)
        ctx.set_state(sb.State(scene.positions, scene.velocities))
        ctx.frame_init(method)

    log(f"warmup {args.warmup} sub-steps")
    ctx.run_sub_steps(method, scene.dt, args.warmup)

    log(f"timing {args.steps} sub-steps")
    t0 = time.perf_counter()
    dev_ms, st = ctx.run_sub_steps(method, scene.dt, args.steps)
    wall = time.perf_counter() - t0
    value = n * args.steps / wall

    # Per-kernel timing pass (HIP events around each launch on the context stream).
    ctx.set_kernel_timing(True)
    prof_steps = 10
    ctx.run_sub_steps(method, scene.dt, prof_steps)
    timing = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    dom = max(timing.items(), key=lambda kv: kv[1][0])
    dom_name, (dom_ms, dom_calls) = dom
    avg_s = dom_ms / dom_calls / 1e3
    per = ALGO_BYTES.get(dom_name)
    achieved = (per * n / avg_s / 1e9) if per else None
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get(method, {}).get(dom_name)
    roofline = {
        "bound": "hbm",
        "kernel": dom_name,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
        "traffic": traffic,
        "avg_launch_us": avg_s * 1e6,
        "algorithmic_bytes_per_particle": per,
    }
    kernels = {k: {"ms_per_step": v[0] / prof_steps, "calls_per_step": v[1] / prof_steps}
               for k, v in sorted(timing.items(), key=lambda kv: -kv[1][0])}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores)")
        cpu = cpu_baseline(scene, cfg, method, gpu_after, args.cpu_budget)

    out = {
        "metric": "IISPH particle-steps/s @1M (1 GPU)" if method == "iisph"
        else f"{method.upper()} particle-steps/s @1M (1 GPU)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "device_ms_per_step": dev_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{method} dam break, 1024x1024 lattice (1,048,576 fluid particles), "
                        f"spacing 1/2560, h = spacing, dt 1.5625e-5, hex walls",
            "particles": n,
            "solids": int(len(scene.solids)),
            "dt": scene.dt,
            "last_iterations": int(st.iterations),
            "sort_passes": int(st.sort_passes),
            "cells": int(st.cell_count),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
