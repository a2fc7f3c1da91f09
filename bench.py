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

The same line carries the other workloads as objects beside the headline:
"batched" (256 ragged scenes, one launch per stage, against per-scene
contexts), "served" (a Poisson request stream over HBM-budgeted pools, p50/p99
step latency), "search" (4,194,304 points, keys, sort, cells and lists alone)
and "small" (the 12,800-particle dam break, 1000 sub-steps, GPU against the
oracle on the host cores, final states compared bit for bit).

  python bench.py [--steps K] [--warmup W]
                  [--method iisph|iisph_ap|iisph_pseudo_ap|dfsph_p|pcisph]
                  [--no-cpu-baseline] [--no-batched] [--no-served] [--no-search]
                  [--no-small]
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

# Algorithmic HBM bytes per fluid particle for one launch of each timed kernel
# class, per solver. Conventions: FP32 SoA; every per-particle field read or
# written once per sweep (neighbour gathers hit cache after the Morton sort);
# the uniform fluid volume costs nothing; neighbour lists and solid-field reads
# are not counted. Neighbour build NB = keys 16 + sort 16 + cells 8 = 40.
# Per sub-step the tables sum to IISPH 160 + 68k (the first iteration's
# sum_j d_ij p_j is known to be zero: pressures start at zero), DFSPH_P
# 172 + 68k_v + 72k_rho, PCISPH 116 + 108k.
NB = {"keys": 16, "radix_sort": 16, "dense_sort": 16, "cells": 8, "step_begin": 0, "loop": 0}
ALGO_BYTES = {
    "iisph": {**NB,
              "build": 48,            # rho 12 + (v += g) 16 + d_ii 20
              "iisph_aii_beta": 84,   # (a_ii, rho*) 36 + (p = 0) 4 + first beta/p'/residual 44
              "iisph_sdp": 24,        # sum_j d_ij p_j, iterations 2..k
              "iisph_beta_p": 44,     # beta, p', residual, iterations 2..k
              "apply_integrate": 56},  # apply 32 + x 24
    "dfsph_p": {**NB,
                "build": 28,          # rho 12 + alpha 16
                "apply": 32,          # warm start / per-iteration kappa application
                "apply_gravity": 48,  # warm start 32 + g 16
                "dfsph_div": 36,      # divergence iteration 68 - apply 32
                "dfsph_dens": 40,     # density iteration 72 - apply 32
                "finalize": 24},      # x 24
    # IISPH_AP / IISPH_PSEUDO_AP (the same sweeps, by the same conventions):
    # build rho 12 + (v += g) 16; diagonal sweep (v, rho) 12 + (a_ii, rho*) 8
    # + (p' of the first iteration) 4; per further iteration accelerations
    # (p, rho) 8 + a 8 and the update (a, a_ii, rho*, p) 20 + p' 4; final
    # (p, rho) 8 + a 8 + (x, v) 16 + (x, v) 16. Sum 100 + 40k.
    "iisph_ap": {**NB, "build": 28, "ap_diag": 24, "ap_accel": 16, "ap_iterate": 24,
                 "ap_final": 48},
    "iisph_pseudo_ap": {**NB, "build": 28, "ap_diag": 24, "ap_accel": 16, "ap_iterate": 24,
                        "ap_final": 48},
    "pcisph": {**NB,
               "build": 0,            # index lists only (not counted)
               "pci_init": 28,        # v_ext 16 + init 12
               "pci_density": 48,     # x* 24 + rho 12 + p 12
               "pci_force": 60,       # F 32 + v* 24 + max error 4
               "pci_finalize": 48},
}


def step_bytes(method: str, k: float, k_div: float) -> float:
    """Algorithmic bytes per particle-step at the measured mean iteration counts."""
    if method == "iisph":
        return 160 + 68 * k
    if method in ("iisph_ap", "iisph_pseudo_ap"):
        return 100 + 40 * k
    if method == "dfsph_p":
        return 172 + 68 * k_div + 72 * k
    return 116 + 108 * k


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


def batched_form(method: str, steps: int, warmup: int, count: int = 256) -> dict:
    """256 dam breaks of uneven size (log-uniform 2k-200k particles, aspect
    0.25-2, per-scene CFL 0.4 dt) stepped as one batch (one launch per sub-step
    stage for all scenes) and as 256 plain per-scene contexts, over the same
    sub-steps from the same states; every scene's state must come out
    bit-identical in both forms."""
    from emerald_amd import scenes
    from emerald_amd import sph2d_box as sb
    log(f"batched form: generating {count} scenes")
    scs = scenes.batch_dam_breaks(count)
    cfgs = [scenes.config_for(sc, method) for sc in scs]
    n_total = sum(len(sc.positions) for sc in scs)
    solids = [sb.Solid_state(sc.solids) for sc in scs]
    states = [sb.State(sc.positions, sc.velocities) for sc in scs]

    log(f"batched form: one batch of {count} scenes, {n_total} particles")
    batch = sb.Batch(cfgs, solids, states)
    batch.set_cfl(0.4)
    batch.frame_init(method)
    batch.run_sub_steps(method, warmup)
    log("batched form: autotune")
    chosen, cand_ms = batch.autotune(method, 2)
    t0 = time.perf_counter()
    batch.run_sub_steps(method, steps)
    t_batch = time.perf_counter() - t0
    last = batch.sub_step(method)  # one more, for the per-scene statistics
    iters = [st.iterations for st in last]

    log(f"batched form: {count} plain per-scene contexts")
    ctxs = []
    for sc, cfg, sol, st in zip(scs, cfgs, solids, states):
        c = sb.Context(cfg, sol, len(sc.positions))
        c.set_cfl(0.4)
        c.set_state(st)
        c.frame_init(method)
        c.run_sub_steps(method, cfg.seconds_per_sub_step, warmup)
        ctxs.append(c)
    t0 = time.perf_counter()
    for c, cfg in zip(ctxs, cfgs):
        c.run_sub_steps(method, cfg.seconds_per_sub_step, steps)
    t_plain = time.perf_counter() - t0
    same = True
    for i, (c, cfg) in enumerate(zip(ctxs, cfgs)):
        c.run_sub_steps(method, cfg.seconds_per_sub_step, 1)
        a, b = c.get_state(), batch.get_state(i)
        same = same and np.array_equal(a.positions.view(np.uint32), b.positions.view(np.uint32)) \
            and np.array_equal(a.velocities.view(np.uint32), b.velocities.view(np.uint32))
        c.close()
    # Per-stage GPU time of the batched sub-step, from the engine's HIP events
    # (timed launches run outside graphs), after the comparison above.
    batch.set_kernel_timing(True)
    batch.run_sub_steps(method, 5)
    stages = batch.kernel_timing()
    batch.set_kernel_timing(False)
    batch.close()
    stage_ms = {k: round(v[0] / 5, 4) for k, v in sorted(stages.items(), key=lambda kv: -kv[1][0])}
    v_batch = n_total * steps / t_batch
    v_plain = n_total * steps / t_plain
    return {
        "scenes": count, "particles": n_total, "sub_steps": steps, "method": method,
        "value": v_batch, "unit": "particle-steps/s", "ms_per_sub_step": t_batch * 1e3 / steps,
        "plain_value": v_plain, "plain_ms_per_sub_step": t_plain * 1e3 / steps,
        "speedup_vs_plain": v_batch / v_plain, "bit_identical_to_plain": bool(same),
        "iterations_min_mean_max": [min(iters), float(np.mean(iters)), max(iters)],
        "dt": "per scene, CFL 0.4 on the device",
        "tuning": {"chosen": chosen, "candidate_ms_per_2_sub_steps": cand_ms},
        "stage_ms_per_sub_step": stage_ms,
    }


def search_form(runs: int = 5, n: int = 4_194_304) -> dict:
    """Z-index keys, sort, cells and 64-capped neighbour lists alone (no
    dynamics) on 4,194,304 points, half uniform and half from 8 Gaussian
    clusters (sigma 0.02), cell 1/2048 (h = 1/4096). Algorithmic bytes per point
    40 + 1 + 4 n (key build 40, count 1, u32 indices; the lists are the
    product)."""
    from emerald_amd import scenes
    from emerald_amd.search import NeighborSearch
    log(f"search form: {n} points")
    xy = scenes.search_points(n, seed=3)
    s = NeighborSearch(n)
    s.set_points(xy)
    s.run(1.0 / 2048.0)  # first run: list capacity and pass count
    ms = [s.run(1.0 / 2048.0) for _ in range(runs)]
    r = s.results(lists=True)
    s.close()
    counts = r["counts"].astype(np.float64)
    nbar = float(counts.mean())
    best = min(ms)
    per_point = 40.0 + 1.0 + 4.0 * nbar
    return {
        "workload": "4,194,304 points (half U[0,1)^2, half 8 Gaussian clusters sigma 0.02), "
                    "cell 1/2048, 64-capped lists",
        "points": n,
        "ms_per_search": {"min": best, "median": float(np.median(ms))},
        "points_per_s": n / (best / 1e3),
        "mean_neighbors": nbar,
        "max_neighbors": int(counts.max()),
        "capped_lists": int((counts >= 64).sum()),
        "algorithmic_bytes_per_point": per_point,
        "achieved_GBps": per_point * n / (best / 1e3) / 1e9,
        "cells": int(len(r["block_begin"])) if "block_begin" in r else None,
    }


def small_form(steps: int = 1000, cpu_budget: float = 20.0) -> dict:
    """The reference's own CPU-runnable case: the 12,800-particle IISPH dam
    break (80 x 160 lattice, spacing 0.005, dt 2e-4), 1000 sub-steps on the GPU
    and the same sub-steps in the oracle on the host cores, with the final
    states compared bit for bit."""
    import oracle
    from emerald_amd import scenes
    from emerald_amd import sph2d_box as sb
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_scene
    log(f"small form: 12,800-particle IISPH dam break, {steps} sub-steps")
    scene = scenes.dam_break_12800()
    cfg = scenes.config_for(scene, "iisph")
    n = len(scene.positions)
    ctx = sb.Context(cfg, sb.Solid_state(scene.solids), n)
    ctx.set_state(sb.State(scene.positions, scene.velocities))
    ctx.frame_init("iisph")
    ctx.iteration_totals(reset=True)
    t0 = time.perf_counter()
    dev_ms, _ = ctx.run_sub_steps("iisph", scene.dt, steps)
    wall = time.perf_counter() - t0
    _, iters, _, _ = ctx.iteration_totals(reset=True)
    gst = ctx.get_state()
    ctx.close()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    o = oracle_scene(scene, cfg)
    o.frame_init("iisph")
    done = 0
    t1 = time.perf_counter()
    while done < steps:
        o.sub_step("iisph", done * scene.dt, scene.dt)
        done += 1
        if time.perf_counter() - t1 > cpu_budget:
            break
    cwall = time.perf_counter() - t1
    same = None
    if done == steps:
        pos, vel = o.state_f32()
        same = bool(np.array_equal(pos.view(np.uint32), gst.positions.view(np.uint32)) and
                    np.array_equal(vel.view(np.uint32), gst.velocities.view(np.uint32)))
    return {
        "workload": "iisph dam break, 80x160 lattice (12,800 particles), spacing 0.005, "
                    "dt 2e-4, hex walls",
        "sub_steps": steps,
        "value": n * steps / wall,
        "unit": "particle-steps/s",
        "ms_per_sub_step": wall * 1e3 / steps,
        "device_ms_per_sub_step": dev_ms / steps,
        "mean_iterations": iters / steps,
        "cpu_value": n * done / cwall,
        "cpu_cores": threads,
        "cpu_sub_steps": done,
        "speedup_vs_cpu": (n * steps / wall) / (n * done / cwall),
        "final_state_bit_identical_to_oracle": same,
    }


def served_form(budget_gb: float, window: float, rate: float = 200.0,
                plain_steps: int = 2) -> dict:
    """Poisson stream of dam-break requests (200/s, 5k-500k particles, IISPH or
    DFSPH_P, 100-2000 steps) served from two device-resident pools (one per
    solver) with continuous batching under an HBM budget. After the window,
    the scenes still live are also stepped as plain per-scene contexts (one at
    a time) for the throughput comparison."""
    from emerald_amd import serve
    from emerald_amd import sph2d_box as sb
    reqs = serve.poisson_requests(rate, window + 5.0)
    res = serve.serve(reqs, budget_gb * 1e9, window, log=log)
    loops = res["_loops"]
    # Plain per-scene contexts on the pools' final live sets.
    log("served form: plain per-scene comparison on the final live sets")
    n_steps = 0
    t_plain = 0.0
    scenes_done = 0
    for m, lp in loops.items():
        live = list(lp.live.items())
        states = {sid: lp.pool.get_state(sid) for sid, _ in live}
        lp.pool.close()
        for sid, v in live:
            req = v[0]
            sc, cfg = serve.request_scene(req)
            c = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
            c.set_cfl(0.4)
            c.set_state(states[sid])
            c.frame_init(m)
            t0 = time.perf_counter()
            c.run_sub_steps(m, cfg.seconds_per_sub_step, plain_steps)
            t_plain += time.perf_counter() - t0
            n_steps += len(sc.positions) * plain_steps
            scenes_done += 1
            c.close()
    serve.close(res)
    res["plain_value_final_live_set"] = n_steps / t_plain if t_plain > 0 else None
    res["plain_scenes"] = scenes_done
    res["speedup_vs_plain"] = (res["value"] / res["plain_value_final_live_set"]
                               if res["plain_value_final_live_set"] else None)
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--method", default="iisph",
                    choices=["iisph", "iisph_ap", "iisph_pseudo_ap", "dfsph_p", "pcisph"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--batched-only", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-only", action="store_true")
    ap.add_argument("--served-window", type=float, default=10.0)
    ap.add_argument("--served-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-search", action="store_true")
    ap.add_argument("--search-only", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only; --gpus above 1 is "
                                   "not supported", "n_gpus": args.gpus}))
        return 2

    from emerald_amd import lib, scenes
    from emerald_amd import sph2d_box as sb
    if lib().emsph_device_count() < 1:
        raise SystemExit("bench: no HIP device")

    method = args.method
    if args.batched_only:
        print(json.dumps(batched_form(method, args.batched_steps, 2)), flush=True)
        return 0
    if args.served_only:
        print(json.dumps(served_form(args.served_budget_gb, args.served_window)), flush=True)
        return 0
    if args.search_only:
        print(json.dumps(search_form()), flush=True)
        return 0
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
    log("autotune")
    chosen, cand_ms = ctx.autotune(method, scene.dt, 5)

    log(f"timing {args.steps} sub-steps")
    ctx.iteration_totals(reset=True)
    t0 = time.perf_counter()
    dev_ms, st = ctx.run_sub_steps(method, scene.dt, args.steps)
    wall = time.perf_counter() - t0
    value = n * args.steps / wall
    _, iters, div_iters, _ = ctx.iteration_totals(reset=True)
    k_mean = iters / args.steps
    kdiv_mean = div_iters / args.steps
    bytes_step = step_bytes(method, k_mean, kdiv_mean)

    # Per-kernel timing pass over the same number of sub-steps: HIP events
    # recorded around each launch on the context's stream, resolved afterwards.
    log(f"kernel timing pass, {args.steps} sub-steps")
    ctx.set_kernel_timing(True)
    prof_steps = args.steps
    ctx.run_sub_steps(method, scene.dt, prof_steps)
    timing = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    table = ALGO_BYTES[method]
    dom = max(timing.items(), key=lambda kv: kv[1][0])
    dom_name, (dom_ms, dom_calls) = dom
    avg_s = dom_ms / dom_calls / 1e3
    per = table.get(dom_name)
    achieved = (per * n / avg_s / 1e9) if per else None
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get(method, {}).get(dom_name)
    valu = None
    vpath = os.path.join(ROOT, "profiles", "valu.json")
    if os.path.exists(vpath):
        with open(vpath) as f:
            valu = json.load(f).get(method, {}).get(dom_name)
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
        "particles_per_launch": n,
        # Measured share of the SIMDs' cycles taken by VALU issue (rocprof
        # PMC): the neighbour build is limited by it, not by HBM bytes.
        "valu_issue_frac": valu,
    }
    kernels = {}
    for k, v in sorted(timing.items(), key=lambda kv: -kv[1][0]):
        e = {"ms_per_step": v[0] / prof_steps, "calls_per_step": v[1] / prof_steps,
             "avg_launch_us": v[0] / v[1] * 1e3}
        if table.get(k):
            e["algorithmic_GBps"] = table[k] * n / (v[0] / v[1] / 1e3) / 1e9
        kernels[k] = e
    whole_step = {
        "algorithmic_bytes_per_particle_step": bytes_step,
        "mean_iterations": k_mean,
        "mean_divergence_iterations": kdiv_mean,
        "achieved_GBps": bytes_step * value / 1e9,
        "frac_of_hbm_peak": bytes_step * value / 1e9 / HBM_PEAK_GBS,
    }

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores)")
        cpu = cpu_baseline(scene, cfg, method, gpu_after, args.cpu_budget)

    batched = None
    if not args.no_batched:
        ctx.close()
        batched = batched_form(method, args.batched_steps, 2)
    served = None
    if not args.no_served:
        ctx.close()
        served = served_form(args.served_budget_gb, args.served_window)
    search = None
    if not args.no_search:
        ctx.close()
        search = search_form()
    small = None
    if not args.no_small:
        ctx.close()
        small = small_form()

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
            "max_neighbors": int(st.max_neighbors),
        },
        "roofline": roofline,
        "whole_step": whole_step,
        "tuning": {"chosen": chosen, "candidate_ms_per_5_sub_steps": cand_ms,
                   "candidates": "neighbour batch 2/4/8 x XCD remap on/off with graphs, "
                                 "then the best without graphs"},
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "search": search,
        "small": small,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
