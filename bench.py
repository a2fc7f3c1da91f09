"""Benchmark: IISPH particle-steps/s on the 1,048,576-particle dam break, one GPU.

A step is one IISPH sub-step of the whole scene at fixed dt (keys, sort, cells,
neighbour lists, densities, the pressure solve with its per-iteration residual
read-back, pressure application and integration). Inputs are resident on the
device before the timed region; the timed region is exactly K sub-steps
bracketed by device synchronisation.

Two windows of the same run are timed:
  * collapse (the headline): K sub-steps from the state after ADVANCE sub-steps
    of the dam break (8,000 sub-steps = 0.125 s of simulated time by default:
    the column is collapsing and the front runs along the floor);
  * rest: K sub-steps right after the warm-up, the lattice still near rest
    (the easiest phase: the solver's minimum of 3 iterations).

Prints one JSON line with the metric, a roofline entry for the dominant kernel
(algorithmic bytes per launch over its average launch time, measured with HIP
events on the context's stream) and the CPU baseline: the oracle, a C++
restatement of the reference's CPU path, timed on this host's cores from the
same collapse state (one warm-up sub-step, the median of 3 runs, and one
sub-step on one thread), its state after those sub-steps compared bit for bit
with the GPU's.

The same line carries the other workloads as objects beside the headline:
"solvers" (PCISPH and DFSPH_P at 1M, GPU and CPU from their own collapse
states), "batched" (256 ragged scenes, one launch per stage, against per-scene
contexts), "served" (a Poisson request stream over HBM-budgeted pools, p50/p99
step latency, FP64 error of sampled results), "search" (4,194,304 points, keys,
sort, cells and lists alone, GPU and CPU) and "small" (the 12,800-particle dam
break, 1000 sub-steps, GPU against the oracle on the host cores, final states
compared bit for bit).

  python bench.py [--steps K] [--warmup W] [--advance A]
                  [--method iisph|iisph_ap|iisph_pseudo_ap|dfsph_p|pcisph]
                  [--no-cpu-baseline] [--no-solvers] [--no-batched] [--no-served]
                  [--no-search] [--no-small]
                  [--batched-only | --served-only | --search-only | --small-only]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Skin candidate lists in the 1M solver windows (--skin-lists on|off; on is
# the library's default for contexts).
SKIN_LISTS = True
# Fixed (neighbour batch, XCD remap) of the 1M solver windows (--tuning), or
# None for the autotuner.
FIXED_TUNING = None

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


def host_cpu() -> dict:
    """The host's CPU: model, the CPUs this process may run on (affinity, the
    cgroup quota) and the threads the CPU legs use: all of them, or the
    OMP_NUM_THREADS share when the environment sets one."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    env = os.environ.get("OMP_NUM_THREADS")
    threads = affinity
    if quota:
        threads = min(threads, max(1, int(math.floor(quota))))
    if env and env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    return {"model": model, "affinity_cpus": affinity, "cgroup_cpus": quota,
            "omp_num_threads_env": env, "threads": threads}


def _bits_equal(a_pos, a_vel, b_pos, b_vel) -> bool:
    return bool(np.array_equal(a_pos.view(np.uint32), b_pos.view(np.uint32)) and
                np.array_equal(a_vel.view(np.uint32), b_vel.view(np.uint32)))


def gpu_states_after(ctx, method, dt, state, counts):
    """The GPU state after each of counts sub-steps from state (a fresh frame)."""
    from emerald_amd import sph2d_box as sb
    ctx.set_state(sb.State(state.positions, state.velocities))
    ctx.frame_init(method)
    out = {}
    done = 0
    for k in sorted(counts):
        ctx.run_sub_steps(method, dt, k - done)
        done = k
        st = ctx.get_state()
        out[k] = (st.positions, st.velocities)
    return out


def cpu_leg(scene, cfg, method, state, gpu_after, cpu, per_run: int = 7, runs: int = 3) -> dict:
    """The oracle from the given state: one warm-up sub-step, then runs x
    per_run sub-steps on cpu["threads"] threads (median of the runs), then one
    sub-step on one thread. The oracle's state after 1 + runs x per_run and
    after one more sub-step is compared bit for bit with the GPU's."""
    import oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_cfg
    threads = cpu["threads"]
    oracle.set_threads(threads)
    o = oracle.Scene(oracle_cfg(cfg), scene.solids, state.positions, state.velocities)
    o.frame_init(method)
    n = len(state.positions)
    dt = scene.dt
    t0 = time.perf_counter()
    o.sub_step(method, 0.0, dt)
    warm = time.perf_counter() - t0
    times = []
    done = 1
    for _ in range(runs):
        t0 = time.perf_counter()
        for _ in range(per_run):
            o.sub_step(method, done * dt, dt)
            done += 1
        times.append(time.perf_counter() - t0)
    pos, vel = o.state_f32()
    parity = None
    if done in gpu_after:
        parity = _bits_equal(pos, vel, *gpu_after[done])
    oracle.set_threads(1)
    t0 = time.perf_counter()
    o.sub_step(method, done * dt, dt)
    one = time.perf_counter() - t0
    done += 1
    oracle.set_threads(threads)
    if parity is not None and done in gpu_after:
        pos, vel = o.state_f32()
        parity = parity and _bits_equal(pos, vel, *gpu_after[done])
    med = statistics.median(times)
    del o
    return {
        "value": n * per_run / med,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "one_thread_value": n / one,
        "runs_s": [round(t, 4) for t in times],
        "warmup_s": round(warm, 4),
        "sample": f"{method} {n}-particle dam break from the GPU's state at the start of the "
                  f"timed window: 1 warm-up sub-step, median of {runs} runs of {per_run} "
                  f"sub-steps on {threads} threads, then 1 sub-step on 1 thread; oracle "
                  f"state after {done - 1} and {done} sub-steps bit-identical to the GPU's: "
                  f"{parity}",
        "bit_identical_to_gpu": parity,
        "host": cpu,
    }


def solver_window(method: str, steps: int, warmup: int, advance: int, cpu: dict | None,
                  autotune: bool, kernel_timing: bool, rest_window: bool,
                  save_state: str | None = None, from_state: str | None = None) -> dict:
    """One solver at 1M: warm-up, optionally the rest window, the advance to the
    collapse state, the timed collapse window, a kernel-timing pass, and the
    CPU leg from the same collapse state. save_state writes the state at the
    start of the timed window (an .npz: positions, velocities, sub-step);
    from_state starts from such a file instead of the initial lattice (the
    warm-up sub-steps then run from it, and there is no rest window or
    advance), so that a profiled run can cover the collapse window alone."""
    from emerald_amd import scenes
    from emerald_amd import sph2d_box as sb
    log(f"{method}: building the 1024 x 1024 lattice dam break")
    scene = scenes.dam_break_1m()
    cfg = scenes.config_for(scene, method)
    n = len(scene.positions)
    ctx = sb.Context(cfg, sb.Solid_state(scene.solids), n)
    ctx.set_skin_lists(SKIN_LISTS)
    done = 0
    if from_state:
        with np.load(from_state) as f:
            pos, vel, done = f["positions"], f["velocities"], int(f["sub_step"])
        if pos.shape != scene.positions.shape or vel.shape != scene.velocities.shape:
            raise SystemExit(f"bench: {from_state} does not hold a state of this scene")
        log(f"{method}: starting from {from_state} (sub-step {done})")
        ctx.set_state(sb.State(pos, vel))
        rest_window = False
        advance = 0
    else:
        ctx.set_state(sb.State(scene.positions, scene.velocities))
    ctx.frame_init(method)
    out = {"method": method, "particles": n, "solids": int(len(scene.solids)), "dt": scene.dt}
    log(f"{method}: warm-up {warmup} sub-steps")
    ctx.run_sub_steps(method, scene.dt, warmup)
    done += warmup
    if autotune and FIXED_TUNING:
        nb, xcd = FIXED_TUNING
        ctx.set_tuning(graphs=False, neighbor_batch=nb, xcd_remap=bool(xcd))
        out["tuning"] = {"chosen": ctx.tuning(), "fixed": True}
    elif autotune:
        log(f"{method}: autotune")
        chosen, cand_ms = ctx.autotune(method, scene.dt, 5)
        out["tuning"] = {"chosen": chosen, "candidate_ms_per_5_sub_steps": cand_ms,
                         "candidates": "neighbour batch 2/4/8 x XCD remap on/off with graphs, "
                                       "then the best without graphs"}
    if rest_window:
        ctx.iteration_totals(reset=True)
        t0 = time.perf_counter()
        dev_ms, _ = ctx.run_sub_steps(method, scene.dt, steps)
        wall = time.perf_counter() - t0
        _, it, dit, _ = ctx.iteration_totals(reset=True)
        out["rest"] = {"sub_steps": [done, done + steps], "value": n * steps / wall,
                       "ms_per_step": wall * 1e3 / steps, "device_ms_per_step": dev_ms / steps,
                       "mean_iterations": it / steps, "mean_divergence_iterations": dit / steps}
        done += steps
    log(f"{method}: advancing to sub-step {advance}")
    if advance > done:
        ctx.run_sub_steps(method, scene.dt, advance - done)
        done = advance
    start = ctx.get_state()
    if save_state:
        np.savez(save_state, positions=start.positions, velocities=start.velocities,
                 sub_step=np.int64(done))
        log(f"{method}: state at sub-step {done} saved to {save_state}")
    ctx.frame_init(method)
    log(f"{method}: timing {steps} sub-steps from sub-step {done}")
    ctx.iteration_totals(reset=True)
    skin0 = ctx.skin_lists()
    t0 = time.perf_counter()
    dev_ms, st = ctx.run_sub_steps(method, scene.dt, steps)
    wall = time.perf_counter() - t0
    skin1 = ctx.skin_lists()
    _, it, dit, redone = ctx.iteration_totals(reset=True)
    k_mean = it / steps
    kdiv_mean = dit / steps
    value = n * steps / wall
    v2 = np.max(np.sum(start.velocities.astype(np.float64) ** 2, axis=1))
    out["collapse"] = {
        "sub_steps": [done, done + steps], "time_s": done * scene.dt,
        "max_speed_at_start": float(math.sqrt(v2)),
        "value": value, "ms_per_step": wall * 1e3 / steps, "device_ms_per_step": dev_ms / steps,
        "mean_iterations": k_mean, "mean_divergence_iterations": kdiv_mean,
        "sub_steps_redone": int(redone),
        "skin_lists": {"on": skin1[0], "rebuilds": skin1[1] - skin0[1],
                       "reuses": skin1[2] - skin0[2]},
        "last_iterations": int(st.iterations), "sort_passes": int(st.sort_passes),
        "cells": int(st.cell_count), "max_neighbors": int(st.max_neighbors),
        "algorithmic_bytes_per_particle_step": step_bytes(method, k_mean, kdiv_mean),
    }
    out["collapse"]["achieved_GBps"] = out["collapse"]["algorithmic_bytes_per_particle_step"] * \
        value / 1e9
    if kernel_timing:
        log(f"{method}: kernel timing pass, {steps} sub-steps")
        ctx.set_kernel_timing(True)
        ctx.run_sub_steps(method, scene.dt, steps)
        out["_timing"] = ctx.kernel_timing()
        out["_timing_steps"] = steps
        ctx.set_kernel_timing(False)
    if cpu is not None:
        per_run, runs = 7, 3
        total = 1 + per_run * runs
        log(f"{method}: GPU states for the CPU parity check")
        gpu_after = gpu_states_after(ctx, method, scene.dt, start, [total, total + 1])
        ctx.close()
        log(f"{method}: CPU leg (oracle on {cpu['threads']} threads, then 1 thread)")
        out["cpu"] = cpu_leg(scene, cfg, method, start, gpu_after, cpu, per_run, runs)
        out["speedup_vs_cpu"] = value / out["cpu"]["value"]
    else:
        ctx.close()
    return out


# The reference's DFSPH_P at this fixed dt ejects a particle at ~1e5 m/s near
# sub-step 2,700 of this dam break (t = 0.042 s) and the run diverges (its
# density and divergence solves then hit their 15-iteration caps); the GPU
# follows it bit for bit. Its window is taken before that.
DFSPH_ADVANCE_LIMIT = 2000


def advance_for(method: str, advance: int) -> int:
    return min(advance, DFSPH_ADVANCE_LIMIT) if method == "dfsph_p" else advance


def roofline_of(method: str, timing: dict, steps: int, n: int) -> tuple:
    """The dominant kernel class by device time: its algorithmic GB/s over its
    average launch, the rocprof HBM bytes per launch and VALU issue share
    recorded in profiles/, and the per-kernel table."""
    table = ALGO_BYTES[method]
    dom_name, (dom_ms, dom_calls) = max(timing.items(), key=lambda kv: kv[1][0])
    avg_s = dom_ms / dom_calls / 1e3
    per = table.get(dom_name)
    achieved = (per * n / avg_s / 1e9) if per else None
    traffic = valu = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(method, {}).get(dom_name)
    vpath = os.path.join(ROOT, "profiles", "valu.json")
    if os.path.exists(vpath):
        with open(vpath) as f:
            valu = json.load(f).get(method, {}).get(dom_name)
    # A kernel whose measured VALU issue takes most of the SIMD cycles is
    # bound by instruction issue; the HBM fraction is reported beside it.
    bound = "valu" if (valu is not None and valu >= 0.6) else "hbm"
    roofline = {
        "bound": bound,
        "kernel": dom_name,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
        "traffic": traffic,
        "avg_launch_us": avg_s * 1e6,
        "algorithmic_bytes_per_particle": per,
        "particles_per_launch": n,
        "valu_issue_frac": valu,
    }
    kernels = {}
    for k, v in sorted(timing.items(), key=lambda kv: -kv[1][0]):
        e = {"ms_per_step": v[0] / steps, "calls_per_step": v[1] / steps,
             "avg_launch_us": v[0] / v[1] * 1e3}
        if table.get(k):
            e["algorithmic_GBps"] = table[k] * n / (v[0] / v[1] / 1e3) / 1e9
        kernels[k] = e
    return roofline, kernels


def batched_form(method: str, steps: int, warmup: int, count: int = 256) -> dict:
    """256 dam breaks of uneven size (log-uniform 2k-200k particles, aspect
    0.25-2, per-scene CFL 0.4 dt) stepped as one batch (one launch per sub-step
    stage for all scenes) and as 256 plain per-scene contexts (one alive at a
    time, each timed over the same sub-steps from the same state); every
    scene's state must come out bit-identical in both forms."""
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
    finals = [batch.get_state(i) for i in range(count)]
    # Per-stage GPU time of the batched sub-step, from the engine's HIP events
    # (timed launches run outside graphs).
    batch.set_kernel_timing(True)
    batch.run_sub_steps(method, 5)
    stages = batch.kernel_timing()
    batch.set_kernel_timing(False)
    batch.close()

    log(f"batched form: {count} plain per-scene contexts, one at a time")
    t_plain = 0.0
    same = True
    for i, (sc, cfg, sol, st) in enumerate(zip(scs, cfgs, solids, states)):
        c = sb.Context(cfg, sol, len(sc.positions))
        c.set_cfl(0.4)
        c.set_state(st)
        c.frame_init(method)
        c.run_sub_steps(method, cfg.seconds_per_sub_step, warmup)
        t0 = time.perf_counter()
        c.run_sub_steps(method, cfg.seconds_per_sub_step, steps)
        t_plain += time.perf_counter() - t0
        c.run_sub_steps(method, cfg.seconds_per_sub_step, 1)
        a = c.get_state()
        same = same and _bits_equal(a.positions, a.velocities, finals[i].positions,
                                    finals[i].velocities)
        c.close()
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


def search_form(cpu: dict | None, runs: int = 5, n: int = 4_194_304) -> dict:
    """Z-index keys, sort, cells and 64-capped neighbour lists alone (no
    dynamics) on 4,194,304 points, half uniform and half from 8 Gaussian
    clusters (sigma 0.02), cell 1/2048 (h = 1/4096). Algorithmic bytes per point
    40 + 1 + 4 n (key build 40, count 1, u32 indices; the lists are the
    product). The CPU leg is the oracle's search over the region the reference
    times (test_neighborhood.cpp:178-231): one warm-up, the median of 3, and one
    run on one thread."""
    import oracle
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
    out = {
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
    if cpu is not None:
        log(f"search form: CPU leg on {cpu['threads']} threads, then 1 thread")
        oracle.set_threads(cpu["threads"])
        _, pairs = oracle.search_seconds(xy, 1.0 / 2048.0)
        times = [oracle.search_seconds(xy, 1.0 / 2048.0)[0] for _ in range(3)]
        oracle.set_threads(1)
        one, _ = oracle.search_seconds(xy, 1.0 / 2048.0)
        oracle.set_threads(cpu["threads"])
        med = statistics.median(times)
        out["cpu"] = {"ms_per_search": med * 1e3, "points_per_s": n / med, "cores": cpu["threads"],
                      "one_thread_ms": one * 1e3, "runs_s": [round(t, 4) for t in times],
                      "neighbour_pairs_equal_gpu": bool(pairs == int(counts.sum())),
                      "kind": "port"}
        out["speedup_vs_cpu"] = (med * 1e3) / float(np.median(ms))
    return out


def small_form(cpu: dict | None, steps: int = 1000, cpu_budget: float = 20.0) -> dict:
    """The reference's own CPU-runnable case: the 12,800-particle IISPH dam
    break (80 x 160 lattice, spacing 0.005, dt 2e-4), 1000 sub-steps on the GPU
    and the same sub-steps in the oracle on the host cores (cpu None: the GPU
    alone), with the final states compared bit for bit."""
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
    # The autotuner's trials start from this state and restore it, so the
    # timed sub-steps below are still the first 1000 of the scene.
    chosen, cand_ms = ctx.autotune("iisph", scene.dt, 5)
    ctx.iteration_totals(reset=True)
    t0 = time.perf_counter()
    dev_ms, _ = ctx.run_sub_steps("iisph", scene.dt, steps)
    wall = time.perf_counter() - t0
    _, iters, _, _ = ctx.iteration_totals(reset=True)
    gst = ctx.get_state()
    ctx.close()
    gpu = {
        "workload": "iisph dam break, 80x160 lattice (12,800 particles), spacing 0.005, "
                    "dt 2e-4, hex walls",
        "sub_steps": steps,
        "value": n * steps / wall,
        "unit": "particle-steps/s",
        "ms_per_sub_step": wall * 1e3 / steps,
        "device_ms_per_sub_step": dev_ms / steps,
        "mean_iterations": iters / steps,
        "tuning": {"chosen": chosen, "candidate_ms_per_5_sub_steps": cand_ms},
    }
    if cpu is None:
        return gpu
    threads = cpu["threads"]
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
        same = _bits_equal(pos, vel, gst.positions, gst.velocities)
    return {
        **gpu,
        "cpu_value": n * done / cwall,
        "cpu_cores": threads,
        "cpu_sub_steps": done,
        "speedup_vs_cpu": (n * steps / wall) / (n * done / cwall),
        "final_state_bit_identical_to_oracle": same,
    }


def fp64_errors(req, state, cpu: dict) -> dict:
    """A completed served request (FP32 on the GPU) against the oracle's FP64
    build of the same scene over the same number of sub-steps, each with its
    own CFL 0.4 dt (cfl.cpp:12-21 on its own velocities): error per state
    field, and the kinetic energies."""
    import oracle
    from emerald_amd import serve
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_scene
    sc, cfg = serve.request_scene(req)
    oracle.set_threads(cpu["threads"])
    o = oracle_scene(sc, cfg, fp64=True)
    o.frame_init(req.method)
    h = float(np.float32(sc.support))
    dt_max = float(np.float32(cfg.seconds_per_sub_step))
    vmax2 = 0.0
    t = 0.0
    for _ in range(req.steps):
        dt = dt_max if vmax2 <= 0 else min(dt_max, float(np.float32(0.4)) * h / math.sqrt(vmax2))
        o.sub_step(req.method, t, dt)
        t += dt
        _, vel = o.state_f64()
        vmax2 = float(np.max(np.sum(vel * vel, axis=1)))
    pos64, vel64 = o.state_f64()
    m = float(np.float32(cfg.mass_per_particle))
    p32 = state.positions.astype(np.float64)
    v32 = state.velocities.astype(np.float64)
    dp = np.hypot(*(p32 - pos64).T)
    dv = np.hypot(*(v32 - vel64).T)
    ke32 = 0.5 * m * np.sum(v32 ** 2)
    ke64 = 0.5 * m * np.sum(vel64 ** 2)
    return {"request": req.rid, "method": req.method, "particles": len(sc.positions),
            "steps": req.steps,
            "positions_max": float(dp.max()), "positions_rms": float(np.sqrt(np.mean(dp ** 2))),
            "velocities_max": float(dv.max()), "velocities_rms": float(np.sqrt(np.mean(dv ** 2))),
            "kinetic_energy_rel": float(abs(ke32 - ke64) / ke64) if ke64 > 0 else None}


def served_form(budget_gb: float, window: float, cpu: dict | None, rate: float = 200.0,
                plain_steps: int = 2, batch_steps: int = 5) -> dict:
    """Poisson stream of dam-break requests (200/s, 5k-500k particles, IISPH or
    DFSPH_P, 100-2000 steps) served from two device-resident pools (one per
    solver) with continuous batching under an HBM budget. After the window,
    the scenes still live are stepped from their final states as one batch
    per solver and as plain per-scene contexts (one at a time), for the
    throughput comparisons at the same states, and the two smallest completed
    requests are compared with the oracle's FP64 build."""
    from emerald_amd import serve
    from emerald_amd import sph2d_box as sb
    reqs = serve.poisson_requests(rate, window + 5.0)
    # Final states are kept for the requests small enough for the FP64 check.
    small = lambda r: r.n * r.steps <= 2e7  # noqa: E731
    res = serve.serve(reqs, budget_gb * 1e9, window, log=log, keep_states=small)
    loops = res["_loops"]
    log("served form: the final live sets as batches and as plain per-scene contexts")
    live_sets = {}
    for m, lp in loops.items():
        live = list(lp.live.items())
        live_sets[m] = [(v[0], lp.pool.get_state(sid)) for sid, v in live]
        lp.pool.close()
    finals = {rid: st for lp in loops.values() for rid, st in lp.finals.items()}
    serve.close(res)
    b_steps = 0
    t_batch = 0.0
    b_iters = []
    n_steps = 0
    t_plain = 0.0
    scenes_done = 0
    for m, items in live_sets.items():
        if not items:
            continue
        built = [serve.request_scene(req) for req, _ in items]
        batch = sb.Batch([cfg for _, cfg in built], [sb.Solid_state(sc.solids) for sc, _ in built],
                         [st for _, st in items])
        batch.set_cfl(0.4)
        batch.frame_init(m)
        batch.run_sub_steps(m, 2)
        batch.iteration_totals(reset=True)
        t0 = time.perf_counter()
        batch.run_sub_steps(m, batch_steps)
        t_batch += time.perf_counter() - t0
        steps, it, dit, _ = batch.iteration_totals(reset=True)
        b_iters.append({"method": m, "iterations_per_sub_step": it / steps,
                        "divergence_iterations_per_sub_step": dit / steps})
        b_steps += sum(len(sc.positions) for sc, _ in built) * batch_steps
        batch.close()
        for (req, st), (sc, cfg) in zip(items, built):
            c = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
            c.set_cfl(0.4)
            c.set_state(st)
            c.frame_init(m)
            c.run_sub_steps(m, cfg.seconds_per_sub_step, 1)
            t0 = time.perf_counter()
            c.run_sub_steps(m, cfg.seconds_per_sub_step, plain_steps)
            t_plain += time.perf_counter() - t0
            n_steps += len(sc.positions) * plain_steps
            scenes_done += 1
            c.close()
    res["batched_value_final_live_set"] = b_steps / t_batch if t_batch > 0 else None
    res["batched_iterations_final_live_set"] = b_iters
    res["plain_value_final_live_set"] = n_steps / t_plain if t_plain > 0 else None
    res["plain_scenes"] = scenes_done
    res["speedup_vs_plain"] = (res["value"] / res["plain_value_final_live_set"]
                               if res["plain_value_final_live_set"] else None)
    res["served_over_batched_same_states"] = (res["value"] / res["batched_value_final_live_set"]
                                              if res["batched_value_final_live_set"] else None)
    if cpu is not None and finals:
        by_id = {r.rid: r for r in reqs}
        picks = sorted(finals, key=lambda rid: by_id[rid].n * by_id[rid].steps)[:2]
        log(f"served form: FP64 oracle on requests {picks}")
        res["fp64_error"] = [fp64_errors(by_id[rid], finals[rid], cpu) for rid in picks]
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--advance", type=int, default=8000,
                    help="sub-steps from the initial state to the timed collapse window")
    ap.add_argument("--method", default="iisph",
                    choices=["iisph", "iisph_ap", "iisph_pseudo_ap", "dfsph_p", "pcisph"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solvers", action="store_true")
    ap.add_argument("--solver-steps", type=int, default=50)
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
    ap.add_argument("--small-only", action="store_true")
    ap.add_argument("--skin-lists", choices=["on", "off"], default="on",
                    help="skin candidate lists in the 1M solver windows (the library default)")
    ap.add_argument("--tuning", default=None, metavar="NB,XCD",
                    help="fixed tuning of the 1M solver windows instead of the autotuner: "
                         "neighbour batch (2, 4 or 8) and XCD remap (0 or 1), without graphs")
    ap.add_argument("--save-state", default=None,
                    help="write the headline solver's state at the start of its timed window "
                         "(.npz)")
    ap.add_argument("--from-state", default=None,
                    help="start the headline solver from a state written by --save-state "
                         "(no rest window or advance)")
    args = ap.parse_args()
    global SKIN_LISTS, FIXED_TUNING
    SKIN_LISTS = args.skin_lists == "on"
    if args.tuning:
        nb, xcd = (int(v) for v in args.tuning.split(","))
        if nb not in (2, 4, 8) or xcd not in (0, 1):
            raise SystemExit("bench: --tuning takes NB,XCD with NB in 2, 4, 8 and XCD in 0, 1")
        FIXED_TUNING = (nb, xcd)
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only; --gpus above 1 is "
                                   "not supported", "n_gpus": args.gpus}))
        return 2

    from emerald_amd import lib
    if lib().emsph_device_count() < 1:
        raise SystemExit("bench: no HIP device")
    cpu = None if args.no_cpu_baseline else host_cpu()

    method = args.method
    if args.batched_only:
        print(json.dumps(batched_form(method, args.batched_steps, 2)), flush=True)
        return 0
    if args.served_only:
        print(json.dumps(served_form(args.served_budget_gb, args.served_window, cpu)), flush=True)
        return 0
    if args.search_only:
        print(json.dumps(search_form(cpu)), flush=True)
        return 0
    if args.small_only:
        print(json.dumps(small_form(cpu)), flush=True)
        return 0

    head = solver_window(method, args.steps, args.warmup, advance_for(method, args.advance), cpu,
                         autotune=True, kernel_timing=True, rest_window=True,
                         save_state=args.save_state, from_state=args.from_state)
    n = head["particles"]
    roofline, kernels = roofline_of(method, head.pop("_timing"), head.pop("_timing_steps"), n)
    col = head["collapse"]
    whole_step = {
        "algorithmic_bytes_per_particle_step": col["algorithmic_bytes_per_particle_step"],
        "mean_iterations": col["mean_iterations"],
        "mean_divergence_iterations": col["mean_divergence_iterations"],
        "achieved_GBps": col["achieved_GBps"],
        "frac_of_hbm_peak": col["achieved_GBps"] / HBM_PEAK_GBS,
    }

    solvers = None
    if not args.no_solvers:
        solvers = {}
        for m in ("iisph", "pcisph", "dfsph_p"):
            if m == method:
                continue
            r = solver_window(m, args.solver_steps, 5, advance_for(m, args.advance), cpu,
                              autotune=True, kernel_timing=False, rest_window=False)
            solvers[m] = r
    batched = None
    if not args.no_batched:
        batched = batched_form(method, args.batched_steps, 2)
    served = None
    if not args.no_served:
        served = served_form(args.served_budget_gb, args.served_window, cpu)
        if batched:
            served["served_over_batched"] = served["value"] / batched["value"]
    search = None
    if not args.no_search:
        search = search_form(cpu)
    small = None
    if not args.no_small and cpu is not None:
        small = small_form(cpu)

    cpu_baseline = None
    if cpu is not None:
        c = head["cpu"]
        cpu_baseline = {k: c[k] for k in ("value", "unit", "cores", "kind", "sample",
                                          "one_thread_value", "runs_s", "bit_identical_to_gpu")}
        cpu_baseline["model"] = cpu["model"]
        cpu_baseline["host_cpus"] = {k: cpu[k] for k in ("affinity_cpus", "cgroup_cpus",
                                                         "omp_num_threads_env")}
    out = {
        "metric": "IISPH particle-steps/s @1M (1 GPU)" if method == "iisph"
        else f"{method.upper()} particle-steps/s @1M (1 GPU)",
        "value": col["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": col["ms_per_step"],
        "device_ms_per_step": col["device_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{method} dam break, 1024x1024 lattice (1,048,576 fluid particles), "
                        f"spacing 1/2560, h = spacing, dt 1.5625e-5, hex walls; timed window "
                        f"sub-steps {col['sub_steps'][0]}-{col['sub_steps'][1]} "
                        f"(t = {col['time_s']:.3f} s, the collapse)",
            "particles": n,
            "solids": head["solids"],
            "dt": head["dt"],
            "mean_iterations": col["mean_iterations"],
            "last_iterations": col["last_iterations"],
            "sort_passes": col["sort_passes"],
            "cells": col["cells"],
            "max_neighbors": col["max_neighbors"],
        },
        "windows": {"collapse": col, "rest": head.get("rest")},
        "roofline": roofline,
        "whole_step": whole_step,
        "tuning": head.get("tuning"),
        "cpu_baseline": cpu_baseline,
        "speedup_vs_cpu": head.get("speedup_vs_cpu"),
        "solvers": solvers,
        "batched": batched,
        "served": served,
        "search": search,
        "small": small,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
