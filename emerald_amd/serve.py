"""Served form: continuous batching of dam-break requests over device-resident
pools.

Requests arrive as a seeded Poisson stream. Each is a dam-break scene of
log-uniform size with a solver (IISPH or DFSPH_P) and a number of steps. One
pool per solver holds the live scenes of that solver on the device (its
capacity is its share of the HBM budget); a thread per pool loops: admit the
requests that have arrived, in arrival order, while they fit (emsph_pool_add
refuses what does not), commit, run one sub-step of every live scene, retire
the scenes that have done their steps. Admitted scenes join the next sub-step
of a pool that is already running (continuous batching), and the two pools'
sub-steps overlap on the device (one HIP stream each). A scene's results do
not depend on what else is live (tests/test_gpu_pool.py).

Per-step latency of a scene = wall time from its previous step (or its
admission) to the end of the sub-step that advanced it.
"""

from __future__ import annotations

import dataclasses
import math
import queue
import threading
import time

import numpy as np

from . import scenes
from . import sph2d_box as sb
from ._capi import EmsphError


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float  # seconds after the start
    n: int          # fluid particles asked for
    aspect: float   # block width / height
    method: str
    steps: int
    seed: int


def poisson_requests(rate: float, duration: float, seed: int = 5, n_min: int = 5000,
                     n_max: int = 500_000, steps_min: int = 100, steps_max: int = 2000,
                     methods=("iisph", "dfsph_p")) -> list:
    """Arrivals of a Poisson process of the given rate over duration seconds
    (exponential gaps from a SplitMix64 stream); sizes log-uniform, aspect
    U[0.25, 2], solver uniform over methods, steps uniform in
    [steps_min, steps_max]."""
    cap = int(rate * duration * 2 + 64)
    u = scenes.uniform01(seed, 5 * cap).reshape(cap, 5)
    out = []
    t = 0.0
    for k in range(cap):
        t += -math.log(1.0 - u[k, 0]) / rate
        if t >= duration:
            break
        n = int(round(math.exp(math.log(n_min) + u[k, 1] * (math.log(n_max) - math.log(n_min)))))
        aspect = 0.25 + u[k, 2] * 1.75
        method = methods[min(int(u[k, 3] * len(methods)), len(methods) - 1)]
        steps = steps_min + min(int(u[k, 4] * (steps_max - steps_min + 1)), steps_max - steps_min)
        out.append(Request(k, t, n, aspect, method, steps, seed * 100003 + k))
    return out


def request_scene(req: Request, cfl: float = 0.4, gravity: float = 9.81):
    """The request's dam break (as in scenes.batch_dam_breaks) and its config."""
    nx = max(1, int(round(math.sqrt(req.n * req.aspect))))
    ny = max(1, int(round(req.n / nx)))
    s = 0.8 / max(nx, ny)
    h = float(np.float32(s))
    dt = cfl * h / math.sqrt(2.0 * gravity * ny * s)
    sc = scenes.dam_break(nx, ny, s, float(np.float32(dt)), seed=req.seed)
    return sc, scenes.config_for(sc, req.method)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q)) if len(xs) else None


class _PoolLoop:
    def __init__(self, method: str, fluid_cap: int, solid_cap: int, max_scenes: int, cfl: float,
                 commit_interval: float, tuning: dict | None = None):
        self.method = method
        self.commit_interval = commit_interval
        self.pool = sb.Pool(fluid_cap, solid_cap, max_scenes, method)
        self.pool.set_cfl(cfl)
        if tuning:
            self.pool.set_tuning(**tuning)
        self.ready = queue.Queue(maxsize=256)  # (Request, DamBreak, config) in arrival order
        self.pending = None                    # head that did not fit yet
        self.live = {}                         # id -> [request, steps done, last time, n]
        self.latency = []                      # per-step latencies (s)
        self.wait = []                         # arrival -> admission (s)
        self.completed = []
        self.admitted = 0
        self.ticks = 0
        self.particle_steps = 0
        self.live_particles = []
        self.finals = {}                       # rid -> final State (when keep_states)
        self.dirty = False                     # retirements staged since the last commit
        self.commit_s = 0.0                    # wall time in commits
        self.step_s = 0.0                      # wall time in sub-steps

    def run(self, t0: float, end: float, keep_states) -> None:
        pool = self.pool
        last_commit = -1e9
        staged = False
        while True:
            now = time.perf_counter() - t0
            if now >= end:
                break
            # Admission, in arrival order, while the head has arrived and fits
            # (staged now, live from the next commit).
            while True:
                if self.pending is None:
                    try:
                        self.pending = self.ready.get_nowait()
                    except queue.Empty:
                        break
                req, sc, cfg = self.pending
                if req.arrival > now:
                    break
                try:
                    sid = pool.add(cfg, sb.Solid_state(sc.solids),
                                   sb.State(sc.positions, sc.velocities))
                except EmsphError:
                    break  # does not fit until something retires
                self.live[sid] = [req, 0, None, len(sc.positions)]
                self.admitted += 1
                self.pending = None
                staged = True
            # Commits (which repack the pool) at most every commit_interval;
            # until then removed scenes ride along in the sort and the
            # neighbour build, unaccounted, but take no solver iterations.
            idle = len(self.live) == 0 or not pool.live()
            if (staged or self.dirty) and (idle or now - last_commit >= self.commit_interval):
                tb = time.perf_counter()
                pool.commit()
                self.dirty = False
                staged = False
                tc = time.perf_counter()
                last_commit = tc - t0
                self.commit_s += tc - tb
                for v in self.live.values():
                    if v[2] is None:
                        v[2] = tc
                        self.wait.append(tc - t0 - v[0].arrival)
            ids = pool.live()
            if not ids:
                time.sleep(0.0005)
                continue
            ts = time.perf_counter()
            pool.sub_step(with_stats=False)
            td = time.perf_counter()
            self.step_s += td - ts
            self.ticks += 1
            n_live = 0
            for sid in ids:
                v = self.live.get(sid)
                if v is None:
                    continue  # retired, removed at the next commit
                self.latency.append(td - v[2])
                v[2] = td
                v[1] += 1
                n_live += v[3]
                if v[1] >= v[0].steps:
                    if keep_states is True or (keep_states and keep_states(v[0])):
                        self.finals[v[0].rid] = pool.get_state(sid)
                    pool.remove(sid)
                    self.completed.append(v[0].rid)
                    del self.live[sid]
                    self.dirty = True
            self.particle_steps += n_live
            self.live_particles.append(n_live)


def serve(requests: list, budget_bytes: float, window: float, cfl: float = 0.4,
          methods=("iisph", "dfsph_p"), max_scenes: int = 4096, solid_ratio: float = 0.25,
          keep_states=False, commit_interval: float = 0.2, log=None,
          tuning: dict | None = None) -> dict:
    """Runs the request stream for window seconds; returns the served metrics
    (and the pools' final states of completed requests when keep_states is
    True, or of those for which keep_states(request) is true). tuning: keyword
    arguments of Pool.set_tuning for every pool (None: the library's
    defaults)."""
    # Device bytes per fluid particle (with solid_ratio solids per fluid
    # particle), measured on a small pool.
    probe = sb.Pool(1 << 20, int(solid_ratio * (1 << 20)), 16, methods[0])
    per_particle = probe.device_bytes / float(1 << 20)
    probe.close()
    share = budget_bytes / len(methods)
    fluid_cap = int(min(share / per_particle, (1 << 30) - 1))
    loops = {m: _PoolLoop(m, fluid_cap, int(solid_ratio * fluid_cap), max_scenes, cfl,
                          commit_interval, tuning)
             for m in methods}
    if log:
        log(f"served: {len(methods)} pools x {fluid_cap} particles "
            f"({sum(lp.pool.device_bytes for lp in loops.values()) / 1e9:.1f} GB), "
            f"{len(requests)} requests over {window:.0f} s")
    stop = threading.Event()

    def generate(method):
        # The client side: each request's scene is built ahead of its arrival.
        for req in requests:
            if req.method != method:
                continue
            if stop.is_set():
                return
            sc, cfg = request_scene(req, cfl)
            while not stop.is_set():
                try:
                    loops[method].ready.put((req, sc, cfg), timeout=0.05)
                    break
                except queue.Full:
                    continue

    gens = [threading.Thread(target=generate, args=(m,), daemon=True) for m in methods]
    for g in gens:
        g.start()
    time.sleep(0.5)  # a head start for the scene builders
    t0 = time.perf_counter()
    threads = [threading.Thread(target=lp.run, args=(t0, window, keep_states)) for lp in
               loops.values()]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t0
    stop.set()
    for g in gens:
        g.join(timeout=5.0)
    lat = [x for lp in loops.values() for x in lp.latency]
    wait = [x for lp in loops.values() for x in lp.wait]
    ps = sum(lp.particle_steps for lp in loops.values())
    arrived = sum(1 for r in requests if r.arrival < window)
    out = {
        "window_s": window, "wall_s": wall, "budget_bytes": budget_bytes,
        "pool_bytes": sum(lp.pool.device_bytes for lp in loops.values()),
        "pool_fluid_capacity": fluid_cap,
        "arrived": arrived,
        "admitted": sum(lp.admitted for lp in loops.values()),
        "completed": sum(len(lp.completed) for lp in loops.values()),
        "live_at_end": sum(len(lp.live) for lp in loops.values()),
        "particle_steps": ps, "value": ps / wall, "unit": "particle-steps/s",
        "step_latency_ms": {"p50": 1e3 * percentile(lat, 50) if lat else None,
                            "p99": 1e3 * percentile(lat, 99) if lat else None,
                            "steps": len(lat)},
        "admission_wait_ms": {"p50": 1e3 * percentile(wait, 50) if wait else None,
                              "p99": 1e3 * percentile(wait, 99) if wait else None},
        "per_pool": {m: {"ticks": lp.ticks, "admitted": lp.admitted,
                         "completed": len(lp.completed),
                         "iterations_per_tick": _iterations_per_tick(lp),
                         "commit_s": lp.commit_s, "sub_step_s": lp.step_s,
                         "mean_live_particles": float(np.mean(lp.live_particles))
                         if lp.live_particles else 0.0}
                     for m, lp in loops.items()},
    }
    out["_loops"] = loops
    return out


def _iterations_per_tick(lp) -> dict:
    """Solver iterations launched per sub-step of a pool (the maximum over its
    live scenes), over the window."""
    steps, it, dit, _ = lp.pool.iteration_totals()
    if not steps:
        return {"iterations": None, "divergence_iterations": None}
    return {"iterations": it / steps, "divergence_iterations": dit / steps}


def close(result: dict) -> None:
    for lp in result.pop("_loops", {}).values():
        lp.pool.close()
