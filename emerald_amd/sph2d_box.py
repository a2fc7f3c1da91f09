"""Python mirror of the reference's sph2d_box step API over the gfx950 C ABI.

Names, argument meaning and error behaviour follow the reference:
  Parameters                 sph2d_box/parameters.h:12-43 (parse_parameters: parameters.cpp:10-44)
  Simulation_config          sph2d_box/config.h:7-18, config.cpp:16-132
  State / Solid_state        sph2d_box/state.h:13-35
  world_walls_initial_solid_state   solids.cpp:139-213
  dam_break_initial_state    initial_state.cpp:14-97
  <method>_simulation_step   iisph.cpp:168-188, pcisph.cpp:166-204, dfsph_p.cpp:376-396
  EZ_EXAMPLE_SIM             simulation.h:24-44, simulation.cpp:14-93

The per-sub-step work runs on the GPU through libemerald_sph.so; this module
only moves host arrays in and out and holds the frame clock (flicks).
"""

from __future__ import annotations

import argparse
import ctypes
import dataclasses
from typing import Optional

import numpy as np

from . import _capi
from ._capi import check, lib

FLICKS_PER_SECOND = 705_600_000
K_FLICKS_ONE_NINETIETH_OF_SECOND = FLICKS_PER_SECOND // 90


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclasses.dataclass
class IisphParameters:
    max_pressure_iterations: int = 30
    error_average_threshold: float = 0.005
    error_max_threshold: float = 0.05
    omega: float = 0.5


@dataclasses.dataclass
class DfsphParameters:
    max_correction_iterations: int = 15
    density_error_average_threshold: float = 0.0035
    density_error_max_threshold: float = 0.03
    divergence_error_average_threshold: float = 0.0035
    divergence_error_max_threshold: float = 0.03


@dataclasses.dataclass
class Parameters:
    """sph2d_box/parameters.h:12-43 (same field names and defaults)."""
    seed: int = 1
    time_per_step: int = 3 * K_FLICKS_ONE_NINETIETH_OF_SECOND  # flicks
    sub_steps: int = 30
    length: float = 1.0
    support: float = 0.025
    gravity: float = 9.81
    viscosity: float = 0.01
    target_density: float = 1000.0
    iisph: IisphParameters = dataclasses.field(default_factory=IisphParameters)
    dfsph: DfsphParameters = dataclasses.field(default_factory=DfsphParameters)
    num_batch_steps: int = 100
    method: str = "iisph_pseudo_ap"

    def to_c(self) -> _capi.Parameters:
        p = _capi.Parameters()
        p.seed = self.seed
        p.time_per_step = self.time_per_step
        p.sub_steps = self.sub_steps
        p.length = self.length
        p.support = self.support
        p.gravity = self.gravity
        p.viscosity = self.viscosity
        p.target_density = self.target_density
        p.iisph_max_pressure_iterations = self.iisph.max_pressure_iterations
        p.iisph_error_average_threshold = self.iisph.error_average_threshold
        p.iisph_error_max_threshold = self.iisph.error_max_threshold
        p.iisph_omega = self.iisph.omega
        p.dfsph_max_correction_iterations = self.dfsph.max_correction_iterations
        p.dfsph_density_error_average_threshold = self.dfsph.density_error_average_threshold
        p.dfsph_density_error_max_threshold = self.dfsph.density_error_max_threshold
        p.dfsph_divergence_error_average_threshold = self.dfsph.divergence_error_average_threshold
        p.dfsph_divergence_error_max_threshold = self.dfsph.divergence_error_max_threshold
        p.num_batch_steps = self.num_batch_steps
        p.method = _capi.METHODS.get(self.method, -1)
        return p


def parse_parameters(argv) -> Parameters:
    """parameters.cpp:10-44: --seed --sub_steps --length --support --gravity
    --viscosity --density -n/--num_steps --method."""
    ap = argparse.ArgumentParser(prog="sph2d_box")
    d = Parameters()
    ap.add_argument("--seed", type=int, default=d.seed)
    ap.add_argument("--sub_steps", type=int, default=d.sub_steps)
    ap.add_argument("--length", type=float, default=d.length)
    ap.add_argument("--support", type=float, default=d.support)
    ap.add_argument("--gravity", type=float, default=d.gravity)
    ap.add_argument("--viscosity", type=float, default=d.viscosity)
    ap.add_argument("--density", type=float, default=d.target_density)
    ap.add_argument("-n", "--num_steps", type=int, default=d.num_batch_steps)
    ap.add_argument("--method", type=str, default=d.method)
    a = ap.parse_args(argv)
    return Parameters(seed=a.seed, sub_steps=a.sub_steps, length=a.length, support=a.support,
                      gravity=a.gravity, viscosity=a.viscosity, target_density=a.density,
                      num_batch_steps=a.num_steps, method=a.method)


@dataclasses.dataclass
class Simulation_config:  # noqa: N801 - the reference's type name
    """config.cpp:16-132. Optional overrides replace the calibrated values."""
    params: Parameters
    seconds_per_sub_step: float = 0.0
    mass_per_particle: float = 0.0
    draw_radius: float = 0.0
    pressure_correction_denom: float = 0.0

    @staticmethod
    def from_parameters(params: Parameters, **overrides) -> "Simulation_config":
        c = _capi.Config()
        cp = params.to_c()
        check(lib().emsph_config_from_parameters(ctypes.byref(cp), ctypes.byref(c)),
              "emsph_config_from_parameters")
        cfg = Simulation_config(params, float(c.seconds_per_sub_step),
                                float(c.mass_per_particle), float(c.draw_radius),
                                float(c.pressure_correction_denom))
        for k, v in overrides.items():
            if not hasattr(cfg, k):
                raise AttributeError(k)
            setattr(cfg, k, v)
        return cfg

    def to_c(self) -> _capi.Config:
        c = _capi.Config()
        c.params = self.params.to_c()
        c.seconds_per_sub_step = self.seconds_per_sub_step
        c.mass_per_particle = self.mass_per_particle
        c.draw_radius = self.draw_radius
        c.pressure_correction_denom = self.pressure_correction_denom
        return c


@dataclasses.dataclass
class State:
    """state.h:13-17."""
    positions: np.ndarray   # [N, 2] float32
    velocities: np.ndarray  # [N, 2] float32
    colors: Optional[np.ndarray] = None  # [N, 4] float32 RGBA, set by a frame step


@dataclasses.dataclass
class Solid_state:  # noqa: N801
    """state.h:19-35: positions (static); volumes are computed on the device."""
    positions: np.ndarray
    velocities: Optional[np.ndarray] = None


def emit_solid_box(box, R: float) -> np.ndarray:
    """solids.cpp:13-50."""
    (x0, y0), (x1, y1) = box
    n = lib().emsph_estimate_solid_box_emission_count(x0, y0, x1, y1, R)
    out = np.empty((max(n, 1), 2), np.float32)
    m = lib().emsph_emit_solid_box(x0, y0, x1, y1, R, n, _ptr(out))
    return out[:m].copy()


def world_walls_initial_solid_state(params: Parameters, random_obstacles: bool = True) -> Solid_state:
    """solids.cpp:139-213 (walls 3h thick; optional 16 random obstacle blocks)."""
    p = params.to_c()
    n = lib().emsph_world_walls_solid_positions(ctypes.byref(p), 1 if random_obstacles else 0,
                                                None, 0)
    out = np.empty((n, 2), np.float32)
    lib().emsph_world_walls_solid_positions(ctypes.byref(p), 1 if random_obstacles else 0,
                                            _ptr(out), n)
    return Solid_state(out, np.zeros_like(out))


def dam_break_initial_state(params: Parameters, solid_state: Solid_state) -> State:
    """initial_state.cpp:14-97: hex block [0, L/2] x [0, 3.25 L/4], then drop
    fluid particles closer than 2R to any solid."""
    p = params.to_c()
    sp = np.ascontiguousarray(solid_state.positions, np.float32).reshape(-1, 2)
    n = lib().emsph_dam_break_positions(ctypes.byref(p), _ptr(sp), len(sp), None, 0)
    pos = np.empty((n, 2), np.float32)
    lib().emsph_dam_break_positions(ctypes.byref(p), _ptr(sp), len(sp), _ptr(pos), n)
    return State(pos, np.zeros_like(pos))


@dataclasses.dataclass
class ForceModel:
    """Device external-force models (emsph_force_model; forces.cpp:62-149),
    accumulated in this order after a host callback: gravity
    (accumulate_gravity_forces), constant pole attraction, simple drag,
    anti-coupling repulsion. gravity alone is default_gravity_forces."""
    gravity_on: bool = True
    gravity: float = 9.81
    pole_magnitude: Optional[float] = None   # None: off
    pole: tuple = (0.5, 0.5)
    drag_magnitude: Optional[float] = None   # None: off
    drag_diameter: float = 0.0
    anti_coupling_max_distance: Optional[float] = None  # None: off
    anti_coupling_magnitude: float = 0.0

    def to_c(self) -> _capi.ForceModel:
        m = _capi.ForceModel()
        m.flags = ((_capi.FORCE_GRAVITY if self.gravity_on else 0)
                   | (_capi.FORCE_POLE if self.pole_magnitude is not None else 0)
                   | (_capi.FORCE_DRAG if self.drag_magnitude is not None else 0)
                   | (_capi.FORCE_ANTI_COUPLING
                      if self.anti_coupling_max_distance is not None else 0))
        m.gravity = self.gravity
        m.pole_magnitude = self.pole_magnitude or 0.0
        m.pole_x, m.pole_y = self.pole
        m.drag_magnitude = self.drag_magnitude or 0.0
        m.drag_diameter = self.drag_diameter
        m.anti_coupling_max_distance = self.anti_coupling_max_distance or 0.0
        m.anti_coupling_magnitude = self.anti_coupling_magnitude
        return m

    @staticmethod
    def from_c(m: _capi.ForceModel) -> "ForceModel":
        f = int(m.flags)
        return ForceModel(
            gravity_on=bool(f & _capi.FORCE_GRAVITY), gravity=float(m.gravity),
            pole_magnitude=float(m.pole_magnitude) if f & _capi.FORCE_POLE else None,
            pole=(float(m.pole_x), float(m.pole_y)),
            drag_magnitude=float(m.drag_magnitude) if f & _capi.FORCE_DRAG else None,
            drag_diameter=float(m.drag_diameter),
            anti_coupling_max_distance=(float(m.anti_coupling_max_distance)
                                        if f & _capi.FORCE_ANTI_COUPLING else None),
            anti_coupling_magnitude=float(m.anti_coupling_magnitude))


def default_gravity_forces(gravity: float) -> ForceModel:
    """forces.cpp:72-83: the default User_forces_function, as the device model."""
    return ForceModel(gravity_on=True, gravity=gravity)


def user_forces_callback(fn):
    """A C callback (emsph_user_forces_fn) around a User_forces_function
    (forces.h:13-21) written in Python: fn(count, t, dt, mass_per_particle,
    radius_per_particle, forces, positions, velocities) with (count, 2)
    float32 views; forces arrives zeroed and fn accumulates into it."""
    def tramp(_user, count, t, dt, mass, radius, f, p, v):
        n = int(count)
        fa = np.ctypeslib.as_array(f, shape=(n, 2))
        pa = np.ctypeslib.as_array(p, shape=(n, 2))
        va = np.ctypeslib.as_array(v, shape=(n, 2))
        fn(n, float(t), float(dt), float(mass), float(radius), fa, pa, va)
    return _capi.USER_FORCES_FN(tramp)


def user_colors_callback(fn):
    """A C callback (emsph_user_colors_fn) around a User_colors_function
    (colors.h:13-21) written in Python: fn(count, t, mass_per_particle,
    radius_per_particle, colors, positions, velocities, densities) with
    (count, 4) float32 colors (written), (count, 2) state and (count,)
    densities."""
    def tramp(_user, count, t, mass, radius, c, p, v, d):
        n = int(count)
        fn(n, float(t), float(mass), float(radius), np.ctypeslib.as_array(c, shape=(n, 4)),
           np.ctypeslib.as_array(p, shape=(n, 2)), np.ctypeslib.as_array(v, shape=(n, 2)),
           np.ctypeslib.as_array(d, shape=(n,)))
    return _capi.USER_COLORS_FN(tramp)


def encode_snapshot(state: State, time_flicks: int = 0, method: Optional[str] = None) -> bytes:
    """A State (and its colors, if set) in the snapshot format of
    emsph_snapshot_encode. No counterpart in the reference."""
    pos = np.ascontiguousarray(state.positions, np.float32).reshape(-1, 2)
    vel = None if state.velocities is None else np.ascontiguousarray(
        state.velocities, np.float32).reshape(-1, 2)
    col = None if state.colors is None else np.ascontiguousarray(
        state.colors, np.float32).reshape(-1, 4)
    flags = (1 if vel is not None else 0) | (2 if col is not None else 0)
    n = len(pos)
    buf = ctypes.create_string_buffer(lib().emsph_snapshot_bytes(n, flags))
    written = ctypes.c_size_t()
    check(lib().emsph_snapshot_encode(_ptr(pos), _ptr(vel), _ptr(col), n, int(time_flicks),
                                      _capi.METHODS.get(method, -1), buf, len(buf),
                                      ctypes.byref(written)), "snapshot_encode")
    return buf.raw[:written.value]


def decode_snapshot(data: bytes):
    """(State, time in flicks, method name or None) from snapshot bytes."""
    h = _capi.SnapshotHeader()
    check(lib().emsph_snapshot_info(data, len(data), ctypes.byref(h)), "snapshot_info")
    n = int(h.count)
    pos = np.empty((n, 2), np.float32)
    vel = np.empty((n, 2), np.float32)
    col = np.empty((n, 4), np.float32) if h.flags & 2 else None
    check(lib().emsph_snapshot_decode(data, len(data), _ptr(pos), _ptr(vel), _ptr(col)),
          "snapshot_decode")
    names = {v: k for k, v in _capi.METHODS.items()}
    return State(pos, vel, col), int(h.time_flicks), names.get(int(h.method))


def save_state(path: str, state: State, time_flicks: int = 0,
               method: Optional[str] = None) -> None:
    data = encode_snapshot(state, time_flicks, method)
    check(lib().emsph_snapshot_write_file(path.encode(), data, len(data)), "snapshot_write")


def load_state(path: str):
    with open(path, "rb") as f:
        return decode_snapshot(f.read())


class Context:
    """Owns one device-resident scene (emsph_context)."""

    def __init__(self, config: Simulation_config, solid_state: Solid_state, capacity: int):
        self.config = config
        c = config.to_c()
        self._cfg = c
        sp = np.ascontiguousarray(solid_state.positions, np.float32).reshape(-1, 2)
        h = ctypes.c_void_p()
        check(lib().emsph_create(ctypes.byref(c), _ptr(sp), len(sp), int(capacity),
                                 ctypes.byref(h)), "emsph_create")
        self._h = h
        self.capacity = int(capacity)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().emsph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def n(self) -> int:
        return int(lib().emsph_fluid_count(self._h))

    def set_store_geometry(self, enable: bool) -> None:
        check(lib().emsph_set_store_geometry(self._h, 1 if enable else 0), "set_store_geometry")

    def set_state(self, state: State) -> None:
        pos = np.ascontiguousarray(state.positions, np.float32).reshape(-1, 2)
        vel = None if state.velocities is None else np.ascontiguousarray(
            state.velocities, np.float32).reshape(-1, 2)
        check(lib().emsph_set_state(self._h, _ptr(pos), _ptr(vel), len(pos)), "set_state")

    def get_state(self) -> State:
        n = self.n
        pos = np.empty((n, 2), np.float32)
        vel = np.empty((n, 2), np.float32)
        check(lib().emsph_get_state(self._h, _ptr(pos), _ptr(vel)), "get_state")
        return State(pos, vel)

    def field(self, name: str) -> np.ndarray:
        fid, comps = _capi.FIELDS[name]
        out = np.empty((self.n, comps), np.float32)
        check(lib().emsph_get_field(self._h, fid, _ptr(out)), f"get_field({name})")
        return out[:, 0].copy() if comps == 1 else out

    def solid_volumes(self) -> np.ndarray:
        n = int(lib().emsph_solid_count(self._h))
        out = np.empty(n, np.float32)
        check(lib().emsph_get_solid_volumes(self._h, _ptr(out)), "get_solid_volumes")
        return out

    def neighborhood(self, which: int = 0, geometry: bool = False):
        n = self.n
        counts = np.empty(n, np.uint8)
        idx = np.empty((n, 64), np.uint32)
        dist = np.empty((n, 64), np.float32) if geometry else None
        vec = np.empty((n, 64, 2), np.float32) if geometry else None
        check(lib().emsph_get_neighborhood(self._h, which, _ptr(counts), _ptr(idx), _ptr(dist),
                                           _ptr(vec)), "get_neighborhood")
        return counts, idx, dist, vec

    def set_force_model(self, model: ForceModel) -> None:
        """Device external-force models (emsph_set_force_model)."""
        m = model.to_c()
        check(lib().emsph_set_force_model(self._h, ctypes.byref(m)), "set_force_model")

    def force_model(self) -> ForceModel:
        m = _capi.ForceModel()
        check(lib().emsph_get_force_model(self._h, ctypes.byref(m)), "get_force_model")
        return ForceModel.from_c(m)

    def set_user_forces(self, fn) -> None:
        """A host User_forces_function (see user_forces_callback), called once
        per sub-step before the device models; None removes it."""
        self._user_cb = user_forces_callback(fn) if fn else None
        check(lib().emsph_set_user_forces(self._h, self._user_cb or _capi.USER_FORCES_FN(),
                                          None), "set_user_forces")

    def save_snapshot(self, path: str, time_flicks: int = 0,
                      method: Optional[str] = None) -> None:
        check(lib().emsph_save_snapshot(self._h, path.encode(), int(time_flicks),
                                        _capi.METHODS.get(method, -1)), "save_snapshot")

    def load_snapshot(self, path: str):
        """Loads the state; returns (time in flicks, method name or None)."""
        t = ctypes.c_int64()
        m = ctypes.c_int32()
        check(lib().emsph_load_snapshot(self._h, path.encode(), ctypes.byref(t), ctypes.byref(m)),
              "load_snapshot")
        names = {v: k for k, v in _capi.METHODS.items()}
        return int(t.value), names.get(int(m.value))

    def set_user_colors(self, fn) -> None:
        """A host User_colors_function (see user_colors_callback); None restores
        target_density_colors on the device."""
        self._colors_cb = user_colors_callback(fn) if fn else None
        check(lib().emsph_set_user_colors(self._h, self._colors_cb or _capi.USER_COLORS_FN(),
                                          None), "set_user_colors")

    def compute_colors(self, time_seconds: float) -> None:
        check(lib().emsph_compute_colors(self._h, float(time_seconds)), "compute_colors")

    def colors(self) -> np.ndarray:
        """State::colors of the last frame (RGBA float32, original order)."""
        out = np.empty((self.n, 4), np.float32)
        check(lib().emsph_get_colors(self._h, _ptr(out)), "get_colors")
        return out

    def colors_u8(self) -> np.ndarray:
        """convert_colors (colors.cpp:57-70) of State::colors."""
        out = np.empty((self.n, 4), np.uint8)
        check(lib().emsph_get_colors_u8(self._h, _ptr(out)), "get_colors_u8")
        return out

    def frame_init(self, method: str) -> None:
        check(lib().emsph_frame_init(self._h, _capi.METHODS[method]), "frame_init")

    def sub_step(self, method: str, t: float, dt: float) -> _capi.StepStats:
        st = _capi.StepStats()
        check(lib().emsph_sub_step(self._h, _capi.METHODS[method], t, dt, ctypes.byref(st)),
              "sub_step")
        return st

    def frame_step(self, method: str, global_time_flicks: int) -> _capi.FrameStats:
        fs = _capi.FrameStats()
        check(lib().emsph_frame_step(self._h, _capi.METHODS[method], int(global_time_flicks),
                                     ctypes.byref(fs)), "frame_step")
        return fs

    def run_sub_steps(self, method: str, dt: float, n: int):
        ms = ctypes.c_float()
        st = _capi.StepStats()
        check(lib().emsph_run_sub_steps(self._h, _capi.METHODS[method], dt, int(n),
                                        ctypes.byref(ms), ctypes.byref(st)), "run_sub_steps")
        return float(ms.value), st

    def set_cfl(self, factor: float) -> None:
        """Per-sub-step dt = min(dt, factor * h / max |v|) on the device (0: off)."""
        check(lib().emsph_set_cfl(self._h, float(factor)), "set_cfl")

    def set_graphs(self, enable: bool) -> None:
        """Replay the fixed segments of each sub-step from captured HIP graphs."""
        check(lib().emsph_set_graphs(self._h, 1 if enable else 0), "set_graphs")

    def set_skin_lists(self, enable: bool) -> None:
        """Skin candidate lists for the neighbour search (emsph_set_skin_lists):
        the same lists, found from candidates recorded at a rebuild while no
        particle has moved h/2 since."""
        check(lib().emsph_set_skin_lists(self._h, 1 if enable else 0), "set_skin_lists")

    def skin_lists(self) -> tuple:
        """(on, sub-steps that rebuilt candidates, sub-steps that reused them)."""
        on, r, u = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().emsph_get_skin_lists(self._h, ctypes.byref(on), ctypes.byref(r),
                                         ctypes.byref(u)), "get_skin_lists")
        return bool(on.value), int(r.value), int(u.value)

    def tuning(self) -> dict:
        t = _capi.Tuning()
        check(lib().emsph_get_tuning(self._h, ctypes.byref(t)), "get_tuning")
        return t.as_dict()

    def set_tuning(self, persistent_loop: bool = False, graphs: bool = True,
                   neighbor_batch: int = 4, xcd_remap: bool = True) -> None:
        t = _capi.Tuning(int(persistent_loop), int(graphs), int(neighbor_batch), int(xcd_remap))
        check(lib().emsph_set_tuning(self._h, ctypes.byref(t)), "set_tuning")

    def autotune(self, method: str, dt: float, trial_sub_steps: int = 5):
        """Times each candidate tuning on this state (which it leaves unchanged)
        and keeps the fastest; returns (chosen, [ms per candidate])."""
        t = _capi.Tuning()
        times = (ctypes.c_double * 8)()
        n = ctypes.c_int()
        check(lib().emsph_autotune(self._h, _capi.METHODS[method], float(dt), int(trial_sub_steps),
                                   ctypes.byref(t), times, ctypes.byref(n)), "autotune")
        return (t.as_dict(), list(times[:n.value]))

    def iteration_totals(self, reset: bool = False) -> tuple:
        """(sub-steps, pressure or density iterations, divergence iterations,
        sub-steps redone after a short sort) since creation or the last reset."""
        out = (ctypes.c_int64 * 4)()
        check(lib().emsph_get_iteration_totals(self._h, 1 if reset else 0, out),
              "get_iteration_totals")
        return tuple(int(x) for x in out)

    def set_kernel_timing(self, enable: bool) -> None:
        check(lib().emsph_set_kernel_timing(self._h, 1 if enable else 0), "set_kernel_timing")

    def kernel_timing(self) -> dict:
        return _kernel_timing(lib().emsph_get_kernel_timing, self._h)


def _kernel_timing(fn, h) -> dict:
    """{kernel class: (total ms, launches)} from the engine's HIP events."""
    m = 64
    names = (ctypes.c_char_p * m)()
    ms = (ctypes.c_double * m)()
    calls = (ctypes.c_int64 * m)()
    k = fn(h, names, ms, calls, m)
    if k < 0:
        raise _capi.EmsphError(-1, "kernel_timing", "reading the timing events failed")
    return {names[i].decode(): (ms[i], calls[i]) for i in range(k)}


def _simulation_step(method: str, global_time: int, config: Simulation_config, state: State,
                     solid_state: Solid_state, ctx: Optional[Context] = None) -> State:
    own = ctx is None
    if own:
        ctx = Context(config, solid_state, len(state.positions))
    try:
        ctx.set_state(state)
        ctx.frame_step(method, global_time)
        out = ctx.get_state()
        out.colors = ctx.colors()
        return out
    finally:
        if own:
            ctx.close()


def iisph_simulation_step(global_time, config, state, solid_state, ctx=None) -> State:
    """iisph.cpp:168-188 (adaptive 60/6, CFL 0.2)."""
    return _simulation_step("iisph", global_time, config, state, solid_state, ctx)


def pcisph_simulation_step(global_time, config, state, solid_state, ctx=None) -> State:
    """pcisph.cpp:166-204 (fixed sub-steps of config.seconds_per_sub_step)."""
    return _simulation_step("pcisph", global_time, config, state, solid_state, ctx)


def dfsph_p_simulation_step(global_time, config, state, solid_state, ctx=None) -> State:
    """dfsph_p.cpp:376-396 (adaptive 60/4, CFL 0.4)."""
    return _simulation_step("dfsph_p", global_time, config, state, solid_state, ctx)


class EZ_EXAMPLE_SIM:  # noqa: N801 - the reference's type name
    """simulation.h:24-44: parameters -> solids -> fluid -> step(method)."""

    def __init__(self, params: Parameters, solid_initial_state=world_walls_initial_solid_state,
                 fluid_initial_state=dam_break_initial_state, config: Simulation_config = None,
                 user_forces=None):
        if params.method not in ("pcisph", "iisph", "iisph_ap", "iisph_pseudo_ap", "dfsph_p"):
            # simulation.cpp:37 EMLD_FAIL
            raise ValueError("Unknown simulation method: " + params.method)
        self.params = params
        self.config = config or Simulation_config.from_parameters(params)
        self.solid_state = solid_initial_state(params)
        self.state = fluid_initial_state(params, self.solid_state)
        self.method = params.method
        self.time = 0  # flicks
        self.ctx = Context(self.config, self.solid_state, max(len(self.state.positions), 1))
        self.ctx.set_state(self.state)
        # user_forces: default_gravity_forces(params.gravity) when None
        # (simulation.cpp:14-24); a ForceModel runs on the device, a callable
        # is a host User_forces_function that replaces gravity.
        if isinstance(user_forces, ForceModel):
            self.ctx.set_force_model(user_forces)
        elif user_forces is not None:
            self.ctx.set_force_model(ForceModel(gravity_on=False, gravity=params.gravity))
            self.ctx.set_user_forces(user_forces)

    def step(self, override_method: Optional[str] = None) -> _capi.FrameStats:
        fs = self.ctx.frame_step(override_method or self.method, self.time)
        self.time += self.params.time_per_step
        return fs

    def fetch_state(self) -> State:
        self.state = self.ctx.get_state()
        self.state.colors = self.ctx.colors()
        return self.state


class Batch:
    """Many independent scenes stepped together (emsph_batch): one launch per
    sub-step stage for all scenes; each scene keeps its own parameters,
    solids, particle count, dt and solver iteration count."""

    def __init__(self, configs, solid_states, states):
        m = len(configs)
        assert m == len(solid_states) == len(states) and m > 0
        cfgs = (_capi.Config * m)(*[c.to_c() for c in configs])
        self._solid = [np.ascontiguousarray(s.positions, np.float32).reshape(-1, 2)
                       for s in solid_states]
        sptr = (ctypes.c_void_p * m)(*[_ptr(a) for a in self._solid])
        scnt = (ctypes.c_size_t * m)(*[len(a) for a in self._solid])
        fcnt = (ctypes.c_size_t * m)(*[len(s.positions) for s in states])
        h = ctypes.c_void_p()
        check(lib().emsph_batch_create(m, cfgs, sptr, scnt, fcnt, ctypes.byref(h)),
              "emsph_batch_create")
        self._h = h
        self.m = m
        self.counts = [len(s.positions) for s in states]
        for i, s in enumerate(states):
            self.set_state(i, s)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().emsph_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, i: int, state: State) -> None:
        pos = np.ascontiguousarray(state.positions, np.float32).reshape(-1, 2)
        vel = None if state.velocities is None else np.ascontiguousarray(
            state.velocities, np.float32).reshape(-1, 2)
        check(lib().emsph_batch_set_state(self._h, i, _ptr(pos), _ptr(vel)), "batch_set_state")

    def get_state(self, i: int) -> State:
        n = self.counts[i]
        pos = np.empty((n, 2), np.float32)
        vel = np.empty((n, 2), np.float32)
        check(lib().emsph_batch_get_state(self._h, i, _ptr(pos), _ptr(vel)), "batch_get_state")
        return State(pos, vel)

    def field(self, i: int, name: str) -> np.ndarray:
        fid, comps = _capi.FIELDS[name]
        out = np.empty((self.counts[i], comps), np.float32)
        check(lib().emsph_batch_get_field(self._h, i, fid, _ptr(out)), "batch_get_field")
        return out[:, 0].copy() if comps == 1 else out

    def set_force_model(self, i: int, model: ForceModel) -> None:
        m = model.to_c()
        check(lib().emsph_batch_set_force_model(self._h, i, ctypes.byref(m)),
              "batch_set_force_model")

    def set_cfl(self, factor: float) -> None:
        """Per-scene, per-sub-step dt = min(config dt, factor * h / max |v|) (0: off)."""
        check(lib().emsph_batch_set_cfl(self._h, float(factor)), "batch_set_cfl")

    def set_graphs(self, enable: bool) -> None:
        check(lib().emsph_batch_set_graphs(self._h, 1 if enable else 0), "batch_set_graphs")

    def set_skin_lists(self, enable: bool) -> None:
        """As Context.set_skin_lists, for every scene of the batch."""
        check(lib().emsph_batch_set_skin_lists(self._h, 1 if enable else 0),
              "batch_set_skin_lists")

    def skin_lists(self) -> tuple:
        on, r, u = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().emsph_batch_get_skin_lists(self._h, ctypes.byref(on), ctypes.byref(r),
                                               ctypes.byref(u)), "batch_get_skin_lists")
        return bool(on.value), int(r.value), int(u.value)

    def set_kernel_timing(self, enable: bool) -> None:
        check(lib().emsph_batch_set_kernel_timing(self._h, 1 if enable else 0),
              "batch_set_kernel_timing")

    def kernel_timing(self) -> dict:
        return _kernel_timing(lib().emsph_batch_get_kernel_timing, self._h)

    def iteration_totals(self, reset: bool = False) -> tuple:
        """As Context.iteration_totals, over the batch's sub-steps."""
        out = (ctypes.c_int64 * 4)()
        check(lib().emsph_batch_get_iteration_totals(self._h, 1 if reset else 0, out),
              "batch_get_iteration_totals")
        return tuple(int(x) for x in out)

    def set_tuning(self, persistent_loop: bool = False, graphs: bool = True,
                   neighbor_batch: int = 4, xcd_remap: bool = True) -> None:
        t = _capi.Tuning(int(persistent_loop), int(graphs), int(neighbor_batch), int(xcd_remap))
        check(lib().emsph_batch_set_tuning(self._h, ctypes.byref(t)), "batch_set_tuning")

    def autotune(self, method: str, trial_sub_steps: int = 3):
        t = _capi.Tuning()
        times = (ctypes.c_double * 8)()
        n = ctypes.c_int()
        check(lib().emsph_batch_autotune(self._h, _capi.METHODS[method], int(trial_sub_steps),
                                         ctypes.byref(t), times, ctypes.byref(n)),
              "batch_autotune")
        return (t.as_dict(), list(times[:n.value]))

    def frame_init(self, method: str) -> None:
        check(lib().emsph_batch_frame_init(self._h, _capi.METHODS[method]), "batch_frame_init")

    def sub_step(self, method: str):
        st = (_capi.StepStats * self.m)()
        check(lib().emsph_batch_sub_step(self._h, _capi.METHODS[method], st), "batch_sub_step")
        return list(st)

    def run_sub_steps(self, method: str, n: int) -> float:
        ms = ctypes.c_float()
        check(lib().emsph_batch_run_sub_steps(self._h, _capi.METHODS[method], int(n),
                                              ctypes.byref(ms)), "batch_run_sub_steps")
        return float(ms.value)


class Pool:
    """Scenes that come and go, stepped together (emsph_pool): one solver per
    pool; adds and removes are staged and applied by commit() between
    sub-steps. Capacity (the HBM budget) is allocated at creation."""

    def __init__(self, fluid_capacity: int, solid_capacity: int, max_scenes: int, method: str):
        h = ctypes.c_void_p()
        before = lib().emsph_device_bytes()
        check(lib().emsph_pool_create(int(fluid_capacity), int(solid_capacity), int(max_scenes),
                                      _capi.METHODS[method], ctypes.byref(h)), "emsph_pool_create")
        self._h = h
        self.method = method
        self.device_bytes = lib().emsph_device_bytes() - before
        self.counts = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().emsph_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, config: Simulation_config, solid_state: Solid_state, state: State) -> int:
        """Stage a scene; returns its id. Raises EmsphError (EMSPH_ERR_CAPACITY)
        when it does not fit."""
        c = config.to_c()
        sol = np.ascontiguousarray(solid_state.positions, np.float32).reshape(-1, 2)
        pos = np.ascontiguousarray(state.positions, np.float32).reshape(-1, 2)
        vel = None if state.velocities is None else np.ascontiguousarray(
            state.velocities, np.float32).reshape(-1, 2)
        sid = ctypes.c_int()
        check(lib().emsph_pool_add(self._h, ctypes.byref(c), _ptr(sol), len(sol), _ptr(pos),
                                   _ptr(vel), len(pos), ctypes.byref(sid)), "pool_add")
        self.counts[sid.value] = len(pos)
        return sid.value

    def remove(self, sid: int) -> None:
        check(lib().emsph_pool_remove(self._h, int(sid)), "pool_remove")

    def commit(self) -> None:
        check(lib().emsph_pool_commit(self._h), "pool_commit")

    def live(self) -> list:
        n = lib().emsph_pool_live(self._h, None, 0)
        ids = (ctypes.c_int * max(n, 1))()
        lib().emsph_pool_live(self._h, ids, n)
        return list(ids[:n])

    def set_cfl(self, factor: float) -> None:
        check(lib().emsph_pool_set_cfl(self._h, float(factor)), "pool_set_cfl")

    def set_tuning(self, persistent_loop: bool = False, graphs: bool = True,
                   neighbor_batch: int = 4, xcd_remap: bool = True) -> None:
        t = _capi.Tuning(int(persistent_loop), int(graphs), int(neighbor_batch), int(xcd_remap))
        check(lib().emsph_pool_set_tuning(self._h, ctypes.byref(t)), "pool_set_tuning")

    def sub_step(self, with_stats: bool = True):
        """One sub-step of every live scene; stats in live() order."""
        if not with_stats:
            check(lib().emsph_pool_sub_step(self._h, None), "pool_sub_step")
            return None
        n = lib().emsph_pool_live(self._h, None, 0)
        st = (_capi.StepStats * max(n, 1))()
        check(lib().emsph_pool_sub_step(self._h, st), "pool_sub_step")
        return list(st[:n])

    def save_snapshot(self, sid: int, path: str, time_flicks: int = 0) -> None:
        check(lib().emsph_pool_save_snapshot(self._h, int(sid), path.encode(), int(time_flicks)),
              "pool_save_snapshot")

    def iteration_totals(self, reset: bool = False) -> tuple:
        """As Context.iteration_totals, over the pool's sub-steps (iterations
        launched: the maximum over the live scenes of each sub-step)."""
        out = (ctypes.c_int64 * 4)()
        check(lib().emsph_pool_get_iteration_totals(self._h, 1 if reset else 0, out),
              "pool_get_iteration_totals")
        return tuple(int(x) for x in out)

    def get_state(self, sid: int) -> State:
        n = self.counts[sid]
        pos = np.empty((n, 2), np.float32)
        vel = np.empty((n, 2), np.float32)
        check(lib().emsph_pool_get_state(self._h, int(sid), _ptr(pos), _ptr(vel)), "pool_get_state")
        return State(pos, vel)

