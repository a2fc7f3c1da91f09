"""ctypes access to the CPU oracle (oracle/liboracle_sph.so).

TEST INFRASTRUCTURE ONLY. Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module; the product path
(``emerald_amd``) never does. The oracle restates the reference's CPU algorithm
(see the header of ``sph_oracle.cpp`` for the file:line map).

Parity status of each oracle function (what pins it):
  * z_index / single-axis spread: reference known-answer tests
    (z_index/tests/test_z_index.cpp:11-70) and the reference's own compiled
    z_index.cpp (oracle/_ref), exhaustively on sampled coordinates.
  * grid coords: reference KATs (sph_common/tests/test_common.cpp:42-63).
  * sort / block indices / blocks: closed forms (stable sort by key,
    run-length ranges) and the invariants of test_neighborhood.cpp:62-175.
  * neighbourhoods: O(N^2) brute force on the reference fixture
    (test_neighborhood.cpp:177-292, mt19937_64{17272}); ordering and the
    64-cap are pinned by closed-form checks (sortedness by (distance, index),
    truncation equal to the first 64 accepted in scan order).
  * kernels W/dW/GradW: closed forms evaluated in FP64 (W(0)=1/(pi h^3),
    W(h)=0.25/(pi h^3), W(2h)=0, dW continuity, GradW antisymmetry).
  * safe_divide: reference KATs (util/tests/test_safe_divide.cpp:40-71,146-182).
  * density / divergence / pressure / IISPH / PCISPH / DFSPH ops: no reference
    test pins them; pinned by invariants (uniform lattice density, pairwise
    antisymmetry, v=0 -> div=0, p=0 -> s=0 and beta=0) and by an independent
    FP64 brute-force restatement on small scenes (tests/test_oracle_solvers.py).
  * Imath Vec2::length() (IlmBase 2.5 ImathVec.h, absent from the image),
    including the lengthTiny branch for |v|^2 < 2*FLT_MIN: the IEEE root of
    |v|^2 above the threshold; below it, within 2 float ulps of the exact
    length where x*x underflows, exact on axis vectors and under power-of-two
    scaling, checks a plain sqrtf(x*x + y*y) fails (tests/test_oracle_kats.py).
  * scene inputs: world_walls_initial_solid_state with its random obstacles
    (solids.cpp:139-206) and dam_break_initial_state (initial_state.cpp:14-97):
    the C++ standard's mt19937_64 known answer, a pure-Python restatement of
    the float draws and the boxes, and an all-pairs kill test
    (tests/test_oracle_scene.py).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sph.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libzindex_ref.so")

METHODS = {"pcisph": 0, "iisph": 1, "iisph_ap": 2, "iisph_pseudo_ap": 3, "dfsph_p": 4}


def build(force: bool = False) -> None:
    """Compile the oracle with oracle/Makefile (g++, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"],
                       check=True)


def build_ref() -> bool:
    """Compile oracle/_ref from the reference tree when it is mounted."""
    if not os.path.isdir("/root/reference/emerald/z_index"):
        return os.path.exists(_REF_PATH)
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "ref"], check=True)
    return os.path.exists(_REF_PATH)


class Stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("divergence_iterations", ctypes.c_int32),
        ("error_average", ctypes.c_float),
        ("error_max", ctypes.c_float),
        ("divergence_error_average", ctypes.c_float),
        ("divergence_error_max", ctypes.c_float),
    ]


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("support", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("viscosity", ctypes.c_double),
        ("target_density", ctypes.c_double),
        ("time_per_step", ctypes.c_int64),
        ("sub_steps", ctypes.c_int32),
        ("iisph_max_iterations", ctypes.c_int32),
        ("iisph_error_average_threshold", ctypes.c_double),
        ("iisph_error_max_threshold", ctypes.c_double),
        ("iisph_omega", ctypes.c_double),
        ("dfsph_max_iterations", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("dfsph_density_avg", ctypes.c_double),
        ("dfsph_density_max", ctypes.c_double),
        ("dfsph_divergence_avg", ctypes.c_double),
        ("dfsph_divergence_max", ctypes.c_double),
        ("seconds_per_sub_step", ctypes.c_double),
        ("mass_per_particle", ctypes.c_double),
        ("pressure_correction_denom", ctypes.c_double),
    ]


_lib = None
_ref = None

_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_F = ctypes.c_float
_D = ctypes.c_double
_I = ctypes.c_int
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_I64 = ctypes.c_int64


# User_forces_function (forces.h:13-21) as a C callback:
# (user, count, t, dt, mass, radius, forces_xy, pos_xy, vel_xy).
USER_FORCES_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                  ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(ctypes.c_float))

FORCE_GRAVITY, FORCE_POLE, FORCE_DRAG, FORCE_ANTI_COUPLING = 1, 2, 4, 8


def wrap_user_forces(fn):
    """A C callback around fn(count, t, dt, mass, radius, forces, pos, vel),
    where forces/pos/vel are (count, 2) float32 views (forces is written)."""
    def tramp(_user, count, t, dt, mass, radius, f, p, v):
        n = int(count)
        fa = np.ctypeslib.as_array(f, shape=(n, 2))
        pa = np.ctypeslib.as_array(p, shape=(n, 2))
        va = np.ctypeslib.as_array(v, shape=(n, 2))
        fn(n, float(t), float(dt), float(mass), float(radius), fa, pa, va)
    return USER_FORCES_FN(tramp)


def anti_coupling_angle(i: int) -> float:
    """forces.cpp:134-141: the angle drawn for particle i (host C++ library)."""
    return float(lib().orc_anti_coupling_angle(int(i)))


def lehmer_first_draw(seed: int) -> int:
    """util/random.h:112-134: the first draw of Lehmer_rand_gen_64{seed}."""
    return int(lib().orc_lehmer_first_draw(int(seed)))


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_z_index": (_U64, [_U32, _U32]),
            "orc_single_axis_index_2": (_U64, [_U32]),
            "orc_length": (_F, [_F, _F]),
            "orc_W": (_F, [_F, _F]),
            "orc_dW": (_F, [_F, _F]),
            "orc_GradW": (None, [_F, _F, _F, _P]),
            "orc_is_safe_divide_f": (_I, [_F, _F]),
            "orc_is_safe_divide_v2f": (_I, [_F, _F, _F]),
            "orc_is_safe_divide_d": (_I, [_D, _D]),
            "orc_grid_coords": (None, [_SZ, _F, _P, _P]),
            "orc_spatial_index": (_SZ, [_SZ, _F, _P, _P, _P, _P, _P, _P, _P, _P]),
            "orc_neighborhoods": (None, [_SZ, _P, _SZ, _P, _I, _F, _P, _P, _P, _P]),
            "orc_simulation_config": (None, [_F, _F, _I64, _I, _P]),
            "orc_estimate_solid_box_emission_count": (_SZ, [_F, _F, _F, _F, _F]),
            "orc_emit_solid_box": (_SZ, [_F, _F, _F, _F, _F, _SZ, _P]),
            "orc_world_walls_positions": (_SZ, [_F, _F, _I, _P, _SZ]),
            "orc_search_seconds": (_D, [_SZ, _P, _F, _P]),
            "orc_dam_break_positions": (_SZ, [_F, _F, _SZ, _P, _P, _SZ]),
            "orc_mt19937_64_uniform_float": (None, [_U64, _F, _F, _SZ, _P]),
            "orc_scene_create": (_P, [_P, _SZ, _P, _SZ, _P, _P, _I]),
            "orc_scene_destroy": (None, [_P]),
            "orc_set_threads": (None, [_I]),
            "orc_get_threads": (_I, []),
            "orc_scene_frame_init": (None, [_P, _I]),
            "orc_scene_sub_step": (None, [_P, _I, _D, _D, _P]),
            "orc_scene_frame_step": (_I, [_P, _I, _I64, _P]),
            "orc_scene_solid_count": (_SZ, [_P]),
            "orc_scene_solid_volumes": (None, [_P, _P]),
            "orc_scene_get_field": (_I, [_P, ctypes.c_char_p, _P]),
            "orc_scene_get_state_f32": (None, [_P, _P, _P]),
            "orc_scene_get_state_f64": (None, [_P, _P, _P]),
            "orc_scene_get_neighborhood": (None, [_P, _I, _P, _P]),
            "orc_fixture_points_mt19937_64": (None, [_U64, _SZ, _F, _F, _P]),
            "orc_anti_coupling_angle": (_F, [_U64]),
            "orc_lehmer_first_draw": (_U64, [_U64]),
            "orc_scene_set_forces": (_I, [_P, _U32, _P, USER_FORCES_FN, _P]),
            "orc_scene_get_colors": (_I, [_P, _P, _P]),
            "orc_convert_colors": (None, [_SZ, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's own compiled z_index (None when it cannot be built)."""
    global _ref
    if _ref is None:
        if not os.path.exists(_REF_PATH):
            try:
                if not build_ref():
                    return None
            except (OSError, subprocess.CalledProcessError):
                return None
        L = ctypes.CDLL(_REF_PATH)
        for name, args in (("ref_z_index_2", [_U32, _U32]),
                           ("ref_single_axis_index_2", [_U32]),
                           ("ref_single_axis_template_2", [_U32])):
            getattr(L, name).restype = _U64
            getattr(L, name).argtypes = args
        _ref = L
    return _ref


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def fixture_points(seed: int, n: int, lo: float, hi: float) -> np.ndarray:
    out = np.empty((n, 2), np.float32)
    lib().orc_fixture_points_mt19937_64(seed, n, lo, hi, _ptr(out))
    return out


def grid_coords(xy: np.ndarray, cell: float) -> np.ndarray:
    xy = np.ascontiguousarray(xy, np.float32)
    out = np.empty((len(xy), 2), np.int32)
    lib().orc_grid_coords(len(xy), cell, _ptr(xy), _ptr(out))
    return out


def spatial_index(xy: np.ndarray, cell: float) -> dict:
    xy = np.ascontiguousarray(xy, np.float32)
    n = len(xy)
    out = {
        "grid": np.empty((n, 2), np.int32),
        "keys": np.empty(n, np.uint64),
        "sorted_keys": np.empty(n, np.uint64),
        "sorted_idx": np.empty(n, np.uint32),
        "block_indices": np.empty(n, np.uint32),
        "block_begin": np.empty(n, np.uint32),
        "block_end": np.empty(n, np.uint32),
    }
    nb = lib().orc_spatial_index(n, cell, _ptr(xy), *(_ptr(out[k]) for k in (
        "grid", "keys", "sorted_keys", "sorted_idx", "block_indices", "block_begin",
        "block_end")))
    out["block_begin"] = out["block_begin"][:nb].copy()
    out["block_end"] = out["block_end"][:nb].copy()
    out["block_count"] = int(nb)
    return out


def neighborhoods(self_xy, other_xy, same_array: bool, max_distance: float):
    self_xy = np.ascontiguousarray(self_xy, np.float32)
    other_xy = np.ascontiguousarray(other_xy, np.float32)
    n = len(self_xy)
    counts = np.empty(n, np.uint8)
    idx = np.zeros((n, 64), np.uint32)
    dist = np.zeros((n, 64), np.float32)
    vec = np.zeros((n, 64, 2), np.float32)
    lib().orc_neighborhoods(n, _ptr(self_xy), len(other_xy), _ptr(other_xy),
                            1 if same_array else 0, max_distance, _ptr(counts),
                            _ptr(idx), _ptr(dist), _ptr(vec))
    return counts, idx, dist, vec


def search_seconds(xy: np.ndarray, cell: float):
    """(seconds, neighbour pairs) of one neighbour search of the points: the
    reference's timed region of test_neighborhood.cpp:178-231."""
    xy = np.ascontiguousarray(xy, np.float32)
    pairs = ctypes.c_uint64()
    sec = lib().orc_search_seconds(len(xy), _ptr(xy), cell, ctypes.byref(pairs))
    return float(sec), int(pairs.value)


def simulation_config(support: float, rho0: float, time_per_step: int,
                      sub_steps: int) -> dict:
    out = np.empty(4, np.float32)
    lib().orc_simulation_config(support, rho0, time_per_step, sub_steps, _ptr(out))
    return {"mass_per_particle": out[0], "seconds_per_sub_step": out[1],
            "draw_radius": out[2], "pressure_correction_denom": out[3]}


def emit_solid_box(box, R: float) -> np.ndarray:
    (x0, y0), (x1, y1) = box
    n = lib().orc_estimate_solid_box_emission_count(x0, y0, x1, y1, R)
    out = np.empty((max(n, 1), 2), np.float32)
    m = lib().orc_emit_solid_box(x0, y0, x1, y1, R, n, _ptr(out))
    return out[:m].copy()


def world_walls_positions(support: float, length: float = 1.0,
                          random_obstacles: bool = True) -> np.ndarray:
    """world_walls_initial_solid_state positions (solids.cpp:139-206)."""
    n = lib().orc_world_walls_positions(support, length, int(random_obstacles), None, 0)
    out = np.empty((max(n, 1), 2), np.float32)
    lib().orc_world_walls_positions(support, length, int(random_obstacles), _ptr(out), n)
    return out[:n].copy()


def dam_break_positions(support: float, solid_xy, length: float = 1.0) -> np.ndarray:
    """dam_break_initial_state positions (initial_state.cpp:14-97)."""
    s = np.ascontiguousarray(solid_xy, np.float32).reshape(-1, 2)
    n = lib().orc_dam_break_positions(support, length, len(s), _ptr(s), None, 0)
    out = np.empty((max(n, 1), 2), np.float32)
    lib().orc_dam_break_positions(support, length, len(s), _ptr(s), _ptr(out), n)
    return out[:n].copy()


def uniform_float_draws(seed: int, a: float, b: float, n: int) -> np.ndarray:
    """n draws of std::uniform_real_distribution<float>{a, b} over
    std::mt19937_64{seed} (the host C++ library's)."""
    out = np.empty(n, np.float32)
    lib().orc_mt19937_64_uniform_float(seed, a, b, n, _ptr(out))
    return out


class Scene:
    """One oracle simulation (float = reference precision, fp64 = FP64 build)."""

    def __init__(self, cfg: dict, solid_xy, fluid_xy, fluid_vel=None, fp64=False):
        c = OrcConfig()
        for name, _ in OrcConfig._fields_:
            if name in cfg:
                setattr(c, name, cfg[name])
        self._cfg = c
        solid_xy = np.ascontiguousarray(solid_xy, np.float32).reshape(-1, 2)
        fluid_xy = np.ascontiguousarray(fluid_xy, np.float32).reshape(-1, 2)
        self.n = len(fluid_xy)
        vel_p = None
        if fluid_vel is not None:
            fluid_vel = np.ascontiguousarray(fluid_vel, np.float32).reshape(-1, 2)
            vel_p = _ptr(fluid_vel)
        self._h = lib().orc_scene_create(ctypes.byref(c), len(solid_xy), _ptr(solid_xy),
                                         self.n, _ptr(fluid_xy), vel_p, 1 if fp64 else 0)
        self.fp64 = fp64

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_scene_destroy(h)
            self._h = None

    def set_forces(self, flags: int = FORCE_GRAVITY, gravity: float = 9.81,
                   pole_magnitude: float = 0.0, pole=(0.0, 0.0), drag_magnitude: float = 0.0,
                   drag_diameter: float = 0.0, anti_coupling_max_distance: float = 0.0,
                   anti_coupling_magnitude: float = 0.0, draw_radius: float = 0.0,
                   user_forces=None) -> None:
        """External force model (forces.cpp:11-149): user_forces (a Python
        callable, see wrap_user_forces) first, then the flagged models."""
        v = np.array([gravity, pole_magnitude, pole[0], pole[1], drag_magnitude, drag_diameter,
                      anti_coupling_max_distance, anti_coupling_magnitude, draw_radius],
                     np.float64)
        self._user_cb = wrap_user_forces(user_forces) if user_forces else USER_FORCES_FN()
        if lib().orc_scene_set_forces(self._h, int(flags), _ptr(v), self._user_cb, None):
            raise ValueError("user force callbacks need a float (not fp64) scene")

    def frame_init(self, method: str) -> None:
        lib().orc_scene_frame_init(self._h, METHODS[method])

    def sub_step(self, method: str, t: float, dt: float) -> Stats:
        st = Stats()
        lib().orc_scene_sub_step(self._h, METHODS[method], t, dt, ctypes.byref(st))
        return st

    def frame_step(self, method: str, global_time_flicks: int):
        st = Stats()
        n = lib().orc_scene_frame_step(self._h, METHODS[method], global_time_flicks,
                                       ctypes.byref(st))
        return n, st

    def field(self, name: str) -> np.ndarray:
        out = np.empty((self.n, 2), np.float64)
        k = lib().orc_scene_get_field(self._h, name.encode(), _ptr(out))
        if k < 0:
            raise KeyError(name)
        return out.ravel()[:self.n].copy() if k == 1 else out

    def colors(self):
        """State::colors after the last frame step: (RGBA floats, bytes)."""
        rgba = np.empty((self.n, 4), np.float32)
        rgba8 = np.empty((self.n, 4), np.uint8)
        if not lib().orc_scene_get_colors(self._h, _ptr(rgba), _ptr(rgba8)):
            raise RuntimeError("no frame step has computed colors")
        return rgba, rgba8

    def state_f32(self):
        pos = np.empty((self.n, 2), np.float32)
        vel = np.empty((self.n, 2), np.float32)
        lib().orc_scene_get_state_f32(self._h, _ptr(pos), _ptr(vel))
        return pos, vel

    def state_f64(self):
        """The state in the scene's own precision, widened to float64."""
        pos = np.empty((self.n, 2), np.float64)
        vel = np.empty((self.n, 2), np.float64)
        lib().orc_scene_get_state_f64(self._h, _ptr(pos), _ptr(vel))
        return pos, vel

    def neighborhood(self, which: int = 0):
        counts = np.empty(self.n, np.uint8)
        idx = np.empty((self.n, 64), np.uint32)
        lib().orc_scene_get_neighborhood(self._h, which, _ptr(counts), _ptr(idx))
        return counts, idx

    def solid_volumes(self) -> np.ndarray:
        n = lib().orc_scene_solid_count(self._h)
        out = np.empty(n, np.float64)
        lib().orc_scene_solid_volumes(self._h, _ptr(out))
        return out


def convert_colors(rgba: np.ndarray) -> np.ndarray:
    """colors.cpp:57-70 on float RGBA -> uint8 RGBA."""
    rgba = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
    out = np.empty(rgba.shape, np.uint8)
    lib().orc_convert_colors(len(rgba), _ptr(rgba), _ptr(out))
    return out


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())
