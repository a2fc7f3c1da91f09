"""ctypes binding of include/emerald_sph.h (libemerald_sph.so, built for gfx950).

The library is loaded from this package directory only; there is no fallback:
if it is missing or cannot be loaded, importing the binding raises.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemerald_sph.so")

OK = 0
STATUS = {
    0: "EMSPH_OK",
    1: "EMSPH_ERR_INVALID_ARGUMENT",
    2: "EMSPH_ERR_OUT_OF_MEMORY",
    3: "EMSPH_ERR_HIP",
    4: "EMSPH_ERR_UNKNOWN_METHOD",
    5: "EMSPH_ERR_CAPACITY",
    6: "EMSPH_ERR_NO_DEVICE",
}

PCISPH, IISPH, IISPH_AP, IISPH_PSEUDO_AP, DFSPH_P = 0, 1, 2, 3, 4
METHODS = {"pcisph": PCISPH, "iisph": IISPH, "iisph_ap": IISPH_AP,
           "iisph_pseudo_ap": IISPH_PSEUDO_AP, "dfsph_p": DFSPH_P}

FIELDS = {
    "density": (0, 1), "density_star": (1, 1), "pressure": (2, 1), "aii": (3, 1),
    "dii": (4, 2), "sum_dij_pj": (5, 2), "alpha": (6, 1), "kappa_density": (7, 1),
    "kappa_divergence": (8, 1), "pressure_force": (9, 2), "pressure_acceleration": (10, 2),
    "external_force": (11, 2),
}


class Parameters(ctypes.Structure):
    """emsph_parameters (sph2d_box/parameters.h:12-43)."""
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("time_per_step", ctypes.c_int64),
        ("sub_steps", ctypes.c_int32),
        ("length", ctypes.c_float),
        ("support", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("viscosity", ctypes.c_float),
        ("target_density", ctypes.c_float),
        ("iisph_max_pressure_iterations", ctypes.c_int32),
        ("iisph_error_average_threshold", ctypes.c_float),
        ("iisph_error_max_threshold", ctypes.c_float),
        ("iisph_omega", ctypes.c_float),
        ("dfsph_max_correction_iterations", ctypes.c_int32),
        ("dfsph_density_error_average_threshold", ctypes.c_float),
        ("dfsph_density_error_max_threshold", ctypes.c_float),
        ("dfsph_divergence_error_average_threshold", ctypes.c_float),
        ("dfsph_divergence_error_max_threshold", ctypes.c_float),
        ("num_batch_steps", ctypes.c_int32),
        ("method", ctypes.c_int32),
    ]


class Config(ctypes.Structure):
    """emsph_config (sph2d_box/config.h:7-18)."""
    _fields_ = [
        ("params", Parameters),
        ("seconds_per_sub_step", ctypes.c_float),
        ("mass_per_particle", ctypes.c_float),
        ("draw_radius", ctypes.c_float),
        ("pressure_correction_denom", ctypes.c_float),
    ]


class StepStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("divergence_iterations", ctypes.c_int32),
        ("error_average", ctypes.c_float),
        ("error_max", ctypes.c_float),
        ("divergence_error_average", ctypes.c_float),
        ("divergence_error_max", ctypes.c_float),
        ("max_velocity_sq", ctypes.c_float),
        ("cell_count", ctypes.c_uint32),
        ("sort_passes", ctypes.c_uint32),
        ("max_neighbors", ctypes.c_uint32),
        ("dt", ctypes.c_float),
        ("reserved", ctypes.c_uint32),
        ("kinetic_energy", ctypes.c_double),
        ("potential_energy", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(ctypes.Structure):
    """emsph_tuning."""
    _fields_ = [("persistent_loop", ctypes.c_int32), ("graphs", ctypes.c_int32),
                ("neighbor_batch", ctypes.c_int32), ("xcd_remap", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 4)]

    def as_dict(self) -> dict:
        return {"graphs": int(self.graphs), "neighbor_batch": int(self.neighbor_batch),
                "xcd_remap": int(self.xcd_remap)}


FORCE_GRAVITY, FORCE_POLE, FORCE_DRAG, FORCE_ANTI_COUPLING = 1, 2, 4, 8


class ForceModel(ctypes.Structure):
    """emsph_force_model (sph2d_box/forces.cpp:62-149)."""
    _fields_ = [("flags", ctypes.c_uint32), ("gravity", ctypes.c_float),
                ("pole_magnitude", ctypes.c_float), ("pole_x", ctypes.c_float),
                ("pole_y", ctypes.c_float), ("drag_magnitude", ctypes.c_float),
                ("drag_diameter", ctypes.c_float),
                ("anti_coupling_max_distance", ctypes.c_float),
                ("anti_coupling_magnitude", ctypes.c_float)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# emsph_user_forces_fn: (user, count, t, dt, mass, radius, forces_xy, pos_xy, vel_xy).
USER_FORCES_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                  ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(ctypes.c_float))


# emsph_user_colors_fn: (user, count, t, mass, radius, colors_rgba, pos_xy, vel_xy, densities).
USER_COLORS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(ctypes.c_float))


class SnapshotHeader(ctypes.Structure):
    """emsph_snapshot_header."""
    _fields_ = [("count", ctypes.c_uint64), ("time_flicks", ctypes.c_int64),
                ("method", ctypes.c_int32), ("flags", ctypes.c_uint32),
                ("bytes", ctypes.c_uint64)]


class FrameStats(ctypes.Structure):
    _fields_ = [
        ("sub_steps", ctypes.c_int32),
        ("total_iterations", ctypes.c_int32),
        ("total_divergence_iterations", ctypes.c_int32),
        ("last_dt", ctypes.c_float),
        ("last", StepStats),
    ]


_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_F = ctypes.c_float
_I = ctypes.c_int
_I64 = ctypes.c_int64
_ST = ctypes.c_int

SIGNATURES = {
    "emsph_last_error": (ctypes.c_char_p, []),
    "emsph_build_info": (ctypes.c_char_p, []),
    "emsph_device_count": (_I, []),
    "emsph_default_parameters": (_ST, [_P]),
    "emsph_config_from_parameters": (_ST, [_P, _P]),
    "emsph_estimate_solid_box_emission_count": (_SZ, [_F, _F, _F, _F, _F]),
    "emsph_emit_solid_box": (_SZ, [_F, _F, _F, _F, _F, _SZ, _P]),
    "emsph_world_walls_solid_positions": (_SZ, [_P, _I, _P, _SZ]),
    "emsph_dam_break_positions": (_SZ, [_P, _P, _SZ, _P, _SZ]),
    "emsph_lattice_positions": (_SZ, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                      ctypes.c_uint64, ctypes.c_double, _P]),
    "emsph_create": (_ST, [_P, _P, _SZ, _SZ, _P]),
    "emsph_destroy": (_ST, [_P]),
    "emsph_set_store_geometry": (_ST, [_P, _I]),
    "emsph_set_state": (_ST, [_P, _P, _P, _SZ]),
    "emsph_get_state": (_ST, [_P, _P, _P]),
    "emsph_fluid_count": (_SZ, [_P]),
    "emsph_solid_count": (_SZ, [_P]),
    "emsph_get_solid_volumes": (_ST, [_P, _P]),
    "emsph_get_field": (_ST, [_P, _I, _P]),
    "emsph_get_neighborhood": (_ST, [_P, _I, _P, _P, _P, _P]),
    "emsph_default_force_model": (_ST, [_P, _P]),
    "emsph_set_force_model": (_ST, [_P, _P]),
    "emsph_get_force_model": (_ST, [_P, _P]),
    "emsph_set_user_forces": (_ST, [_P, USER_FORCES_FN, _P]),
    "emsph_batch_set_force_model": (_ST, [_P, _I, _P]),
    "emsph_set_user_colors": (_ST, [_P, USER_COLORS_FN, _P]),
    "emsph_compute_colors": (_ST, [_P, _F]),
    "emsph_get_colors": (_ST, [_P, _P]),
    "emsph_get_colors_u8": (_ST, [_P, _P]),
    "emsph_frame_init": (_ST, [_P, _I]),
    "emsph_sub_step": (_ST, [_P, _I, _F, _F, _P]),
    "emsph_frame_step": (_ST, [_P, _I, _I64, _P]),
    "emsph_run_sub_steps": (_ST, [_P, _I, _F, _I, _P, _P]),
    "emsph_set_cfl": (_ST, [_P, _F]),
    "emsph_set_graphs": (_ST, [_P, _I]),
    "emsph_set_skin_lists": (_ST, [_P, _I]),
    "emsph_get_skin_lists": (_ST, [_P, _P, _P, _P]),
    "emsph_get_tuning": (_ST, [_P, _P]),
    "emsph_set_tuning": (_ST, [_P, _P]),
    "emsph_autotune": (_ST, [_P, _I, _F, _I, _P, _P, _P]),
    "emsph_batch_set_tuning": (_ST, [_P, _P]),
    "emsph_batch_autotune": (_ST, [_P, _I, _I, _P, _P, _P]),
    "emsph_pool_set_tuning": (_ST, [_P, _P]),
    "emsph_batch_set_graphs": (_ST, [_P, _I]),
    "emsph_batch_set_skin_lists": (_ST, [_P, _I]),
    "emsph_batch_get_skin_lists": (_ST, [_P, _P, _P, _P]),
    "emsph_batch_set_kernel_timing": (_ST, [_P, _I]),
    "emsph_batch_get_kernel_timing": (_I, [_P, _P, _P, _P, _I]),
    "emsph_batch_get_iteration_totals": (_ST, [_P, _I, _P]),
    "emsph_batch_set_cfl": (_ST, [_P, _F]),
    "emsph_set_kernel_timing": (_ST, [_P, _I]),
    "emsph_get_kernel_timing": (_I, [_P, _P, _P, _P, _I]),
    "emsph_get_iteration_totals": (_ST, [_P, _I, _P]),
    "emsph_batch_create": (_ST, [_I, _P, _P, _P, _P, _P]),
    "emsph_batch_destroy": (_ST, [_P]),
    "emsph_batch_scene_count": (_I, [_P]),
    "emsph_batch_set_state": (_ST, [_P, _I, _P, _P]),
    "emsph_batch_get_state": (_ST, [_P, _I, _P, _P]),
    "emsph_batch_get_field": (_ST, [_P, _I, _I, _P]),
    "emsph_batch_frame_init": (_ST, [_P, _I]),
    "emsph_batch_sub_step": (_ST, [_P, _I, _P]),
    "emsph_batch_run_sub_steps": (_ST, [_P, _I, _I, _P]),
    "emsph_pool_create": (_ST, [_SZ, _SZ, _I, _I, _P]),
    "emsph_pool_destroy": (_ST, [_P]),
    "emsph_pool_add": (_ST, [_P, _P, _P, _SZ, _P, _P, _SZ, _P]),
    "emsph_pool_remove": (_ST, [_P, _I]),
    "emsph_pool_commit": (_ST, [_P]),
    "emsph_pool_live": (_I, [_P, _P, _I]),
    "emsph_pool_set_cfl": (_ST, [_P, _F]),
    "emsph_pool_sub_step": (_ST, [_P, _P]),
    "emsph_pool_get_state": (_ST, [_P, _I, _P, _P]),
    "emsph_device_bytes": (_I64, []),
    "emsph_snapshot_bytes": (_SZ, [_SZ, ctypes.c_uint32]),
    "emsph_snapshot_encode": (_ST, [_P, _P, _P, _SZ, _I64, ctypes.c_int32, _P, _SZ, _P]),
    "emsph_snapshot_info": (_ST, [_P, _SZ, _P]),
    "emsph_snapshot_decode": (_ST, [_P, _SZ, _P, _P, _P]),
    "emsph_snapshot_write_file": (_ST, [ctypes.c_char_p, _P, _SZ]),
    "emsph_save_snapshot": (_ST, [_P, ctypes.c_char_p, _I64, ctypes.c_int32]),
    "emsph_load_snapshot": (_ST, [_P, ctypes.c_char_p, _P, _P]),
    "emsph_pool_save_snapshot": (_ST, [_P, _I, ctypes.c_char_p, _I64]),
    "emsph_search_create": (_ST, [_SZ, _P]),
    "emsph_search_destroy": (_ST, [_P]),
    "emsph_search_set_points": (_ST, [_P, _P, _SZ]),
    "emsph_search_run": (_ST, [_P, _F, _I, _P]),
    "emsph_search_get": (_ST, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "emsph_search_lists_noalloc": (_ST, [_P, _SZ, _F, _P, _P, _P]),
    "emsph_pool_get_iteration_totals": (_ST, [_P, _I, _P]),
}

_lib = None


class EmsphError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: {STATUS.get(status, status)}: {message}")
        self.status = status


def lib() -> ctypes.CDLL:
    """The loaded HIP library. Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m emerald_amd.build` "
                "(hipcc, gfx950). There is no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int, where: str) -> None:
    if status != OK:
        raise EmsphError(status, where, lib().emsph_last_error().decode(errors="replace"))
