"""Shared helpers for the parity tests (test code only)."""

import numpy as np

import oracle
from emerald_amd import scenes


def oracle_cfg(cfg) -> dict:
    """oracle.OrcConfig fields from an emerald_amd Simulation_config."""
    p = cfg.params
    return {
        "support": p.support,
        "gravity": p.gravity,
        "viscosity": p.viscosity,
        "target_density": p.target_density,
        "time_per_step": p.time_per_step,
        "sub_steps": p.sub_steps,
        "iisph_max_iterations": p.iisph.max_pressure_iterations,
        "iisph_error_average_threshold": p.iisph.error_average_threshold,
        "iisph_error_max_threshold": p.iisph.error_max_threshold,
        "iisph_omega": p.iisph.omega,
        "dfsph_max_iterations": p.dfsph.max_correction_iterations,
        "dfsph_density_avg": p.dfsph.density_error_average_threshold,
        "dfsph_density_max": p.dfsph.density_error_max_threshold,
        "dfsph_divergence_avg": p.dfsph.divergence_error_average_threshold,
        "dfsph_divergence_max": p.dfsph.divergence_error_max_threshold,
        "seconds_per_sub_step": cfg.seconds_per_sub_step,
        "mass_per_particle": cfg.mass_per_particle,
        "pressure_correction_denom": cfg.pressure_correction_denom,
    }


def oracle_scene(scene, cfg, fp64=False):
    return oracle.Scene(oracle_cfg(cfg), scene.solids, scene.positions, scene.velocities, fp64=fp64)


def small_dam_break(nx=24, ny=40, seed=5, spacing=0.005, dt=2e-4):
    return scenes.dam_break(nx, ny, spacing, dt, seed=seed)


def assert_bits_equal(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} != {b.shape}"
    if a.dtype == np.float64 or b.dtype == np.float64:
        a = a.astype(np.float32)
        b = b.astype(np.float32)
    av = a.view(np.uint32) if a.dtype == np.float32 else a
    bv = b.view(np.uint32) if b.dtype == np.float32 else b
    bad = np.flatnonzero(av.ravel() != bv.ravel())
    if len(bad):
        i = bad[0]
        raise AssertionError(
            f"{what}: {len(bad)} of {av.size} differ; first at flat {i}: "
            f"{a.ravel()[i]!r} vs {b.ravel()[i]!r}")
