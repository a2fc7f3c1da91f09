"""FP32 on the GPU against the FP64 build of the reference path (the oracle
instantiated on double) over 1000 sub-steps: mean density error, kinetic
energy and solver iteration totals within 1 %, and the FP64 error per field
reported. Cases: the 12,800-particle dam break (dt 2e-4) with IISPH, PCISPH
and DFSPH_P, and the 1,048,576-particle dam break's block at 1/16 of its
particles (256 x 256 at spacing 1/640, dt 6.25e-5: the same CFL number) with
IISPH."""

import os

import numpy as np
import pytest

import oracle
from emerald_amd import scenes
from emerald_amd import sph2d_box as sb
from helpers import oracle_scene

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

STEPS = 1000
EVERY = 10


def _compare(sc, method):
    oracle.set_threads(min(16, os.cpu_count() or 1))
    cfg = scenes.config_for(sc, method)
    rho0 = sc.rho0
    m = float(np.float32(cfg.mass_per_particle))
    ctx = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
    ctx.set_state(sb.State(sc.positions, sc.velocities))
    ctx.frame_init(method)
    o64 = oracle_scene(sc, cfg, fp64=True)
    o64.frame_init(method)
    its32 = its64 = dits32 = dits64 = 0
    err32, err64, ke32, ke64 = [], [], [], []
    for k in range(STEPS):
        st = ctx.sub_step(method, k * sc.dt, sc.dt)
        ost = o64.sub_step(method, k * sc.dt, sc.dt)
        its32 += st.iterations
        its64 += ost.iterations
        dits32 += st.divergence_iterations
        dits64 += ost.divergence_iterations
        if k % EVERY == EVERY - 1:
            r32 = ctx.field("density").astype(np.float64)
            r64 = o64.field("density")
            err32.append(np.mean(np.abs(r32 / rho0 - 1.0)))
            err64.append(np.mean(np.abs(r64 / rho0 - 1.0)))
            ke32.append(st.kinetic_energy)
            v64 = o64.field("vel")
            ke64.append(0.5 * m * np.sum(v64 * v64))
    pos32 = ctx.get_state().positions.astype(np.float64)
    vel32 = ctx.get_state().velocities.astype(np.float64)
    pos64 = o64.field("pos")
    vel64 = o64.field("vel")
    r32 = ctx.field("density").astype(np.float64)
    r64 = o64.field("density")
    rel = lambda a, b: abs(a - b) / abs(b) if b else abs(a - b)
    d_err = rel(np.mean(err32), np.mean(err64))
    d_ke = rel(np.mean(ke32), np.mean(ke64))
    d_it = rel(its32, its64)
    d_dit = rel(dits32, dits64)
    print(f"[fp64] {method}, {len(sc.positions)} particles: mean density error FP32 "
          f"{np.mean(err32):.6e} FP64 {np.mean(err64):.6e} (rel {d_err:.2e}); mean KE FP32 "
          f"{np.mean(ke32):.6e} FP64 {np.mean(ke64):.6e} (rel {d_ke:.2e}); iterations FP32 "
          f"{its32} FP64 {its64} (rel {d_it:.2e}); divergence iterations FP32 {dits32} FP64 "
          f"{dits64}")
    print(f"[fp64] after {STEPS} sub-steps, FP32 - FP64 per field: position max |dx| "
          f"{np.max(np.abs(pos32 - pos64)):.3e} (rms {np.sqrt(np.mean((pos32 - pos64) ** 2)):.3e}); "
          f"velocity max |dv| {np.max(np.abs(vel32 - vel64)):.3e} (rms "
          f"{np.sqrt(np.mean((vel32 - vel64) ** 2)):.3e}); density max rel "
          f"{np.max(np.abs(r32 / r64 - 1.0)):.3e} (mean {np.mean(np.abs(r32 / r64 - 1.0)):.3e})",
          flush=True)
    assert d_err <= 0.01
    assert d_ke <= 0.01
    assert d_it <= 0.01
    assert d_dit <= 0.01


@pytest.mark.parametrize("method", ["iisph", "pcisph", "dfsph_p"])
def test_fp32_gpu_vs_fp64_reference_1000_steps(gpu, method):
    _compare(scenes.dam_break_12800(), method)


def test_fp32_gpu_vs_fp64_reference_1000_steps_1m_block_at_1_16(gpu):
    _compare(scenes.dam_break(256, 256, 1.0 / 640.0, 6.25e-5, seed=2), "iisph")
