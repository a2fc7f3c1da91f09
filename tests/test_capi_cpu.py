"""The C ABI library loads, exports every symbol include/emerald_sph.h declares,
and its host-side scene/config functions match the oracle (no GPU compute)."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from emerald_amd import _capi, lib
from emerald_amd import sph2d_box as sb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "emerald_sph.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(emsph_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert len(names) >= 25
    L = lib()
    for n in names:
        assert hasattr(L, n), n
    # And the Python binding covers each of them.
    assert set(names) <= set(_capi.SIGNATURES)


def test_no_device_is_reported_loudly():
    if lib().emsph_device_count() > 0:
        pytest.skip("a GPU is visible")
    cfg = sb.Simulation_config.from_parameters(sb.Parameters()).to_c()
    h = ctypes.c_void_p()
    st = lib().emsph_create(ctypes.byref(cfg), None, 0, 16, ctypes.byref(h))
    assert st == 6  # EMSPH_ERR_NO_DEVICE
    assert b"device" in lib().emsph_last_error()


def test_default_parameters_match_reference_defaults():
    p = _capi.Parameters()
    assert lib().emsph_default_parameters(ctypes.byref(p)) == 0
    # sph2d_box/parameters.h:12-43
    assert p.time_per_step == 23520000 and p.sub_steps == 30
    assert np.float32(p.support) == np.float32(0.025)
    assert p.iisph_max_pressure_iterations == 30 and p.dfsph_max_correction_iterations == 15
    assert p.method == 3  # iisph_pseudo_ap


@pytest.mark.parametrize("support", [0.025, 0.01, 0.005, 1.0 / 2560.0])
def test_config_calibration_matches_oracle(support):
    params = sb.Parameters(support=support)
    c = sb.Simulation_config.from_parameters(params)
    o = oracle.simulation_config(np.float32(support), 1000.0, params.time_per_step,
                                 params.sub_steps)
    assert np.float32(c.mass_per_particle) == o["mass_per_particle"]
    assert np.float32(c.seconds_per_sub_step) == o["seconds_per_sub_step"]
    assert np.float32(c.draw_radius) == o["draw_radius"]
    assert np.float32(c.pressure_correction_denom) == o["pressure_correction_denom"]


def test_emit_solid_box_matches_oracle():
    for box, R in [(((-0.075, -0.075), (0.0, 1.075)), 0.012375),
                   (((0.0, 0.0), (0.5, 0.8125)), 0.012375),
                   (((-0.003, -0.003), (1.0015, 0.0)), 0.0024750001)]:
        a = sb.emit_solid_box(box, R)
        b = oracle.emit_solid_box(box, R)
        assert np.array_equal(a, b)


def test_world_walls_counts():
    p = sb.Parameters()
    w = sb.world_walls_initial_solid_state(p, random_obstacles=False).positions
    w2 = sb.world_walls_initial_solid_state(p, random_obstacles=True).positions
    assert 0 < len(w) < len(w2)
    assert np.array_equal(w2[:len(w)], w)


def test_lattice_positions_match_restatement():
    # The C generator against a numpy restatement of the same SplitMix64
    # stream and FP64 expression.
    from emerald_amd import scenes
    for nx, ny, s, seed, jit in [(80, 160, 0.005, 1, 1e-3), (33, 17, 0.004, 9, 1e-3),
                                 (7, 5, 0.1, 2 ** 63 + 5, 0.25)]:
        u = scenes.uniform01(seed, 2 * nx * ny).reshape(ny * nx, 2)
        jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        base = np.stack([(ii.ravel() + 0.5) * s, (jj.ravel() + 0.5) * s], axis=1)
        want = (base + (2.0 * u - 1.0) * (jit * s)).astype(np.float32)
        got = scenes.lattice_positions(nx, ny, s, seed, jit)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (nx, ny)

