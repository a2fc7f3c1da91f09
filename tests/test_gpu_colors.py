"""State::colors on the GPU (sph2d_box/colors.cpp:9-70) vs the oracle.

* frame steps of IISPH, PCISPH and DFSPH_P compute target_density_colors at
  the frame's end (adaptive_time_step.h:105-110, pcisph.cpp:186-192) from the
  last sub-step's densities: RGBA floats and convert_colors bytes bit for bit;
* a host User_colors_function receives the state and densities of the frame's
  end in original order (equal to get_state and the density field) and the
  frame's end time, and its colours are what get_colors returns.
"""

import numpy as np
import pytest

from emerald_amd import scenes
from emerald_amd import sph2d_box as sb
from helpers import assert_bits_equal, oracle_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["iisph", "pcisph", "dfsph_p"])
def test_frame_colors_bit_exact(gpu, method):
    scene = scenes.dam_break(16, 24, 0.005, 2e-4, seed=23)
    cfg = scenes.config_for(scene, method)
    ctx = sb.Context(cfg, sb.Solid_state(scene.solids), len(scene.positions))
    ctx.set_state(sb.State(scene.positions, scene.velocities))
    orc = oracle_scene(scene, cfg)
    print(f"[colors] {method} one frame", flush=True)
    ctx.frame_step(method, 0)
    orc.frame_step(method, 0)
    rgba, rgba8 = orc.colors()
    assert_bits_equal(ctx.colors(), rgba, f"{method} colors")
    assert np.array_equal(ctx.colors_u8(), rgba8)


def test_user_colors_callback(gpu):
    method = "iisph"
    scene = scenes.dam_break(12, 20, 0.005, 2e-4, seed=29)
    cfg = scenes.config_for(scene, method)
    ctx = sb.Context(cfg, sb.Solid_state(scene.solids), len(scene.positions))
    ctx.set_state(sb.State(scene.positions, scene.velocities))
    seen = {}

    def speed_colors(n, t, mass, radius, colors, pos, vel, dens):
        seen.update(n=n, t=t, mass=mass, pos=pos.copy(), vel=vel.copy(), dens=dens.copy())
        s = np.sqrt(np.sum(vel.astype(np.float64) ** 2, axis=1)).astype(np.float32)
        colors[:, 0] = s
        colors[:, 1] = dens / np.float32(1000.0)
        colors[:, 2] = pos[:, 0]
        colors[:, 3] = np.float32(1.0)

    ctx.set_user_colors(speed_colors)
    t0 = 5 * cfg.params.time_per_step
    ctx.frame_step(method, t0)
    st = ctx.get_state()
    assert seen["n"] == len(scene.positions)
    assert np.float32(seen["t"]) == np.float32((t0 + cfg.params.time_per_step) / 705600000.0)
    assert_bits_equal(seen["pos"], st.positions, "callback positions")
    assert_bits_equal(seen["vel"], st.velocities, "callback velocities")
    assert_bits_equal(seen["dens"], ctx.field("density"), "callback densities")
    got = ctx.colors()
    assert_bits_equal(got[:, 2], st.positions[:, 0], "callback colours")
    ctx.set_user_colors(None)
    ctx.compute_colors(0.0)
    c = ctx.colors()
    assert np.all(c[:, 2] == 1.0) and np.all(c[:, 0] == c[:, 1])
