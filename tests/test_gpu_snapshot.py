"""Snapshots of device state (emsph_save_snapshot / emsph_load_snapshot /
emsph_pool_save_snapshot): a run checkpointed after a frame and resumed in a
fresh context equals the uninterrupted run bit for bit; the saved colors are
the frame's; a pool scene's snapshot equals its state."""

import numpy as np
import pytest

from emerald_amd import scenes
from emerald_amd import sph2d_box as sb
from helpers import assert_bits_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["iisph", "dfsph_p"])
def test_checkpoint_resume_bit_exact(gpu, tmp_path, method):
    scene = scenes.dam_break(16, 24, 0.005, 2e-4, seed=31)
    cfg = scenes.config_for(scene, method)
    T = cfg.params.time_per_step

    def fresh():
        c = sb.Context(cfg, sb.Solid_state(scene.solids), len(scene.positions))
        c.set_state(sb.State(scene.positions, scene.velocities))
        return c

    a = fresh()
    a.frame_step(method, 0)
    path = str(tmp_path / "ckpt.emsph")
    a.save_snapshot(path, T, method)
    a.frame_step(method, T)
    a.frame_step(method, 2 * T)
    st_a = a.get_state()

    snap, t, m = sb.load_state(path)
    assert t == T and m == method and snap.colors is not None
    b = sb.Context(cfg, sb.Solid_state(scene.solids), len(scene.positions))
    t2, m2 = b.load_snapshot(path)
    assert (t2, m2) == (T, method)
    b.frame_step(method, t2)
    b.frame_step(method, t2 + T)
    st_b = b.get_state()
    assert_bits_equal(st_b.positions, st_a.positions, "resumed positions")
    assert_bits_equal(st_b.velocities, st_a.velocities, "resumed velocities")


def test_pool_scene_snapshot(gpu, tmp_path):
    scene = scenes.dam_break(12, 20, 0.005, 2e-4, seed=37)
    cfg = scenes.config_for(scene, "iisph")
    pool = sb.Pool(4096, 4096, 4, "iisph")
    sid = pool.add(cfg, sb.Solid_state(scene.solids), sb.State(scene.positions, scene.velocities))
    pool.commit()
    for _ in range(3):
        pool.sub_step()
    path = str(tmp_path / "pool.emsph")
    pool.save_snapshot(sid, path, 99)
    snap, t, m = sb.load_state(path)
    st = pool.get_state(sid)
    assert t == 99 and m == "iisph"
    assert_bits_equal(snap.positions, st.positions, "pool snapshot positions")
    assert_bits_equal(snap.velocities, st.velocities, "pool snapshot velocities")
    pool.close()
