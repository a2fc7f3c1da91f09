"""Batched scenes: every scene of a mixed batch is bit-identical to the same
scene run alone (single-scene context) and to the oracle."""

import numpy as np
import pytest

from emerald_amd import scenes
from emerald_amd import sph2d_box as sb
from helpers import assert_bits_equal, oracle_scene

pytestmark = pytest.mark.gpu


def _mixed_scenes():
    # Different sizes, spacings (so h, cell and dt differ), seeds and aspect ratios.
    specs = [(20, 30, 0.005, 11), (34, 17, 0.004, 12), (12, 40, 0.006, 13), (25, 25, 0.005, 14)]
    out = []
    for nx, ny, s, seed in specs:
        sc = scenes.dam_break(nx, ny, s, 0.04 * s, seed=seed)
        out.append(sc)
    return out


@pytest.mark.parametrize("method", ["iisph", "dfsph_p", "pcisph"])
def test_batch_equals_alone(gpu, method):
    scs = _mixed_scenes()
    cfgs = [scenes.config_for(sc, method) for sc in scs]
    print(f"[batch parity] {method} scenes={[len(s.positions) for s in scs]}", flush=True)
    batch = sb.Batch(cfgs, [sb.Solid_state(s.solids) for s in scs],
                     [sb.State(s.positions, s.velocities) for s in scs])
    batch.frame_init(method)
    alone = []
    orcs = []
    for sc, cfg in zip(scs, cfgs):
        ctx = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
        ctx.set_state(sb.State(sc.positions, sc.velocities))
        ctx.frame_init(method)
        alone.append(ctx)
        o = oracle_scene(sc, cfg)
        o.frame_init(method)
        orcs.append(o)
    for k in range(25):
        bst = batch.sub_step(method)
        for i, (sc, ctx, o) in enumerate(zip(scs, alone, orcs)):
            st = ctx.sub_step(method, k * sc.dt, sc.dt)
            ost = o.sub_step(method, k * sc.dt, sc.dt)
            assert bst[i].iterations == st.iterations == ost.iterations, (k, i)
            assert bst[i].divergence_iterations == st.divergence_iterations
    for i, (ctx, o) in enumerate(zip(alone, orcs)):
        b = batch.get_state(i)
        a = ctx.get_state()
        pos, vel = o.state_f32()
        assert_bits_equal(b.positions, a.positions, f"scene {i} batch vs alone positions")
        assert_bits_equal(b.velocities, a.velocities, f"scene {i} batch vs alone velocities")
        assert_bits_equal(b.positions, pos, f"scene {i} batch vs oracle positions")
        assert_bits_equal(batch.field(i, "density"), ctx.field("density"), f"scene {i} density")
