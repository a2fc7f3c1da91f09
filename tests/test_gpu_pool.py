"""Pools (the served form): scenes admitted and retired between sub-steps give
bit-identical results to each scene stepped alone, whatever else is live."""

import numpy as np
import pytest

from emerald_amd import scenes
from emerald_amd import sph2d_box as sb
from emerald_amd._capi import EmsphError
from helpers import assert_bits_equal

pytestmark = pytest.mark.gpu


def _scenes():
    specs = [(20, 30, 0.005, 21), (34, 17, 0.004, 22), (12, 40, 0.006, 23), (25, 25, 0.005, 24)]
    return [scenes.dam_break(nx, ny, s, 0.04 * s, seed=seed) for nx, ny, s, seed in specs]


def _alone(sc, cfg, method, steps):
    ctx = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
    ctx.set_state(sb.State(sc.positions, sc.velocities))
    ctx.frame_init(method)
    its = []
    for _ in range(steps):
        its.append(ctx.sub_step(method, 0.0, cfg.seconds_per_sub_step).iterations)
    st = ctx.get_state()
    ctx.close()
    return st, its


@pytest.mark.parametrize("method", ["iisph", "dfsph_p"])
def test_pool_admit_retire_equals_alone(gpu, method):
    scs = _scenes()
    cfgs = [scenes.config_for(sc, method) for sc in scs]
    print(f"[pool parity] {method} sizes={[len(s.positions) for s in scs]}", flush=True)
    pool = sb.Pool(20_000, 40_000, 8, method)
    try:
        add = lambda i: pool.add(cfgs[i], sb.Solid_state(scs[i].solids),
                                 sb.State(scs[i].positions, scs[i].velocities))
        ids = {0: add(0), 1: add(1)}
        pool.commit()
        steps = {0: 0, 1: 0}
        its = {0: [], 1: [], 2: [], 3: []}

        def step(n):
            for _ in range(n):
                live = pool.live()
                st = pool.sub_step()
                for sid, s in zip(live, st):
                    i = [k for k, v in ids.items() if v == sid][0]
                    steps[i] += 1
                    its[i].append(s.iterations)

        step(5)
        retired = pool.get_state(ids[1])
        ids[2] = add(2)
        pool.remove(ids[1])
        steps[2] = 0
        pool.commit()
        assert pool.live() == [ids[0], ids[2]]
        step(7)
        ids[3] = add(3)
        steps[3] = 0
        pool.commit()
        step(4)
        assert steps == {0: 16, 1: 5, 2: 11, 3: 4}
        for i in (0, 2, 3):
            want, want_its = _alone(scs[i], cfgs[i], method, steps[i])
            got = pool.get_state(ids[i])
            assert its[i] == want_its, (i, its[i], want_its)
            assert_bits_equal(got.positions, want.positions, f"scene {i} positions")
            assert_bits_equal(got.velocities, want.velocities, f"scene {i} velocities")
        want, _ = _alone(scs[1], cfgs[1], method, 5)
        assert_bits_equal(retired.positions, want.positions, "retired scene positions")
    finally:
        pool.close()


def test_pool_capacity(gpu):
    scs = _scenes()
    cfg = scenes.config_for(scs[0], "iisph")
    n = len(scs[0].positions)
    pool = sb.Pool(n + 10, 10 * len(scs[0].solids), 4, "iisph")
    try:
        pool.add(cfg, sb.Solid_state(scs[0].solids), sb.State(scs[0].positions, scs[0].velocities))
        with pytest.raises(EmsphError, match="CAPACITY"):
            pool.add(cfg, sb.Solid_state(scs[0].solids),
                     sb.State(scs[0].positions, scs[0].velocities))
        pool.commit()
        assert len(pool.live()) == 1
        pool.sub_step()
        # An empty pool steps as a no-op.
        pool.remove(pool.live()[0])
        pool.commit()
        assert pool.live() == []
        assert pool.sub_step() == []
    finally:
        pool.close()


@pytest.mark.parametrize("method", ["iisph", "dfsph_p"])
def test_removed_scene_takes_no_iterations_until_commit(gpu, method):
    # A scene removed but not yet committed rides along in the sort and the
    # neighbour build but takes no solver iterations; the other scenes'
    # results stay those of each scene alone.
    scs = _scenes()
    cfgs = [scenes.config_for(sc, method) for sc in scs]
    pool = sb.Pool(40_000, 60_000, 8, method)
    try:
        ids = [pool.add(cfgs[i], sb.Solid_state(scs[i].solids),
                        sb.State(scs[i].positions, scs[i].velocities)) for i in range(3)]
        pool.commit()
        for _ in range(4):
            pool.sub_step()
        pool.remove(ids[1])
        assert pool.live() == ids  # staged: still live until the commit
        for _ in range(3):
            st = pool.sub_step()
            assert st[1].iterations == 0 and st[1].divergence_iterations == 0
            assert st[0].iterations > 0 and st[2].iterations > 0
        pool.commit()
        assert pool.live() == [ids[0], ids[2]]
        pool.sub_step()
        for i in (0, 2):
            want, _ = _alone(scs[i], cfgs[i], method, 8)
            got = pool.get_state(ids[i])
            assert_bits_equal(got.positions, want.positions, f"scene {i} positions")
            assert_bits_equal(got.velocities, want.velocities, f"scene {i} velocities")
    finally:
        pool.close()
