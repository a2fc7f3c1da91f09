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


@pytest.mark.parametrize("method", ["iisph", "iisph_ap", "iisph_pseudo_ap", "dfsph_p", "pcisph"])
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


def test_batch_cfl_dt_and_energy(gpu):
    # Per-scene CFL dt on the device and the FP64 energy sums: identical alone
    # and in the batch, dt per cfl.cpp:12-21 over the previous max |v|^2 (the
    # first from the uploaded velocities), and the energies equal to a host
    # FP64 sum of the downloaded state. The scenes start in a uniform flow fast
    # enough for the CFL bound to be below the largest dt.
    method = "iisph"
    factor = 0.04
    scs = _mixed_scenes()
    cfgs = [scenes.config_for(sc, method) for sc in scs]
    states = []
    for k, sc in enumerate(scs):
        u = scenes.uniform01(77 + k, 2 * len(sc.positions)).reshape(-1, 2)
        v = np.array([1.5, -0.5]) + 0.2 * (u - 0.5)
        states.append(sb.State(sc.positions, v.astype(np.float32)))
    batch = sb.Batch(cfgs, [sb.Solid_state(s.solids) for s in scs], states)
    batch.set_cfl(factor)
    batch.frame_init(method)
    alone = []
    for sc, cfg, st0 in zip(scs, cfgs, states):
        ctx = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
        ctx.set_cfl(factor)
        ctx.set_state(st0)
        ctx.frame_init(method)
        alone.append(ctx)
    prev_vmax2 = [float(np.max(np.sum(s.velocities.astype(np.float32) ** 2, axis=1)))
                  for s in states]
    cfl_bound = 0
    for k in range(20):
        bst = batch.sub_step(method)
        for i, (sc, cfg, ctx) in enumerate(zip(scs, cfgs, alone)):
            st = ctx.sub_step(method, 0.0, cfg.seconds_per_sub_step)
            assert bst[i].dt == st.dt and bst[i].iterations == st.iterations, (k, i)
            assert bst[i].kinetic_energy == st.kinetic_energy, (k, i)
            assert bst[i].potential_energy == st.potential_energy, (k, i)
            dt_max = np.float32(cfg.seconds_per_sub_step)
            want = dt_max
            if prev_vmax2[i] > 0:
                # The factor is held as a float on the device.
                d = (float(np.float32(factor)) * float(np.float32(sc.support))
                     / np.sqrt(np.float64(prev_vmax2[i])))
                if d < float(dt_max):
                    want = np.float32(d)
                    cfl_bound += 1
            assert np.float32(st.dt) == want, (k, i, st.dt, want)
            prev_vmax2[i] = float(np.float32(st.max_velocity_sq))
            s = ctx.get_state()
            m = np.float64(np.float32(cfg.mass_per_particle))
            v = s.velocities.astype(np.float64)
            ke = 0.5 * m * np.sum(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1])
            pe = m * np.float64(np.float32(sc.gravity)) * np.sum(s.positions[:, 1].astype(np.float64))
            assert abs(st.kinetic_energy - ke) <= 1e-12 * abs(ke), (k, i)
            assert abs(st.potential_energy - pe) <= 1e-12 * abs(pe), (k, i)
    assert cfl_bound > 0, "the CFL bound never went below the largest dt"
    for i, ctx in enumerate(alone):
        assert_bits_equal(batch.get_state(i).positions, ctx.get_state().positions, f"scene {i}")


def test_cfl_sub_step_redone_after_short_sort(gpu):
    # A sub-step sorts with the previous sub-step's digit count and is redone
    # when the key range has grown past it. The redone sub-step must take the
    # same CFL dt. A 12 x 12 block near the origin has keys below 256 (one
    # digit); one particle just left of grid column 16 moving in +x crosses it
    # in the first sub-step, so the second needs two digits and is redone.
    # Reference for every sub-step: the same state stepped with a fresh
    # upload before each sub-step (exact digit count, max |v|^2 recomputed).
    method = "iisph"
    factor = 0.4
    sc = scenes.dam_break(12, 12, 0.005, 2e-3, seed=21)
    cfg = scenes.config_for(sc, method)
    cell = 2.0 * float(np.float32(sc.support))
    pos = sc.positions.astype(np.float32).copy()
    assert np.max(np.floor(pos / cell)) < 16
    u = scenes.uniform01(91, 2 * len(pos)).reshape(-1, 2)
    vel = (np.array([1.5, -0.5]) + 0.2 * (u - 0.5)).astype(np.float32)
    pos[0] = (np.float32(16 * cell - 0.0005), np.float32(0.05))
    vel[0] = (np.float32(1.5), np.float32(0.0))
    st0 = sb.State(pos, vel)
    a = sb.Context(cfg, sb.Solid_state(sc.solids), len(pos))
    b = sb.Context(cfg, sb.Solid_state(sc.solids), len(pos))
    for ctx in (a, b):
        ctx.set_skin_lists(False)  # the full path's sort, every sub-step
        ctx.set_cfl(factor)
        ctx.set_state(st0)
        ctx.frame_init(method)
        ctx.iteration_totals(reset=True)
    for k in range(12):
        sa = a.sub_step(method, 0.0, cfg.seconds_per_sub_step)
        sb_ = b.sub_step(method, 0.0, cfg.seconds_per_sub_step)
        assert np.float32(sa.dt) == np.float32(sb_.dt), (k, sa.dt, sb_.dt)
        ga, gb = a.get_state(), b.get_state()
        assert_bits_equal(ga.positions, gb.positions, f"sub-step {k} positions")
        assert_bits_equal(ga.velocities, gb.velocities, f"sub-step {k} velocities")
        b.set_state(gb)
    redone = a.iteration_totals()[3]
    assert redone > 0, "no sub-step was redone: the scene did not grow its key range"
    assert sa.dt < cfg.seconds_per_sub_step


def test_changed_dt_keeps_scene_totals_and_cfl(gpu):
    # A sub-step with another dt than the last re-uploads the scene's
    # constants; its tiles (max |v|^2, energies) and CFL factor stay.
    method = "iisph"
    sc = scenes.dam_break(16, 20, 0.005, 2e-4, seed=23)
    cfg = scenes.config_for(sc, method)
    ctx = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
    ctx.set_state(sb.State(sc.positions, sc.velocities))
    ctx.frame_init(method)
    m = np.float64(np.float32(cfg.mass_per_particle))
    for k, dt in enumerate([2e-4, 1e-4, 1.5e-4]):
        st = ctx.sub_step(method, 0.0, dt)
        s = ctx.get_state()
        v = s.velocities.astype(np.float64)
        ke = 0.5 * m * np.sum(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1])
        vmax2 = np.max(np.sum(s.velocities.astype(np.float32) ** 2, axis=1, dtype=np.float32))
        assert abs(st.kinetic_energy - ke) <= 1e-12 * abs(ke), (k, st.kinetic_energy, ke)
        assert np.float32(st.max_velocity_sq) == np.float32(vmax2), (k, st.max_velocity_sq, vmax2)
    # CFL with a dt change: the bound applies to the new dt.
    ctx.set_cfl(1e-4)
    st = ctx.sub_step(method, 0.0, 3e-4)
    assert np.float32(st.dt) < np.float32(3e-4)


@pytest.mark.skip(reason="persistent cooperative solve disabled: a 1,792-workgroup cooperative "
                         "launch of it did not finish on the MI355X and the cause was not "
                         "established; set_tuning(persistent_loop=True) now raises")
def test_persistent_solve_equals_launches(gpu):
    # The IISPH pressure solve in one cooperative launch against one launch per
    # sweep, on a mixed batch whose scenes converge at different iterations.
    method = "iisph"
    scs = _mixed_scenes()
    cfgs = [scenes.config_for(sc, method) for sc in scs]
    out = []
    for persistent in (True, False):
        batch = sb.Batch(cfgs, [sb.Solid_state(s.solids) for s in scs],
                         [sb.State(s.positions, s.velocities) for s in scs])
        batch.set_tuning(persistent_loop=persistent)
        batch.set_cfl(0.4)
        batch.frame_init(method)
        its = [[st.iterations for st in batch.sub_step(method)] for _ in range(15)]
        out.append((its, [batch.get_state(i) for i in range(len(scs))]))
        batch.close()
    assert out[0][0] == out[1][0]
    assert len({tuple(x) for x in out[0][0]}) > 1 or len(set(out[0][0][0])) > 1
    for i, (a, b) in enumerate(zip(out[0][1], out[1][1])):
        assert_bits_equal(a.positions, b.positions, f"scene {i} positions")
        assert_bits_equal(a.velocities, b.velocities, f"scene {i} velocities")


def test_autotune_leaves_state(gpu):
    method = "iisph"
    sc = scenes.dam_break(30, 40, 0.005, 2e-4, seed=31)
    cfg = scenes.config_for(sc, method)
    runs = []
    for tune in (True, False):
        ctx = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
        ctx.set_state(sb.State(sc.positions, sc.velocities))
        ctx.frame_init(method)
        for k in range(3):
            ctx.sub_step(method, 0.0, sc.dt)
        before = ctx.get_state()
        if tune:
            chosen, times = ctx.autotune(method, sc.dt, 3)
            assert len(times) == 7 and chosen == ctx.tuning()
            assert_bits_equal(ctx.get_state().positions, before.positions, "state after autotune")
        its = [ctx.sub_step(method, 0.0, sc.dt).iterations for _ in range(6)]
        runs.append((its, ctx.get_state()))
        ctx.close()
    assert runs[0][0] == runs[1][0]
    assert_bits_equal(runs[0][1].positions, runs[1][1].positions, "positions")


@pytest.mark.parametrize("method", ["iisph", "dfsph_p"])
def test_tuning_shapes_equal(gpu, method):
    # Every tuning shape (neighbour batch depth, XCD remap) gives the same bits.
    scs = _mixed_scenes()
    cfgs = [scenes.config_for(sc, method) for sc in scs]
    ref = None
    for nb in (2, 4, 8):
        for xcd in (True, False):
            batch = sb.Batch(cfgs, [sb.Solid_state(s.solids) for s in scs],
                             [sb.State(s.positions, s.velocities) for s in scs])
            batch.set_tuning(neighbor_batch=nb, xcd_remap=xcd)
            batch.frame_init(method)
            its = [[st.iterations for st in batch.sub_step(method)] for _ in range(8)]
            got = (its, [batch.get_state(i).positions for i in range(len(scs))])
            batch.close()
            if ref is None:
                ref = got
                continue
            assert got[0] == ref[0], (nb, xcd)
            for i, (a, b) in enumerate(zip(got[1], ref[1])):
                assert_bits_equal(a, b, f"nb {nb} xcd {xcd} scene {i}")



def test_batch_dense_index_used(gpu):
    """After its first (exact) sub-step a batch indexes its cells with the
    dense index, each scene in its own table (no radix passes), and stays
    bit-identical to the scenes run alone."""
    method = "iisph"
    scs = _mixed_scenes()
    cfgs = [scenes.config_for(sc, method) for sc in scs]
    batch = sb.Batch(cfgs, [sb.Solid_state(s.solids) for s in scs],
                     [sb.State(s.positions, s.velocities) for s in scs])
    batch.set_skin_lists(False)  # the index runs every sub-step
    batch.frame_init(method)
    batch.run_sub_steps(method, 3)
    batch.set_kernel_timing(True)
    batch.run_sub_steps(method, 2)
    stages = batch.kernel_timing()
    batch.set_kernel_timing(False)
    print(f"[batch dense] stages {sorted(stages)}", flush=True)
    assert "dense_sort" in stages and "radix_sort" not in stages
    for i, (sc, cfg) in enumerate(zip(scs, cfgs)):
        ctx = sb.Context(cfg, sb.Solid_state(sc.solids), len(sc.positions))
        ctx.set_state(sb.State(sc.positions, sc.velocities))
        ctx.frame_init(method)
        for k in range(5):
            ctx.sub_step(method, k * sc.dt, sc.dt)
        assert_bits_equal(batch.get_state(i).positions, ctx.get_state().positions, f"scene {i}")
        ctx.close()
    batch.close()


def test_batch_dense_table_growth_redo(gpu):
    """A scene whose block, right of its walls (cells up to 101), moves right
    across cell coordinate 128 leaves the 128 x 128-cell table its keys and
    walls asked for: the batch's sub-step is flagged and redone with the radix
    sort, the scene's table grows, and every scene equals the oracle bit for
    bit."""
    method = "iisph"
    base = scenes.dam_break(12, 16, 0.005, 2e-4, seed=53)
    pos = base.positions + np.array([1.2, 0.0], np.float32)  # cells 120..126 in x
    vel = np.tile(np.array([[8.0, 0.0]], np.float32), (len(pos), 1))
    moving = scenes.DamBreak(np.ascontiguousarray(pos), vel, base.solids, base.support,
                             base.spacing, base.dt, base.mass)
    still = scenes.dam_break(20, 30, 0.005, 2e-4, seed=54)
    scs = [moving, still]
    cfgs = [scenes.config_for(sc, method) for sc in scs]
    batch = sb.Batch(cfgs, [sb.Solid_state(s.solids) for s in scs],
                     [sb.State(s.positions, s.velocities) for s in scs])
    orcs = [oracle_scene(sc, cfg) for sc, cfg in zip(scs, cfgs)]
    batch.set_skin_lists(False)  # the index runs every sub-step
    batch.frame_init(method)
    for o in orcs:
        o.frame_init(method)
    batch.iteration_totals(reset=True)
    for k in range(30):
        bst = batch.sub_step(method)
        for i, (sc, o) in enumerate(zip(scs, orcs)):
            ost = o.sub_step(method, k * sc.dt, sc.dt)
            assert bst[i].iterations == ost.iterations, (k, i)
    for i, o in enumerate(orcs):
        p, v = o.state_f32()
        st = batch.get_state(i)
        assert_bits_equal(st.positions, p, f"scene {i} positions")
        assert_bits_equal(st.velocities, v, f"scene {i} velocities")
    assert float(orcs[0].state_f32()[0][:, 0].max()) > 1.28  # crossed cell coordinate 128
    assert batch.iteration_totals()[3] >= 1  # a sub-step was redone
    batch.close()
