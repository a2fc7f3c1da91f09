"""Skin candidate lists (emsph_set_skin_lists): sub-steps that take their
neighbour lists from candidates recorded at a rebuild give the same results,
bit for bit, as the full path that builds them from the cell windows
(neighborhood.h:198-282) every sub-step: for every solver, in batches, across
set_state, when a particle suddenly moves too far (the sub-step is redone by
the full path), and against the oracle."""

import numpy as np
import pytest

from emerald_amd import scenes
from emerald_amd import sph2d_box as sb
from helpers import assert_bits_equal, oracle_scene, small_dam_break

pytestmark = pytest.mark.gpu

FIELDS = {
    "iisph": ["density", "pressure", "aii", "dii"],
    "iisph_ap": ["density", "pressure"],
    "iisph_pseudo_ap": ["density", "pressure"],
    "pcisph": ["density", "pressure", "pressure_force"],
    "dfsph_p": ["density", "alpha", "kappa_density", "kappa_divergence"],
}


def _run(scene, method, skin, steps, forces=None, restate_at=None):
    cfg = scenes.config_for(scene, method)
    ctx = sb.Context(cfg, sb.Solid_state(scene.solids), len(scene.positions))
    try:
        ctx.set_skin_lists(skin)
        if forces is not None:
            ctx.set_user_forces(forces)
        ctx.set_state(sb.State(scene.positions, scene.velocities))
        ctx.frame_init(method)
        its = []
        for k in range(steps):
            if restate_at is not None and k == restate_at:
                st = ctx.get_state()
                ctx.set_state(sb.State(st.positions[::-1].copy(), st.velocities[::-1].copy()))
            s = ctx.sub_step(method, k * scene.dt, scene.dt)
            its.append((s.iterations, s.divergence_iterations))
        out = {"state": ctx.get_state(), "its": its,
               "fields": {f: ctx.field(f) for f in FIELDS[method]},
               "skin": ctx.skin_lists(), "totals": ctx.iteration_totals()}
    finally:
        ctx.close()
    return out


def _same(a, b, what):
    assert a["its"] == b["its"], what
    assert_bits_equal(a["state"].positions, b["state"].positions, f"{what} positions")
    assert_bits_equal(a["state"].velocities, b["state"].velocities, f"{what} velocities")
    for f in a["fields"]:
        assert_bits_equal(a["fields"][f], b["fields"][f], f"{what} {f}")


@pytest.mark.parametrize("method", ["iisph", "pcisph", "dfsph_p", "iisph_ap", "iisph_pseudo_ap"])
def test_skin_lists_equal_full_path(gpu, method):
    scene = small_dam_break(48, 80, seed=9)
    print(f"[skin] {method} n={len(scene.positions)}", flush=True)
    full = _run(scene, method, False, 150)
    skin = _run(scene, method, True, 150)
    on, rebuilds, reuses = skin["skin"]
    print(f"[skin] {method}: on={on} rebuilds={rebuilds} reuses={reuses}", flush=True)
    assert on and rebuilds >= 2 and reuses >= 40
    assert full["skin"] == (False, 0, 0)
    _same(full, skin, method)


def test_skin_lists_against_oracle(gpu):
    scene = small_dam_break(30, 50, seed=4)
    method = "iisph"
    cfg = scenes.config_for(scene, method)
    ctx = sb.Context(cfg, sb.Solid_state(scene.solids), len(scene.positions))
    orc = oracle_scene(scene, cfg)
    try:
        ctx.set_skin_lists(True)
        ctx.set_state(sb.State(scene.positions, scene.velocities))
        ctx.frame_init(method)
        orc.frame_init(method)
        for k in range(60):
            st = ctx.sub_step(method, k * scene.dt, scene.dt)
            ost = orc.sub_step(method, k * scene.dt, scene.dt)
            assert st.iterations == ost.iterations, k
        pos, vel = orc.state_f32()
        g = ctx.get_state()
        assert_bits_equal(g.positions, pos, "positions")
        assert_bits_equal(g.velocities, vel, "velocities")
        assert_bits_equal(ctx.field("density"), orc.field("density"), "density")
        assert ctx.skin_lists()[2] >= 20
    finally:
        ctx.close()


def test_skin_lists_redo_when_a_particle_jumps(gpu):
    # A host force kicks one particle upwards at 20 m/s in sub-step 40: in the
    # next sub-step it has moved past h/2 unforeseen, so that sub-step is
    # redone by the full path; the results still equal the full path's.
    scene = small_dam_break(40, 60, seed=6)

    def kick(n, t, dt, mass, radius, f, p, v):
        if abs(t - 40 * scene.dt) < 0.5 * scene.dt:
            f[0, 1] += mass * 20.0 / dt

    full = _run(scene, "iisph", False, 80, forces=kick)
    skin = _run(scene, "iisph", True, 80, forces=kick)
    print(f"[skin] kick: skin={skin['skin']} totals={skin['totals']}", flush=True)
    assert skin["totals"][3] > full["totals"][3]  # a redone sub-step
    _same(full, skin, "kick")


def test_skin_lists_set_state_mid_run(gpu):
    scene = small_dam_break(36, 50, seed=8)
    full = _run(scene, "dfsph_p", False, 90, restate_at=45)
    skin = _run(scene, "dfsph_p", True, 90, restate_at=45)
    assert skin["skin"][1] >= 3
    _same(full, skin, "set_state")


@pytest.mark.parametrize("method,cfl", [("iisph", 0.0), ("iisph", 0.4), ("dfsph_p", 0.0),
                                        ("pcisph", 0.0)])
def test_skin_lists_batch(gpu, method, cfl):
    # Fixed dt: the candidates serve many sub-steps once the first splashes
    # of the jittered lattices calm down. CFL 0.4: each scene's dt (at most the
    # fixed one) lets its fastest particle move up to 0.4 h per sub-step, and
    # rebuilds come sooner. The results are the same either way. DFSPH_P's
    # reuses also gather its stiffnesses into the rebuild's order.
    scs = [small_dam_break(36 + 8 * k, 48 + 6 * k, seed=20 + k) for k in range(4)]
    out = []
    for skin in (False, True):
        cfgs = [scenes.config_for(sc, method) for sc in scs]
        b = sb.Batch(cfgs, [sb.Solid_state(sc.solids) for sc in scs],
                     [sb.State(sc.positions, sc.velocities) for sc in scs])
        try:
            b.set_cfl(cfl)
            b.set_skin_lists(skin)
            b.frame_init(method)
            its = [[(s.iterations, s.dt) for s in b.sub_step(method)] for _ in range(160)]
            out.append((its, [b.get_state(i) for i in range(len(scs))], b.skin_lists()))
        finally:
            b.close()
    print(f"[skin batch] {method} cfl {cfl}: {out[1][2]}", flush=True)
    assert out[0][0] == out[1][0]
    on, rebuilds, reuses = out[1][2]
    # DFSPH_P's and PCISPH's splashes end intervals sooner here, and short
    # intervals pause the candidates for a while (the full path alone).
    assert on and rebuilds >= 2 and reuses >= (20 if method == "iisph" else 10)
    for i in range(len(scs)):
        assert_bits_equal(out[0][1][i].positions, out[1][1][i].positions, f"scene {i} positions")
        assert_bits_equal(out[0][1][i].velocities, out[1][1][i].velocities,
                          f"scene {i} velocities")


def test_skin_lists_1m_against_oracle(gpu):
    sc = scenes.dam_break_1m()
    method = "iisph"
    cfg = scenes.config_for(sc, method)
    n = len(sc.positions)
    ctx = sb.Context(cfg, sb.Solid_state(sc.solids), n)
    orc = oracle_scene(sc, cfg)
    try:
        ctx.set_skin_lists(True)
        ctx.set_state(sb.State(sc.positions, sc.velocities))
        ctx.frame_init(method)
        orc.frame_init(method)
        for k in range(6):
            st = ctx.sub_step(method, k * sc.dt, sc.dt)
            ost = orc.sub_step(method, k * sc.dt, sc.dt)
            print(f"[skin 1M] sub-step {k}: {st.iterations} iterations", flush=True)
            assert st.iterations == ost.iterations, k
        pos, vel = orc.state_f32()
        g = ctx.get_state()
        assert_bits_equal(g.positions, pos, "1M positions")
        assert_bits_equal(g.velocities, vel, "1M velocities")
        for f in ("density", "pressure", "dii", "aii"):
            assert_bits_equal(ctx.field(f), orc.field(f), f"1M {f}")
        assert ctx.skin_lists()[2] >= 3
    finally:
        ctx.close()
        del orc
