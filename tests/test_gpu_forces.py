"""External forces on the GPU (sph2d_box/forces.cpp:11-149) vs the oracle.

* the device force models (gravity, constant pole attraction, simple drag,
  the combination commented out at forces.cpp:31-47) over 12 sub-steps of
  every method, bit for bit: state, the external-force field and iterations;
* a host User_forces_function (a time-dependent force written in numpy) through
  the D2H / callback / H2D path, bit for bit, with the same call arguments as
  the oracle's (t, or t + dt for DFSPH_P);
* anti-coupling repulsion on a scene with nearly coincident particles: the
  force within one float ulp of cos/sin per kick (the host library's float
  cos/sin against the device's double-rounded ones), and the state after one
  sub-step to 1e-6;
* a batch of scenes with different force models equals each scene alone.
"""

import numpy as np
import pytest

import oracle
from emerald_amd import scenes
from emerald_amd import sph2d_box as sb
from helpers import assert_bits_equal, oracle_scene

pytestmark = pytest.mark.gpu

ALL = ["iisph", "iisph_ap", "iisph_pseudo_ap", "pcisph", "dfsph_p"]


def _pair(method, nx=20, ny=32, seed=13):
    scene = scenes.dam_break(nx, ny, 0.005, 2e-4, seed=seed)
    cfg = scenes.config_for(scene, method)
    ctx = sb.Context(cfg, sb.Solid_state(scene.solids), len(scene.positions))
    ctx.set_state(sb.State(scene.positions, scene.velocities))
    orc = oracle_scene(scene, cfg)
    return scene, cfg, ctx, orc


def _device_models(cfg):
    m = np.float32(cfg.mass_per_particle)
    g = np.float32(cfg.params.gravity)
    h = np.float32(cfg.params.support)
    # forces.cpp:31-47: pole attraction of a quarter of gravity at the box
    # centre, drag 0.025 of the support, and gravity.
    return sb.ForceModel(gravity_on=True, gravity=float(g),
                         pole_magnitude=float(np.float32(0.25) * g * m), pole=(0.5, 0.5),
                         drag_magnitude=0.025, drag_diameter=float(h))


def _orc_set(orc, fm: sb.ForceModel, **kw):
    c = fm.to_c()
    orc.set_forces(flags=int(c.flags), gravity=c.gravity, pole_magnitude=c.pole_magnitude,
                   pole=(c.pole_x, c.pole_y), drag_magnitude=c.drag_magnitude,
                   drag_diameter=c.drag_diameter,
                   anti_coupling_max_distance=c.anti_coupling_max_distance,
                   anti_coupling_magnitude=c.anti_coupling_magnitude, **kw)


@pytest.mark.parametrize("method", ALL)
def test_device_force_models_bit_exact(gpu, method):
    scene, cfg, ctx, orc = _pair(method)
    fm = _device_models(cfg)
    ctx.set_force_model(fm)
    _orc_set(orc, fm, draw_radius=cfg.draw_radius)
    got = ctx.force_model()
    assert got.pole_magnitude == pytest.approx(fm.pole_magnitude) and got.drag_magnitude
    ctx.frame_init(method)
    orc.frame_init(method)
    print(f"[forces] {method} n={len(scene.positions)} device models, 12 sub-steps", flush=True)
    for k in range(12):
        st = ctx.sub_step(method, k * scene.dt, scene.dt)
        ost = orc.sub_step(method, k * scene.dt, scene.dt)
        assert (st.iterations, st.divergence_iterations) == \
            (ost.iterations, ost.divergence_iterations), f"step {k}"
    assert_bits_equal(ctx.field("external_force"), orc.field("external_force"),
                      f"{method} external force")
    pos, vel = orc.state_f32()
    g = ctx.get_state()
    assert_bits_equal(g.positions, pos, f"{method} positions")
    assert_bits_equal(g.velocities, vel, f"{method} velocities")


@pytest.mark.parametrize("method", ["iisph", "iisph_pseudo_ap", "dfsph_p", "pcisph"])
def test_host_callback_bit_exact(gpu, method):
    scene, cfg, ctx, orc = _pair(method, nx=16, ny=24, seed=17)
    g = np.float32(cfg.params.gravity)
    logs = {"gpu": [], "orc": []}

    def make(which):
        def fn(n, t, dt, mass, radius, forces, pos, vel):
            logs[which].append((n, np.float32(t), np.float32(dt), np.float32(mass),
                                np.float32(radius)))
            m = np.float32(mass)
            forces[:, 0] += m * np.float32(3.0) * np.float32(np.sin(np.float32(40.0) * t))
            forces[:, 1] -= m * g
            forces[:, 1] += np.float32(0.5) * m * (pos[:, 0] - np.float32(0.2))
        return fn

    ctx.set_force_model(sb.ForceModel(gravity_on=False, gravity=float(g)))
    ctx.set_user_forces(make("gpu"))
    orc.set_forces(flags=0, gravity=float(g), draw_radius=cfg.draw_radius,
                   user_forces=make("orc"))
    ctx.frame_init(method)
    orc.frame_init(method)
    print(f"[forces] {method} host callback, 6 sub-steps", flush=True)
    t0 = 0.125
    for k in range(6):
        st = ctx.sub_step(method, t0 + k * scene.dt, scene.dt)
        ost = orc.sub_step(method, t0 + k * scene.dt, scene.dt)
        assert st.iterations == ost.iterations, f"step {k}"
    assert logs["gpu"] == logs["orc"]
    assert len(logs["gpu"]) == 6
    pos, vel = orc.state_f32()
    gs = ctx.get_state()
    assert_bits_equal(gs.positions, pos, f"{method} callback positions")
    assert_bits_equal(gs.velocities, vel, f"{method} callback velocities")
    # Removing the callback restores the device path.
    ctx.set_user_forces(None)
    ctx.set_force_model(sb.default_gravity_forces(float(g)))
    ctx.sub_step(method, 0.0, scene.dt)


@pytest.mark.parametrize("method", ["iisph", "dfsph_p"])
def test_anti_coupling_within_an_ulp(gpu, method):
    scene = scenes.dam_break(16, 24, 0.005, 2e-4, seed=19)
    pos = scene.positions.copy()
    # Nearly coincident copies of every 7th particle (appended), so some
    # particles have one or more neighbours closer than tiny_h.
    dup = pos[::7] + np.float32(1e-7)
    pos = np.ascontiguousarray(np.concatenate([pos, dup]).astype(np.float32))
    vel = np.zeros_like(pos)
    sc2 = scenes.DamBreak(pos, vel, scene.solids, scene.support, scene.spacing, scene.dt,
                          scene.mass)
    cfg = scenes.config_for(sc2, method)
    ctx = sb.Context(cfg, sb.Solid_state(sc2.solids), len(pos))
    ctx.set_state(sb.State(pos, vel))
    orc = oracle_scene(sc2, cfg)
    h = np.float32(cfg.params.support)
    tiny = np.float32(0.001) * h
    m = np.float32(cfg.mass_per_particle)
    # forces.cpp:41-47: an offset of tiny_h per step
    mag = m * tiny / np.float32(np.float32(scene.dt) * np.float32(scene.dt))
    fm = sb.ForceModel(gravity_on=True, gravity=float(cfg.params.gravity),
                       anti_coupling_max_distance=float(tiny), anti_coupling_magnitude=float(mag))
    ctx.set_force_model(fm)
    _orc_set(orc, fm)
    ctx.frame_init(method)
    orc.frame_init(method)
    ctx.sub_step(method, 0.0, scene.dt)
    orc.sub_step(method, 0.0, scene.dt)
    fg = ctx.field("external_force").astype(np.float64)
    fo = orc.field("external_force")
    kicked = np.abs(fo[:, 0]) > 0
    assert kicked.sum() >= 2 * len(dup) - 2
    # Each kick is mag * (cos a, sin a): one float ulp of cos/sin per kick.
    counts = np.rint(np.hypot(fo[:, 0], fo[:, 1] + m * cfg.params.gravity) / mag)
    tol = (counts + 1) * float(mag) * 2.0 ** -23 + 1e-12
    assert np.all(np.abs(fg[:, 0] - fo[:, 0]) <= tol)
    assert np.all(np.abs(fg[:, 1] - fo[:, 1]) <= tol)
    po, vo = orc.state_f32()
    gs = ctx.get_state()
    np.testing.assert_allclose(gs.velocities, vo, rtol=1e-6, atol=1e-6 * float(np.abs(vo).max()))
    np.testing.assert_allclose(gs.positions, po, rtol=1e-6, atol=1e-7)


def test_batch_mixed_force_models_equal_alone(gpu):
    method = "iisph"
    sc = [scenes.dam_break(14 + 3 * k, 20 + 2 * k, 0.005, 2e-4, seed=30 + k) for k in range(3)]
    cfgs = [scenes.config_for(s, method) for s in sc]
    models = [None, _device_models(cfgs[1]),
              sb.ForceModel(gravity_on=False, gravity=float(cfgs[2].params.gravity),
                            drag_magnitude=0.05, drag_diameter=0.005)]
    b = sb.Batch(cfgs, [sb.Solid_state(s.solids) for s in sc],
                 [sb.State(s.positions, s.velocities) for s in sc])
    for i, fm in enumerate(models):
        if fm is not None:
            b.set_force_model(i, fm)
    b.frame_init(method)
    for _ in range(8):
        b.sub_step(method)
    for i, s in enumerate(sc):
        ctx = sb.Context(cfgs[i], sb.Solid_state(s.solids), len(s.positions))
        ctx.set_state(sb.State(s.positions, s.velocities))
        if models[i] is not None:
            ctx.set_force_model(models[i])
        ctx.frame_init(method)
        for _ in range(8):
            ctx.sub_step(method, 0.0, cfgs[i].seconds_per_sub_step)
        a = ctx.get_state()
        bb = b.get_state(i)
        assert_bits_equal(bb.positions, a.positions, f"scene {i} positions")
        assert_bits_equal(bb.velocities, a.velocities, f"scene {i} velocities")
        ctx.close()
    b.close()


@pytest.mark.parametrize("method", ["iisph", "dfsph_p"])
def test_host_callback_in_frame_steps(gpu, method):
    """The callback's time inside a frame step comes from the device-side
    frame loop (the sub-step start, t + dt for DFSPH_P): the same calls as the
    oracle's frame step."""
    scene, cfg, ctx, orc = _pair(method, nx=12, ny=18, seed=43)
    g = np.float32(cfg.params.gravity)
    logs = {"gpu": [], "orc": []}

    def make(which):
        def fn(n, t, dt, mass, radius, forces, pos, vel):
            logs[which].append((n, np.float32(t), np.float32(dt)))
            forces[:, 1] -= np.float32(mass) * g
            forces[:, 0] += np.float32(mass) * np.float32(np.cos(np.float32(t)))
        return fn

    ctx.set_force_model(sb.ForceModel(gravity_on=False, gravity=float(g)))
    ctx.set_user_forces(make("gpu"))
    orc.set_forces(flags=0, gravity=float(g), draw_radius=cfg.draw_radius,
                   user_forces=make("orc"))
    T = cfg.params.time_per_step
    for f in range(2):
        fs = ctx.frame_step(method, 7 * T + f * T)
        n, _ = orc.frame_step(method, 7 * T + f * T)
        assert fs.sub_steps == n
    assert logs["gpu"] == logs["orc"]
    pos, vel = orc.state_f32()
    gs = ctx.get_state()
    assert_bits_equal(gs.positions, pos, f"{method} frame callback positions")
    assert_bits_equal(gs.velocities, vel, f"{method} frame callback velocities")
