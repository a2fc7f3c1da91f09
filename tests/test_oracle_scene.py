"""The reference's scene generators in the oracle and in the library.

world_walls_initial_solid_state (sph2d_box/solids.cpp:139-213, with its 16
random obstacle boxes from std::mt19937_64{1919}) and dam_break_initial_state
(sph2d_box/initial_state.cpp:14-97: a hex block minus the fluid within 2R of a
solid) are restated in the oracle and in the library's host code. The oracle is
pinned here independently of both:

* the engine: a pure-Python mt19937_64 checked against the C++ standard's
  required value ([rand.predef]: the 10000th output of a default-constructed
  mt19937_64 is 9981545732273789042);
* the float draws: libstdc++'s generate_canonical<float, 24> over a 64-bit
  engine (one engine call x, float(x) * 2^-64, clamped below 1) and
  uniform_real_distribution's u * (b - a) + a, in float32 arithmetic;
* the obstacle boxes: solids.cpp:160-176 evaluated in float32 with the draws in
  left-to-right order (the reference's clang build);
* the kill step: an all-pairs float32 distance test, which equals the
  reference's 3x3-cell search because 2R < the cell size 2H.
"""

import numpy as np
import pytest

import oracle
from emerald_amd import sph2d_box as sb

F32 = np.float32
MASK64 = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister of the C++ standard ([rand.eng.mers],
    [rand.predef] mt19937_64 parameters)."""
    N, M = 312, 156
    MATRIX_A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed=5489):
        self.mt = [0] * self.N
        self.mt[0] = seed & MASK64
        for i in range(1, self.N):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & MASK64
        self.i = self.N

    def _twist(self):
        mt = self.mt
        for k in range(self.N):
            y = (mt[k] & self.UPPER) | (mt[(k + 1) % self.N] & self.LOWER)
            v = mt[(k + self.M) % self.N] ^ (y >> 1)
            if y & 1:
                v ^= self.MATRIX_A
            mt[k] = v
        self.i = 0

    def __call__(self):
        if self.i >= self.N:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & MASK64


def uniform_float(gen, a, b):
    """uniform_real_distribution<float>{a, b}(gen) as libstdc++ computes it."""
    x = gen()
    u = F32(x) * F32(2.0 ** -64)  # float(x) rounds to nearest; the scale is exact
    if u >= F32(1.0):
        u = np.nextafter(F32(1.0), F32(0.0))
    return F32(u * (F32(b) - F32(a))) + F32(a)


def test_mt19937_64_standard_value():
    g = MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_obstacle_draws_match_library_distribution():
    g = MT19937_64(1919)
    want = np.array([uniform_float(g, -0.06, 0.06) for _ in range(48)], np.float32)
    got = oracle.uniform_float_draws(1919, -0.06, 0.06, 48)
    assert np.array_equal(want.view(np.uint32), got.view(np.uint32))
    assert np.all(np.abs(got) <= F32(0.06))


def _wall_boxes(support, length, obstacles):
    """solids.cpp:139-176 in float32."""
    H = F32(support)
    L = F32(length)
    b = H * F32(3.0)
    boxes = [((-b, -b), (F32(0.0), L + b)),
             ((L, -b), (L + b, L + b)),
             ((F32(-0.5) * b, -b), (L + F32(0.5) * b, F32(0.0))),
             ((F32(-0.5) * b, L), (L + F32(0.5) * b, L + b))]
    if obstacles:
        g = MT19937_64(1919)
        for by in range(4):
            byf = F32(by + 1) / F32(5)
            for bx in range(4):
                bxf = F32(bx + 1) / F32(5)
                ux = bxf + uniform_float(g, -0.06, 0.06)
                uy = byf + uniform_float(g, -0.06, 0.06)
                roff = uniform_float(g, -0.06, 0.06)
                r = F32(0.04) + abs(roff)
                boxes.append(((L * (ux - r), L * (uy - r)), (L * (ux + r), L * (uy + r))))
    return boxes


def _radius(support):
    return F32(0.99) * (F32(support) * F32(0.5))


@pytest.mark.parametrize("support", [0.025, 0.01, 0.004])
@pytest.mark.parametrize("obstacles", [False, True])
def test_world_walls_oracle_against_restatement(support, obstacles):
    R = _radius(support)
    parts = [oracle.emit_solid_box(box, R) for box in _wall_boxes(support, 1.0, obstacles)]
    want = np.concatenate(parts)
    got = oracle.world_walls_positions(support, 1.0, obstacles)
    assert np.array_equal(want.view(np.uint32), got.view(np.uint32))


@pytest.mark.parametrize("support", [0.025, 0.01, 0.004])
@pytest.mark.parametrize("obstacles", [False, True])
def test_world_walls_library_equals_oracle(support, obstacles):
    params = sb.Parameters(support=support)
    got = sb.world_walls_initial_solid_state(params, random_obstacles=obstacles).positions
    want = oracle.world_walls_positions(F32(support), 1.0, obstacles)
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _kill_brute_force(fluid, solids, two_r):
    """True where some solid is closer than 2R (float32 Imath length)."""
    kill = np.zeros(len(fluid), bool)
    for k in range(0, len(fluid), 512):
        f = fluid[k:k + 512]
        dx = (solids[None, :, 0] - f[:, None, 0]).astype(np.float32)
        dy = (solids[None, :, 1] - f[:, None, 1]).astype(np.float32)
        l2 = dx * dx + dy * dy
        assert np.all(l2 >= F32(2.0) * np.finfo(np.float32).tiny)  # no lengthTiny branch here
        kill[k:k + 512] = np.any(np.sqrt(l2) < two_r, axis=1)
    return kill


@pytest.mark.parametrize("support,length,obstacles",
                         [(0.025, 1.0, True), (0.025, 1.0, False), (0.01, 1.0, True),
                          (0.02, 0.75, True)])
def test_dam_break_oracle_against_brute_force(support, length, obstacles):
    solids = oracle.world_walls_positions(F32(support), F32(length), obstacles)
    R = _radius(support)
    L = F32(length)
    block = ((F32(0.0), F32(0.0)), (L / F32(2.0), F32(3.25) * L / F32(4.0)))
    emitted = oracle.emit_solid_box(block, R)
    kill = _kill_brute_force(emitted, solids, F32(2) * R)
    want = emitted[~kill]
    got = oracle.dam_break_positions(F32(support), solids, F32(length))
    # Without obstacles the block's first row and column sit exactly 2R or more
    # from the walls' last ones; the obstacles remove particles.
    assert not kill.all() and kill.any() == obstacles
    assert np.array_equal(want.view(np.uint32), got.view(np.uint32))


@pytest.mark.parametrize("support,obstacles", [(0.025, True), (0.025, False), (0.01, True),
                                               (0.005, True)])
def test_dam_break_library_equals_oracle(support, obstacles):
    params = sb.Parameters(support=support)
    solids = sb.world_walls_initial_solid_state(params, random_obstacles=obstacles)
    got = sb.dam_break_initial_state(params, solids).positions
    want = oracle.dam_break_positions(F32(support), solids.positions)
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_reference_default_scene_inputs():
    # EZ_EXAMPLE_SIM's own inputs (simulation.cpp:14-39: solid_init, then
    # fluid_init from the solids) with the default parameters (support 0.025,
    # length 1): the library's generated scene is the oracle's.
    params = sb.Parameters()
    solids = sb.world_walls_initial_solid_state(params)
    state = sb.dam_break_initial_state(params, solids)
    ws = oracle.world_walls_positions(F32(params.support), F32(params.length), True)
    fl = oracle.dam_break_positions(F32(params.support), ws, F32(params.length))
    assert np.array_equal(solids.positions.view(np.uint32), ws.view(np.uint32))
    assert np.array_equal(state.positions.view(np.uint32), fl.view(np.uint32))
    assert np.all(state.velocities == 0)
