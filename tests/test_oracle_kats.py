"""Pins the oracle against the reference's own known-answer tests, its compiled
z_index translation unit, O(N^2) brute force and closed forms (CPU only)."""

import math

import numpy as np
import pytest

import oracle

L = oracle.lib


# --- z_index: z_index/tests/test_z_index.cpp:11-70 -------------------------
def test_z_index_encode_kats():
    assert L().orc_z_index(0b111, 0b111) == 0b111111
    assert L().orc_z_index(0xFFFF, 0xFFFF) == 0xFFFFFFFF
    assert L().orc_z_index(0, 0xFFFF) == 0xAAAAAAAA
    assert L().orc_z_index(0xFFFFFFFF, 0xFFFFFFFF) == 0xFFFFFFFFFFFFFFFF


def _template_encode(i: int) -> int:
    # Single_axis_index<2, 31>::encode (z_index.h:11-24): bit b -> bit 2b.
    return sum(((i >> b) & 1) << (2 * b) for b in range(32))


def test_single_axis_kats():
    assert L().orc_single_axis_index_2(21430205) == _template_encode(21430205)
    assert L().orc_single_axis_index_2(0b111) == 0b10101
    assert L().orc_single_axis_index_2(0xFFFF) == 0x55555555
    assert L().orc_single_axis_index_2(0xFF00) == 0x55550000


def test_z_index_against_reference_build():
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("reference z_index.cpp not buildable here (reference tree absent)")
    rng = np.random.default_rng(11)
    xs = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64).tolist()
    ys = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64).tolist()
    edge = [0, 1, 0xFFFF, 0x10000, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF, 0xFFFFFFFE]
    for x in xs + edge:
        assert L().orc_single_axis_index_2(x) == ref.ref_single_axis_index_2(x)
        assert ref.ref_single_axis_template_2(x) == ref.ref_single_axis_index_2(x)
    for x, y in zip(xs + edge, ys + edge[::-1]):
        assert L().orc_z_index(x, y) == ref.ref_z_index_2(x, y)


# --- grid coords: sph_common/tests/test_common.cpp:42-63 --------------------
def test_grid_coord_kats():
    g = oracle.grid_coords(np.array([[-2.4, 9.11]], np.float32), 2.95)
    assert g.tolist() == [[-1, 3]]
    g = oracle.grid_coords(np.array([[4.1, 5.9], [-4.1, -5.9]], np.float32), 1.0)
    assert g.tolist() == [[4, 5], [-5, -6]]


# --- safe_divide: util/tests/test_safe_divide.cpp:40-71, 146-182 -------------
def test_safe_divide_kats():
    f = L().orc_is_safe_divide_f
    assert not f(0.0, 0.0)
    assert f(1e-37, 1e-37)
    assert not f(1e37, 1e-2)
    assert not f(1e38, 1e-1)
    assert f(-16.0, -4.0)
    assert f(1.0, -10.0)
    v = L().orc_is_safe_divide_v2f
    assert not v(0.0, 0.0, 0.0)
    assert v(1e-37, 1e-37, 1e-37)
    assert not v(1e37, 1e37, 1e-2)
    assert not v(1e38, 1e38, 1e-1)
    assert v(-16.0, 1.0, -4.0)
    assert v(1.0, -0.1, -10.0)
    d = L().orc_is_safe_divide_d
    assert not d(0.0, 0.0)
    assert d(1e-307, 1e-307)


# --- kernels: closed forms of sph_common/kernels.h:12-64 ----------------------
PI_F = float(np.float32(3.14159274101257324218750))


@pytest.mark.parametrize("h", [0.005, 1.0 / 2560.0, 0.025, 1.0])
def test_kernel_closed_forms(h):
    h = float(np.float32(h))
    k = h ** 3 * PI_F
    assert math.isclose(L().orc_W(0.0, h), 1.0 / k, rel_tol=2e-7)
    assert math.isclose(L().orc_W(h, h), 0.25 / k, rel_tol=2e-7)
    assert L().orc_W(2.0 * h, h) == 0.0
    assert L().orc_W(2.5 * h, h) == 0.0
    assert L().orc_dW(2.5 * h, h) == 0.0
    # dW is continuous at q = 1: -3/h^2 h + 9/4 h^2/h^3 = -3/4/h = -3 (1)^2 / (4h)
    assert math.isclose(L().orc_dW(h * (1 - 1e-6), h), L().orc_dW(h * (1 + 1e-6), h), rel_tol=1e-4)
    assert math.isclose(L().orc_dW(h, h), -0.75 / h / k, rel_tol=2e-6)
    # dW is the derivative of W (central difference, float64 reference).
    for q in (0.3, 0.8, 1.4, 1.9):
        r = q * h
        e = 1e-4 * h
        num = (L().orc_W(r + e, h) - L().orc_W(r - e, h)) / (2 * e)
        assert math.isclose(L().orc_dW(r, h), num, rel_tol=2e-2)


def test_gradw_antisymmetry_and_zero():
    out = np.empty(2, np.float32)
    out2 = np.empty(2, np.float32)
    rng = np.random.default_rng(3)
    for _ in range(200):
        x, y = (rng.random(2) - 0.5).astype(np.float32) * np.float32(0.02)
        L().orc_GradW(float(x), float(y), 0.01, out.ctypes.data)
        L().orc_GradW(float(-x), float(-y), 0.01, out2.ctypes.data)
        assert np.array_equal(out, -out2)
    L().orc_GradW(0.0, 0.0, 0.01, out.ctypes.data)
    assert out.tolist() == [0.0, 0.0]


def test_length_matches_ieee_sqrt():
    rng = np.random.default_rng(5)
    v = (rng.random((2000, 2)) - 0.5).astype(np.float32)
    ref = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1])  # float32 IEEE
    got = np.array([L().orc_length(float(a), float(b)) for a, b in v], np.float32)
    assert np.array_equal(got, ref)


# --- Imath Vec2<float>::length() below |v|^2 = 2 FLT_MIN --------------------
# Imath (IlmBase 2.5, ImathVec.h; absent from the image) returns lengthTiny()
# there: the components divided by the larger magnitude, the root of their
# squares, times that magnitude. The mathematics pins it: the result is the
# Euclidean length to within 2 float ulps even where x*x underflows, axis
# vectors come out exact, and scaling by a power of two scales the result
# exactly. A plain sqrtf(x*x + y*y) fails every one of these checks.
_TINY2 = np.float32(2.0) * np.finfo(np.float32).tiny


def _orc_len(v):
    return np.array([L().orc_length(float(a), float(b)) for a, b in v], np.float32)


def _plain_len(v):
    v = v.astype(np.float32)
    return np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1])


def _ulps(a, exact):
    """|a - exact| in units of the float32 spacing at exact (subnormal spacing
    2^-149 below FLT_MIN)."""
    a = a.astype(np.float64)
    m, e = np.frexp(exact)
    ulp = np.ldexp(1.0, np.maximum(e - 24, -149))
    return np.abs(a - exact) / ulp


def _tiny_vectors(n, seed):
    rng = np.random.default_rng(seed)
    ex = rng.integers(-149, -64, size=(n, 2))
    mant = rng.random((n, 2)) + 0.5
    sign = np.where(rng.random((n, 2)) < 0.5, -1.0, 1.0)
    v = (sign * np.ldexp(mant, ex)).astype(np.float32)
    v[rng.random(n) < 0.1, 1] = 0.0
    return v


def test_length_tiny_zero_and_axes():
    assert L().orc_length(0.0, 0.0) == 0.0
    assert L().orc_length(-0.0, 0.0) == 0.0
    for d in [1e-45, 3e-45, 1.2e-40, 5.877e-39, 1e-30, 1e-20]:
        d = float(np.float32(d))
        for v in [(d, 0.0), (0.0, d), (-d, 0.0), (0.0, -d)]:
            assert L().orc_length(*v) == np.float32(d), v
    # the plain formula loses these: d*d underflows to zero or a subnormal
    assert _plain_len(np.array([[1e-30, 0.0]], np.float32))[0] == 0.0


def test_length_tiny_against_exact_length():
    v = _tiny_vectors(4000, 11)
    exact = np.hypot(v[:, 0].astype(np.float64), v[:, 1].astype(np.float64))
    assert np.all(v[:, 0].astype(np.float64) ** 2 + v[:, 1].astype(np.float64) ** 2 < _TINY2)
    got = _orc_len(v)
    assert np.max(_ulps(got, exact)) <= 2.0
    plain = _plain_len(v)
    assert np.mean(_ulps(plain, exact) > 2.0) > 0.5


def test_length_tiny_power_of_two_scaling():
    # vectors in the tiny branch whose lengths are normal floats
    rng = np.random.default_rng(12)
    v = ((rng.random((500, 2)) - 0.5) * 2.0 ** -66).astype(np.float32)
    v = v[np.abs(v).max(axis=1) > 2.0 ** -70]
    base = _orc_len(v)
    for k in (1, 3, 7):
        s = np.float32(2.0 ** -k)
        assert np.array_equal(_orc_len(v * s), base * s)
    assert not np.array_equal(_plain_len(v * np.float32(2.0 ** -7)),
                              _plain_len(v) * np.float32(2.0 ** -7))


def test_length_branch_threshold():
    # At |v|^2 >= 2 FLT_MIN the result is the IEEE root of the float |v|^2.
    rng = np.random.default_rng(13)
    v = ((rng.random((2000, 2)) + 0.25) * 2.0 ** -63).astype(np.float32)
    l2 = v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]
    big = l2 >= _TINY2
    assert big.any() and (~big).any()
    got = _orc_len(v)
    assert np.array_equal(got[big], np.sqrt(l2[big]))
    exact = np.hypot(v[:, 0].astype(np.float64), v[:, 1].astype(np.float64))
    assert np.max(_ulps(got, exact)) <= 2.0


# --- sort / blocks: neighborhood.cpp:10-124 closed forms --------------------
def test_sort_and_blocks_closed_form():
    xy = oracle.fixture_points(17272, 7927, -100.0, 100.0)
    r = oracle.spatial_index(xy, 2.1)
    order = np.argsort(r["keys"], kind="stable")
    assert np.array_equal(r["sorted_idx"], order.astype(np.uint32))
    assert np.array_equal(r["sorted_keys"], r["keys"][order])
    change = np.concatenate([[0], (r["sorted_keys"][1:] != r["sorted_keys"][:-1]).astype(np.int64)])
    assert np.array_equal(r["block_indices"], np.cumsum(change).astype(np.uint32))
    nb = r["block_count"]
    assert nb == len(np.unique(r["keys"]))
    assert r["block_begin"][0] == 0 and r["block_end"][-1] == len(xy)
    assert np.array_equal(r["block_begin"][1:], r["block_end"][:-1])
    # Every block holds one key.
    for b in range(0, nb, 37):
        ks = r["sorted_keys"][r["block_begin"][b]:r["block_end"][b]]
        assert np.all(ks == ks[0])


# --- neighbourhoods: brute force of test_neighborhood.cpp:177-292 -------------
def test_neighborhood_brute_force_fixture():
    xy = oracle.fixture_points(17272, 7927, -100.0, 100.0)
    cell = np.float32(2.1)
    counts, idx, dist, vec = oracle.neighborhoods(xy, xy, True, float(cell))
    n = len(xy)
    for i in range(n):
        d = xy - xy[i]                      # float32 vector_to
        r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])
        js = np.flatnonzero(r < cell)
        js = js[js != i]
        c = counts[i]
        assert c == len(js)
        lst = idx[i, :c]
        assert set(lst.tolist()) == set(js.tolist())
        assert np.array_equal(dist[i, :c], r[lst])
        assert np.array_equal(vec[i, :c], d[lst])
        # neighbor_order (neighborhood.h:172-180): distance, then index
        keyd = dist[i, :c]
        ordered = np.lexsort((lst, keyd))
        assert np.array_equal(ordered, np.arange(c))


def _scan_order_first64(xy, i, cell):
    """Candidates of particle i in the reference's scan order (cells j-outer,
    i-inner; sorted slots within a cell) and the first 64 accepted."""
    g = np.floor(xy / np.float32(cell)).astype(np.int64)
    gi = g[i]
    acc = []
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            cand = np.flatnonzero((g[:, 0] == gi[0] + di) & (g[:, 1] == gi[1] + dj))
            for j in cand:  # increasing index == sorted-slot order in one cell
                if j == i:
                    continue
                d = xy[j] - xy[i]
                r = np.sqrt(d[0] * d[0] + d[1] * d[1])
                if r >= np.float32(cell):
                    continue
                acc.append((r, j))
                if len(acc) == 64:
                    return acc
    return acc


def test_neighborhood_cap_is_scan_order_truncation():
    rng = np.random.default_rng(19)
    xy = np.concatenate([
        (0.5 + 0.01 * rng.standard_normal((700, 2))),
        rng.random((300, 2)),
    ]).astype(np.float32)
    cell = 0.05
    counts, idx, dist, _ = oracle.neighborhoods(xy, xy, True, cell)
    assert counts.max() == 64
    checked = 0
    for i in np.flatnonzero(counts == 64)[:25]:
        acc = _scan_order_first64(xy, int(i), cell)
        expect = sorted(acc, key=lambda t: (t[0], t[1]))
        assert [j for _, j in expect] == idx[i].tolist()
        checked += 1
    assert checked == 25
