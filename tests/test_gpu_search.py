"""GPU z-index keys, sort, cell ranges and neighbour lists vs the oracle, bit for bit.

Reference behaviour: sph_common/common.cpp:43-76 (grid coords, keys),
neighborhood.cpp:10-124 (pairs, sort, block indices, blocks, map),
neighborhood.h:198-282 (64-capped lists in (distance, index) order).
"""

import numpy as np
import pytest

import oracle
from emerald_amd import scenes
from emerald_amd.search import NeighborSearch
from helpers import assert_bits_equal

pytestmark = pytest.mark.gpu


def _compare(xy, cell, geometry=True, s=None):
    xy = np.ascontiguousarray(xy, np.float32)
    n = len(xy)
    print(f"[search parity] n={n} cell={cell}", flush=True)
    own = s is None
    if own:
        s = NeighborSearch(max(n, 1))
    s.set_points(xy)
    s.run(cell, store_geometry=geometry)
    g = s.results(geometry=geometry)
    o = oracle.spatial_index(xy, cell)
    assert g["block_count"] == o["block_count"]
    for k in ("grid", "keys", "sorted_keys", "sorted_idx", "block_indices", "block_begin",
              "block_end"):
        assert np.array_equal(g[k], o[k]), k
    oc, oi, od, ov = oracle.neighborhoods(xy, xy, True, cell)
    assert np.array_equal(g["counts"], oc), "counts"
    mask = np.arange(64)[None, :] < oc[:, None]
    assert np.array_equal(np.where(mask, g["indices"], 0), np.where(mask, oi, 0)), "indices"
    if geometry:
        assert_bits_equal(np.where(mask, g["distances"], 0), np.where(mask, od, 0), "distances")
        m2 = mask[:, :, None]
        assert_bits_equal(np.where(m2, g["vectors"], 0), np.where(m2, ov, 0), "vectors")
    if own:
        s.close()
    return g


def test_reference_fixture_7927(gpu):
    # sph_common/tests/test_neighborhood.cpp:18-31 (mt19937_64{17272}, U(-100,100)^2,
    # cell 2.1): negative cells give keys with the high bits set.
    xy = oracle.fixture_points(17272, 7927, -100.0, 100.0)
    g = _compare(xy, 2.1)
    assert (g["keys"] > np.uint64(1 << 62)).any()


def test_clustered_points_cap_64(gpu):
    # Many candidates per particle: the 64-cap in scan order is live.
    rng = np.random.default_rng(7)
    core = (0.5 + 0.002 * rng.standard_normal((6000, 2))).astype(np.float32)
    bg = rng.random((20000, 2)).astype(np.float32)
    xy = np.concatenate([bg, core])
    g = _compare(xy, 1.0 / 256.0)
    assert g["counts"].max() == 64


def test_list_capacity_switches(gpu):
    # One search object run by turns on dense and sparse points: the build's
    # LDS list capacity follows the previous run's longest list (32 on the
    # first run, so the dense run defers its long lists to the full-capacity
    # pass; then 64; then back down to 16 on sparse points).
    rng = np.random.default_rng(11)
    core = (0.5 + 0.002 * rng.standard_normal((6000, 2))).astype(np.float32)
    dense = np.concatenate([rng.random((20000, 2)).astype(np.float32), core])
    sparse = rng.random((len(dense), 2)).astype(np.float32)
    s = NeighborSearch(len(dense))
    try:
        for xy, cell in ((dense, 1.0 / 256.0), (dense, 1.0 / 256.0), (sparse, 1.0 / 512.0),
                         (sparse, 1.0 / 512.0), (dense, 1.0 / 256.0)):
            _compare(xy, cell, s=s)
    finally:
        s.close()


def test_mostly_long_lists_take_the_full_build(gpu):
    # Nearly every list is capped at 64 in scan order: the first run defers
    # them from the 32-entry build to the full-capacity pass, the next runs
    # build every list at the full 64-entry capacity.
    rng = np.random.default_rng(13)
    xy = (0.5 + 0.01 * rng.standard_normal((12000, 2))).astype(np.float32)
    s = NeighborSearch(len(xy))
    try:
        for _ in range(3):
            g = _compare(xy, 1.0 / 256.0, s=s)
        c = g["counts"]
        assert (c > 32).mean() > 0.5 and c.max() == 64
    finally:
        s.close()


def test_few_long_lists_deferred_at_cap_16(gpu):
    # Sparse points with a small tight cluster: the cluster's lists (above 16
    # entries) are under 1/64 of the particles, so the second run keeps the
    # 16-entry build and defers them to the full-capacity pass.
    rng = np.random.default_rng(12)
    cluster = (0.3 + 0.0005 * rng.standard_normal((200, 2))).astype(np.float32)
    xy = np.concatenate([rng.random((20000, 2)).astype(np.float32), cluster])
    s = NeighborSearch(len(xy))
    try:
        for _ in range(3):
            g = _compare(xy, 1.0 / 100.0, s=s)
        assert g["counts"].max() > 16
    finally:
        s.close()


def test_lattice_distance_ties(gpu):
    # An unjittered lattice has exact distance ties, broken by original index.
    i, j = np.meshgrid(np.arange(60), np.arange(50), indexing="ij")
    xy = np.stack([(i.ravel() + 0.5) * 0.005, (j.ravel() + 0.5) * 0.005], 1).astype(np.float32)
    _compare(xy, 0.02)


def test_near_coincident_tiny_lengths(gpu):
    # Points near the origin at separations far below sqrt(2 FLT_MIN): their
    # float |v|^2 underflows, so every distance among them is Imath's
    # lengthTiny (neighborhood.h:263); multiples of 2^-100, of the smallest
    # subnormal, and offsets around the branch threshold (|v| ~ 1.5e-19).
    rng = np.random.default_rng(21)
    i, j = np.meshgrid(np.arange(5), np.arange(4), indexing="ij")
    grid = np.stack([np.ldexp(1.0 + i.ravel(), -100), np.ldexp(1.0 + j.ravel(), -100)], 1)
    sub = np.stack([np.arange(1, 9) * 1.4e-45, (np.arange(8) % 3) * 1.4e-45], 1)
    thr = rng.random((12, 2)) * 3e-19
    bg = rng.random((3000, 2))
    xy = np.concatenate([bg, grid, sub, thr]).astype(np.float32)
    g = _compare(xy, 1.0 / 64.0)
    assert g["counts"][3000:].min() >= 39  # the 40 points near the origin see each other


def test_search_points_sample(gpu):
    xy = scenes.search_points(200_000, seed=3)
    _compare(xy, 1.0 / 2048.0, geometry=False)


def test_single_and_empty(gpu):
    _compare(np.array([[0.25, 0.75]], np.float32), 0.1)
    s = NeighborSearch(16)
    s.set_points(np.zeros((0, 2), np.float32))
    s.run(0.1)
    assert s.results()["block_count"] == 0


@pytest.mark.parametrize("n", [4_194_304])
def test_full_size_search_sortedness(gpu, n):
    """4,194,304 points, h = 1/4096, cell 1/2048: sorted order and lists against
    size-independent properties (the oracle check runs on a sample)."""
    xy = scenes.search_points(n, seed=3)
    print(f"[search full] n={n}", flush=True)
    s = NeighborSearch(n)
    s.set_points(xy)
    s.run(1.0 / 2048.0)
    r = s.results()
    sk = r["sorted_keys"]
    si = r["sorted_idx"].astype(np.int64)
    assert np.all(sk[1:] >= sk[:-1])
    same = sk[1:] == sk[:-1]
    assert np.all(si[1:][same] > si[:-1][same])       # ties keep index order
    assert np.array_equal(np.sort(si), np.arange(n))  # a permutation
    # Lists on a sample of particles against the oracle's full-array lists.
    oc, oi, _, _ = oracle.neighborhoods(xy, xy, True, 1.0 / 2048.0)
    assert np.array_equal(r["counts"], oc)
    mask = np.arange(64)[None, :] < oc[:, None]
    assert np.array_equal(np.where(mask, r["indices"], 0), np.where(mask, oi, 0))


def test_repeatability(gpu):
    """test_repeatability.cpp:53-112: grid coordinates, keys, sorted pairs,
    blocks and neighbour indices identical over repeated runs (here 20 runs of
    one search object, the first with the exact sort and list capacity unknown,
    the later ones on the dense index and the learned capacity)."""
    xy = scenes.search_points(100_000, seed=11)
    s = NeighborSearch(len(xy))
    s.set_points(xy)
    first = None
    for k in range(20):
        s.run(1.0 / 2048.0)
        r = s.results()
        mask = np.arange(64)[None, :] < r["counts"][:, None]
        r["indices"] = np.where(mask, r["indices"], 0)
        if first is None:
            first = r
            continue
        for key, v in first.items():
            assert np.array_equal(np.asarray(r[key]), np.asarray(v)), f"run {k}: {key}"
    s.close()
