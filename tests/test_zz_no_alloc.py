"""emsph_search_lists_noalloc (the search's lists with no allocation) against
the allocating search entry points (emsph_search_run + emsph_search_get), bit
for bit, on two calls in a row, the second with a smaller input.

Device buffers for the inputs and outputs are the caller's: they are allocated
here through the HIP runtime the library itself loaded."""

import ctypes

import numpy as np
import pytest

from emerald_amd import lib
from emerald_amd.search import NeighborSearch

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    lib()
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                h = ctypes.CDLL(line.split()[-1])
                h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
                h.hipFree.argtypes = [ctypes.c_void_p]
                h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.c_int]
                return h
    raise RuntimeError("the library's HIP runtime is not loaded")


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(nbytes, 1)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def _points(n, seed):
    rng = np.random.default_rng(seed)
    core = (0.4 + 0.004 * rng.standard_normal((n // 5, 2)))
    bg = rng.random((n - n // 5, 2))
    return np.concatenate([bg, core]).astype(np.float32)


def _allocating(xy, cell):
    s = NeighborSearch(len(xy))
    try:
        s.set_points(xy)
        s.run(cell)
        r = s.results(lists=True)
        return r["counts"], r["indices"]
    finally:
        s.close()


def _noalloc(hip, bufs, xy, cell):
    n = len(xy)
    bufs["xy"].put(xy)
    st = lib().emsph_search_lists_noalloc(bufs["xy"].p, n, cell, bufs["counts"].p,
                                          bufs["idx"].p, None)
    assert st == 0, lib().emsph_last_error()
    counts = bufs["counts"].get(np.empty(n, np.uint8))
    idx = bufs["idx"].get(np.empty((n, 64), np.uint32))
    return counts, idx


def test_noalloc_lists_equal_allocating_search(gpu):
    hip = _hip()
    nmax = 65536
    bufs = {"xy": DevBuf(hip, nmax * 8), "counts": DevBuf(hip, nmax),
            "idx": DevBuf(hip, nmax * 64 * 4)}
    try:
        for n, cell, seed in ((60000, 1.0 / 300.0, 1), (7000, 1.0 / 64.0, 2)):
            xy = _points(n, seed)
            print(f"[noalloc] n={n} cell={cell}", flush=True)
            gc, gi = _noalloc(hip, bufs, xy, cell)
            ac, ai = _allocating(xy, cell)
            assert np.array_equal(gc, ac)
            assert np.array_equal(gi, ai)
            assert gc.max() == 64  # the cluster's lists are capped
        # Above the scratch's maximum: a clear error, nothing run.
        st = lib().emsph_search_lists_noalloc(bufs["xy"].p, nmax + 1, 0.01, bufs["counts"].p,
                                              bufs["idx"].p, None)
        assert st == 5 and b"65536" in lib().emsph_last_error()
    finally:
        for b in bufs.values():
            b.free()
