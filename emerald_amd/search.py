"""Stand-alone z-index sort and neighbour search on the GPU (no dynamics).

The reference's compute_all_neighbhorhoods fluid part
(sph2d_box/neighborhoods_from_state.cpp:9-53): grid coords, 64-bit Morton keys,
sorted (key, index) pairs, block indices and ranges, and 64-capped neighbour
lists ordered by (distance, index).
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._capi import check, lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class NeighborSearch:
    def __init__(self, capacity: int):
        h = ctypes.c_void_p()
        check(lib().emsph_search_create(int(capacity), ctypes.byref(h)), "emsph_search_create")
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().emsph_search_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_points(self, xy: np.ndarray) -> None:
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        check(lib().emsph_search_set_points(self._h, _ptr(xy), len(xy)), "search_set_points")
        self.n = len(xy)

    def run(self, cell_size: float, store_geometry: bool = False) -> float:
        ms = ctypes.c_float()
        check(lib().emsph_search_run(self._h, float(cell_size), 1 if store_geometry else 0,
                                     ctypes.byref(ms)), "search_run")
        return float(ms.value)

    def results(self, geometry: bool = False, lists: bool = True) -> dict:
        n = self.n
        out = {
            "grid": np.empty((n, 2), np.int32),
            "keys": np.empty(n, np.uint64),
            "sorted_keys": np.empty(n, np.uint64),
            "sorted_idx": np.empty(n, np.uint32),
            "block_indices": np.empty(n, np.uint32),
            "block_begin": np.empty(max(n, 1), np.uint32),
            "block_end": np.empty(max(n, 1), np.uint32),
        }
        counts = np.empty(n, np.uint8) if lists else None
        idx = np.empty((n, 64), np.uint32) if lists else None
        dist = np.empty((n, 64), np.float32) if geometry else None
        vec = np.empty((n, 64, 2), np.float32) if geometry else None
        nb = ctypes.c_size_t()
        check(lib().emsph_search_get(self._h, _ptr(out["grid"]), _ptr(out["keys"]),
                                     _ptr(out["sorted_keys"]), _ptr(out["sorted_idx"]),
                                     _ptr(out["block_indices"]), _ptr(out["block_begin"]),
                                     _ptr(out["block_end"]), ctypes.byref(nb), _ptr(counts),
                                     _ptr(idx), _ptr(dist), _ptr(vec)), "search_get")
        out["block_count"] = int(nb.value)
        out["block_begin"] = out["block_begin"][:nb.value].copy()
        out["block_end"] = out["block_end"][:nb.value].copy()
        out["counts"], out["indices"], out["distances"], out["vectors"] = counts, idx, dist, vec
        return out
