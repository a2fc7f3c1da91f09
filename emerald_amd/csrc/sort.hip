// Stable LSD radix sort of (64-bit Morton key, 32-bit particle index) pairs and
// the device-wide exclusive scan it is built on, written for gfx950 wave64.
//
// The reference sorts (key, index) pairs with tbb::parallel_sort
// (sph_common/neighborhood.cpp:18-26). Pairs are unique and the indices enter
// in increasing order, so a stable sort by key yields the same sequence.
//
// One pass sorts one 8-bit digit:
//   rs_hist    per-tile digit counts (wave-aggregated LDS histograms)
//   scan       exclusive scan of the digit-major [256][tiles] count matrix
//   rs_scatter stable per-tile ranking (wave match via 8 ballots, per-wave
//              counts in LDS) and scatter to the scanned offsets
// Only the digits below the highest bit in which min(key) and max(key) differ
// are sorted: every key shares the bits above it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace emsph {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 4;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 4;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;
// Up to this many scan blocks each block sums the block sums before it
// itself (at most ~8M block-sum reads), instead of a separate top-level scan.
constexpr uint32_t SCAN_DIRECT_BLOCKS = 4096;

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t const lane = __lane_id();
    return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

// Lanes of this wave holding the same 8-bit digit (and valid).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        bool const bit = (d >> b) & 1u;
        uint64_t const m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// Digit of element i: bits [shift, shift+8) of its key, or of scene_of[val]
// for the scene passes.
__device__ __forceinline__ uint32_t sort_digit(uint64_t key, uint32_t val, uint32_t shift,
                                               uint32_t const* __restrict__ scene_of) {
    if (scene_of) { return (scene_of[val] >> shift) & 255u; }
    return static_cast<uint32_t>((key >> shift) & 255u);
}

__global__ __launch_bounds__(RS_THREADS) void rs_hist(uint64_t const* __restrict__ keys,
                                                      uint32_t const* __restrict__ vals,
                                                      uint32_t n, uint32_t shift,
                                                      uint32_t const* __restrict__ scene_of,
                                                      uint32_t* __restrict__ counts,
                                                      uint32_t tiles) {
    __shared__ uint32_t h[RS_WAVES][256];
    uint32_t const tid = threadIdx.x;
    uint32_t const wave = tid / 64;
    for (int w = 0; w < RS_WAVES; ++w) { h[w][tid] = 0; }
    __syncthreads();
    uint32_t const base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        uint32_t const i = base + r * RS_THREADS + tid;
        bool const valid = i < n;
        uint32_t const d =
          valid ? sort_digit(scene_of ? 0ull : keys[i], scene_of ? vals[i] : 0u, shift, scene_of)
                : 0u;
        uint64_t const peers = match_digit(d, valid);
        if (valid && (peers & lanemask_lt()) == 0) {
            atomicAdd(&h[wave][d], static_cast<uint32_t>(__popcll(peers)));
        }
    }
    __syncthreads();
    uint32_t s = 0;
    for (int w = 0; w < RS_WAVES; ++w) { s += h[w][tid]; }
    counts[tid * tiles + blockIdx.x] = s;
}

__global__ __launch_bounds__(RS_THREADS) void rs_scatter(
  uint64_t const* __restrict__ kin, uint32_t const* __restrict__ vin,
  uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, uint32_t shift,
  uint32_t const* __restrict__ scene_of, uint32_t const* __restrict__ offsets, uint32_t tiles) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_WAVES][256];
    uint32_t const tid = threadIdx.x;
    uint32_t const wave = tid / 64;
    run[tid] = offsets[tid * tiles + blockIdx.x];
    for (int w = 0; w < RS_WAVES; ++w) { wcnt[w][tid] = 0; }
    __syncthreads();
    uint32_t const base = blockIdx.x * RS_TILE;
    uint64_t const lt = lanemask_lt();
    for (int r = 0; r < RS_ITEMS; ++r) {
        uint32_t const i = base + r * RS_THREADS + tid;
        bool const valid = i < n;
        uint64_t const k = valid ? kin[i] : 0ull;
        uint32_t const v = valid ? vin[i] : 0u;
        uint32_t const d = valid ? sort_digit(k, v, shift, scene_of) : 0u;
        uint64_t const peers = match_digit(d, valid);
        bool const leader = valid && (peers & lt) == 0;
        uint32_t const cnt = static_cast<uint32_t>(__popcll(peers));
        if (leader) { wcnt[wave][d] = cnt; }
        __syncthreads();
        uint32_t pos = 0;
        if (valid) {
            pos = run[d] + static_cast<uint32_t>(__popcll(peers & lt));
            for (uint32_t w = 0; w < wave; ++w) { pos += wcnt[w][d]; }
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&run[d], cnt);
            wcnt[wave][d] = 0;
        }
        __syncthreads();
        if (valid) {
            kout[pos] = k;
            vout[pos] = v;
        }
    }
}

//------------------------------------------------------------------------------
// Exclusive scan of uint32 (reduce -> top scan -> downsweep).
template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
    uint32_t const lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T const y = __shfl_up(x, off);
        if (lane >= static_cast<uint32_t>(off)) { x += y; }
    }
    return x;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
template <int THREADS, class T>
__device__ __forceinline__ T block_exclusive_scan(T x, T& total) {
    __shared__ T wsum[THREADS / 64];
    uint32_t const lane = __lane_id();
    uint32_t const wave = threadIdx.x / 64;
    T const inc = wave_inclusive_scan(x);
    if (lane == 63) { wsum[wave] = inc; }
    __syncthreads();
    T prefix = 0;
    T tot = 0;
    for (int w = 0; w < THREADS / 64; ++w) {
        T const s = wsum[w];
        if (static_cast<uint32_t>(w) < wave) { prefix += s; }
        tot += s;
    }
    __syncthreads();
    total = tot;
    return prefix + inc - x;
}

// Scan element of a count: the count itself (uint32), or for the cell scan
// the count in the low CELL_SHIFT bits and 1 above them when it is nonzero,
// so one uint64 scan gives both the first slot of every bin and the number of
// nonempty bins (cells) before it (counts sum to below 2^CELL_SHIFT).
constexpr int CELL_SHIFT = 33;  // counts and cells are at most 2^30
template <class T>
__device__ __forceinline__ T scan_item(uint32_t c) {
    if constexpr (sizeof(T) == 8) {
        return static_cast<T>(c) | (c ? (T{1} << CELL_SHIFT) : T{0});
    } else {
        return c;
    }
}

template <class T>
__global__ __launch_bounds__(SC_THREADS) void scan_reduce(uint32_t const* __restrict__ in,
                                                          uint32_t n, T* __restrict__ bsum) {
    uint32_t const base = blockIdx.x * SC_TILE;
    T s = 0;
    for (int r = 0; r < SC_ITEMS; ++r) {
        uint32_t const i = base + r * SC_THREADS + threadIdx.x;
        if (i < n) { s += scan_item<T>(in[i]); }
    }
    for (int off = 32; off > 0; off >>= 1) { s += __shfl_xor(s, off); }
    __shared__ T ws[SC_THREADS / 64];
    if (__lane_id() == 0) { ws[threadIdx.x / 64] = s; }
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = 0;
        for (int w = 0; w < SC_THREADS / 64; ++w) { t += ws[w]; }
        bsum[blockIdx.x] = t;
    }
}

// The scan's outputs: element prefixes into out (uint32; for the cell scan
// the slot part) and, for the cell scan, cell prefixes into cells; the grand
// total into *total and, for the cell scan, the cell total into *cell_total
// (each when not null).
template <class T>
__device__ __forceinline__ void scan_store(T v, uint32_t* out, uint32_t* cells, uint32_t i) {
    if constexpr (sizeof(T) == 8) {
        out[i] = static_cast<uint32_t>(v & ((T{1} << CELL_SHIFT) - 1));
        cells[i] = static_cast<uint32_t>(v >> CELL_SHIFT);
    } else {
        out[i] = v;
    }
}
template <class T>
__device__ __forceinline__ void scan_total(T v, uint32_t* total, uint32_t* cell_total) {
    if constexpr (sizeof(T) == 8) {
        if (total) { *total = static_cast<uint32_t>(v & ((T{1} << CELL_SHIFT) - 1)); }
        if (cell_total) { *cell_total = static_cast<uint32_t>(v >> CELL_SHIFT); }
    } else {
        if (total) { *total = v; }
    }
}

// Single workgroup: exclusive scan of nb block sums in place; writes the grand
// total(s) (scan_total).
template <class T>
__global__ __launch_bounds__(1024) void scan_top(T* __restrict__ bsum, uint32_t nb,
                                                 uint32_t* __restrict__ total,
                                                 uint32_t* __restrict__ cell_total) {
    uint32_t const per = (nb + 1023) / 1024;
    uint32_t const lo = threadIdx.x * per;
    T s = 0;
    for (uint32_t i = 0; i < per; ++i) {
        uint32_t const j = lo + i;
        if (j < nb) { s += bsum[j]; }
    }
    T tot = 0;
    T run = block_exclusive_scan<1024>(s, tot);
    for (uint32_t i = 0; i < per; ++i) {
        uint32_t const j = lo + i;
        if (j < nb) {
            T const v = bsum[j];
            bsum[j] = run;
            run += v;
        }
    }
    if (threadIdx.x == 0) { scan_total(tot, total, cell_total); }
}

// Downsweep. With nb > 0 the block's offset is the sum of the block sums
// before it (read here, so no separate top-level scan runs; used while nb is
// small), and the last block writes the grand total(s) (scan_total); with
// nb == 0, bsum already holds the scanned block offsets.
template <class T>
__global__ __launch_bounds__(SC_THREADS) void scan_down(uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out,
                                                        uint32_t* __restrict__ cells, uint32_t n,
                                                        T const* __restrict__ bsum, uint32_t nb,
                                                        uint32_t* __restrict__ total,
                                                        uint32_t* __restrict__ cell_total,
                                                        int clear_in) {
    __shared__ T tile[SC_TILE];
    __shared__ T block_offset;
    if (nb) {
        T part = 0;
        for (uint32_t i = threadIdx.x; i < blockIdx.x; i += SC_THREADS) { part += bsum[i]; }
        for (int off = 32; off > 0; off >>= 1) { part += __shfl_xor(part, off); }
        __shared__ T ws[SC_THREADS / 64];
        if (__lane_id() == 0) { ws[threadIdx.x / 64] = part; }
        __syncthreads();
        if (threadIdx.x == 0) {
            T o = 0;
            for (int w = 0; w < SC_THREADS / 64; ++w) { o += ws[w]; }
            block_offset = o;
            if (blockIdx.x == nb - 1) { scan_total<T>(o + bsum[blockIdx.x], total, cell_total); }
        }
    } else if (threadIdx.x == 0) {
        block_offset = bsum[blockIdx.x];
    }
    uint32_t const base = blockIdx.x * SC_TILE;
    for (int r = 0; r < SC_ITEMS; ++r) {
        uint32_t const i = base + r * SC_THREADS + threadIdx.x;
        tile[r * SC_THREADS + threadIdx.x] = (i < n) ? scan_item<T>(in[i]) : T{0};
        if (clear_in && i < n) { in[i] = 0u; }  // each element is read by this thread only
    }
    __syncthreads();
    // Each thread scans SC_ITEMS consecutive elements.
    uint32_t const t0 = threadIdx.x * SC_ITEMS;
    T s = 0;
    for (int r = 0; r < SC_ITEMS; ++r) { s += tile[t0 + r]; }
    T tot = 0;
    __syncthreads();  // block_offset
    T run = block_exclusive_scan<SC_THREADS>(s, tot) + block_offset;
    for (int r = 0; r < SC_ITEMS; ++r) {
        T const v = tile[t0 + r];
        tile[t0 + r] = run;
        run += v;
    }
    __syncthreads();
    for (int r = 0; r < SC_ITEMS; ++r) {
        uint32_t const i = base + r * SC_THREADS + threadIdx.x;
        if (i < n) { scan_store<T>(tile[r * SC_THREADS + threadIdx.x], out, cells, i); }
    }
}

//------------------------------------------------------------------------------
// Host launchers.

size_t scan_scratch_words(size_t n) { return (n + SC_TILE - 1) / SC_TILE + 1; }

namespace {
template <class T>
void exclusive_scan(uint32_t* in, uint32_t* out, uint32_t* cells, uint32_t n, T* scratch,
                    uint32_t* total, uint32_t* cell_total, int clear_in, hipStream_t s) {
    if (n == 0) { return; }
    uint32_t const nb = (n + SC_TILE - 1) / SC_TILE;
    hipLaunchKernelGGL(scan_reduce<T>, dim3(nb), dim3(SC_THREADS), 0, s, in, n, scratch);
    if (nb <= SCAN_DIRECT_BLOCKS) {
        hipLaunchKernelGGL(scan_down<T>, dim3(nb), dim3(SC_THREADS), 0, s, in, out, cells, n,
                           scratch, nb, total, cell_total, clear_in);
    } else {
        hipLaunchKernelGGL(scan_top<T>, dim3(1), dim3(1024), 0, s, scratch, nb, total,
                           cell_total);
        hipLaunchKernelGGL(scan_down<T>, dim3(nb), dim3(SC_THREADS), 0, s, in, out, cells, n,
                           scratch, 0u, nullptr, nullptr, clear_in);
    }
}
}  // namespace

void launch_exclusive_scan(uint32_t const* in, uint32_t* out, uint32_t n, uint32_t* scratch,
                           uint32_t* total, hipStream_t s) {
    exclusive_scan<uint32_t>(const_cast<uint32_t*>(in), out, nullptr, n, scratch, total, nullptr,
                             0, s);
}

void launch_exclusive_scan_clear(uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch,
                                 uint32_t* total, hipStream_t s) {
    exclusive_scan<uint32_t>(in, out, nullptr, n, scratch, total, nullptr, 1, s);
}

void launch_cell_scan_clear(uint32_t* counts, uint32_t* first, uint32_t* cells, uint32_t n,
                            uint32_t* scratch, uint32_t* total, uint32_t* cell_total,
                            hipStream_t s) {
    exclusive_scan<unsigned long long>(counts, first, cells, n,
                                       reinterpret_cast<unsigned long long*>(scratch), total,
                                       cell_total, 1, s);
}

size_t radix_counts_words(size_t n) {
    size_t const tiles = (n + RS_TILE - 1) / RS_TILE;
    return 256 * tiles;
}

int radix_pass_count(uint64_t kmin, uint64_t kmax) {
    uint64_t const diff = kmin ^ kmax;
    if (diff == 0) { return 0; }
    int const bits = 64 - __builtin_clzll(diff);
    return (bits + 7) / 8;
}

int launch_radix_sort(uint64_t* keys0, uint32_t* vals0, uint64_t* keys1, uint32_t* vals1,
                      uint32_t n, int key_passes, uint32_t const* scene_of, int scene_passes,
                      uint32_t* counts, uint32_t* scan_scratch, hipStream_t s) {
    uint32_t const tiles = (n + RS_TILE - 1) / RS_TILE;
    uint64_t* kin = keys0;
    uint32_t* vin = vals0;
    uint64_t* kout = keys1;
    uint32_t* vout = vals1;
    int const passes = key_passes + (scene_of ? scene_passes : 0);
    for (int p = 0; p < passes; ++p) {
        bool const scene_pass = p >= key_passes;
        uint32_t const shift = 8u * static_cast<uint32_t>(scene_pass ? p - key_passes : p);
        uint32_t const* so = scene_pass ? scene_of : nullptr;
        hipLaunchKernelGGL(rs_hist, dim3(tiles), dim3(RS_THREADS), 0, s, kin, vin, n, shift, so,
                           counts, tiles);
        launch_exclusive_scan(counts, counts, 256u * tiles, scan_scratch, nullptr, s);
        hipLaunchKernelGGL(rs_scatter, dim3(tiles), dim3(RS_THREADS), 0, s, kin, vin, kout, vout,
                           n, shift, so, counts, tiles);
        uint64_t* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    // Result buffer index: 0 if in keys0/vals0, 1 otherwise.
    return passes & 1;
}

}  // namespace emsph
