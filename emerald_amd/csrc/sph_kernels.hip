// HIP kernels of the 2D SPH sub-step for gfx950, over a batch of scenes.
//
// Layout: particle fields live in sorted order ("slots"): scenes one after
// another, each in Morton order, for the duration of a sub-step; neighbour
// lists are slot-major ([k * stride + t]) so that the thread of slot t reads
// entry k of consecutive slots with one coalesced load. Fluid kernels run one
// workgroup per scene-aligned tile and one thread per particle, summing its
// neighbours in the reference's list order (distance, then original index), so
// every float result matches the reference CPU path bit for bit.
//
// The reference runs separate fluid and solid passes for each sum
// (e.g. density.cpp:84-109); both passes only write the particle's own element
// and read others' inputs that the pass does not modify, so one thread can run
// the fluid pass and then the solid pass of its particle without changing any
// result.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>

#include "device_math.hpp"
#include "kernels.hpp"
#include "sort_networks.hpp"

namespace emsph {

// Never launched: its size places the neighbour build that follows it in this
// file's code object (k_build<BUILD_IISPH, 16, 64, false>) at byte 3840 of a
// 4 KB page. With the same instructions, that kernel's time at 1M particles
// depends on its offset in the page (104-121 us over the offsets measured,
// repeatably; the hardware mechanism is not established), and 3840 was the
// fastest. The build step checks the offset and warns when an edit moved it.
__global__ void k_layout_pad() { asm volatile(".rept 960\n s_nop 0\n.endr"); }

constexpr int BUILD_THREADS = 64;  // the full-capacity build and deferred pass
constexpr int BUILD_WG = 64;        // the capped builds (16 or 32 entries)

// The tile of this workgroup (XCD-aware order: each XCD walks one contiguous
// run of tiles, i.e. of Morton-ordered space, and keeps its gathers in its L2).
struct TileRef {
    uint32_t scene;
    uint32_t t;       // this thread's slot
    uint32_t tile;
    bool valid;
};
// This workgroup's tile: remapped so that each XCD walks one contiguous run of
// tiles (Batch::xcd_remap, a tuning choice), or in launch order.
__device__ __forceinline__ uint32_t tile_index(Batch const& b) {
    return b.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
}

// This workgroup's tile in a launch over ActiveTiles (uniform per workgroup,
// so the search runs on scalar registers).
__device__ __forceinline__ uint32_t active_tile(Batch const& b, ActiveTiles const& at) {
    uint32_t const k = tile_index(b);
    if (at.n == 0) { return k; }
    uint32_t j = 0;
    for (uint32_t i = 1; i < at.n; ++i) { j = (at.start[i] <= k) ? i : j; }
    return at.first[j] + (k - at.start[j]);
}

__device__ __forceinline__ TileRef tile_ref_of(Batch const& b, uint32_t k) {
    Tile const tl = b.tiles[k];
    TileRef r;
    r.scene = tl.scene;
    r.tile = k;
    r.t = tl.start + threadIdx.x;
    r.valid = threadIdx.x < tl.count;
    return r;
}

__device__ __forceinline__ TileRef tile_ref(Batch const& b, uint32_t threads) {
    (void)threads;
    return tile_ref_of(b, tile_index(b));
}

// Workgroup partials, written to the workgroup's own tile slot (no atomics:
// one address taken by every workgroup serialises at the device-coherent level
// across the XCDs). The consumer reduces a scene's tile slots: the loop
// decision (residuals) and the end-of-sub-step scene totals (max |v|^2).
// Quantised residual: integer sum and max of the (non-negative) float bits.
template <int THREADS>
__device__ __forceinline__ void block_residual(unsigned long long v, unsigned int m,
                                               Reduction* out) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off);
        unsigned int const o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    __shared__ unsigned long long sv[THREADS / 64];
    __shared__ unsigned int sm[THREADS / 64];
    if (__lane_id() == 0) {
        sv[threadIdx.x / 64] = v;
        sm[threadIdx.x / 64] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tv = 0;
        unsigned int tm = 0;
        for (int w = 0; w < THREADS / 64; ++w) {
            tv += sv[w];
            tm = sm[w] > tm ? sm[w] : tm;
        }
        out->isum = tv;
        out->max_bits = tm;
        out->pad = 0;
    }
}

template <int THREADS>
__device__ __forceinline__ void block_max(unsigned int m, unsigned int* out) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned int const o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    __shared__ unsigned int sm[THREADS / 64];
    if (__lane_id() == 0) { sm[threadIdx.x / 64] = m; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int tm = 0;
        for (int w = 0; w < THREADS / 64; ++w) { tm = sm[w] > tm ? sm[w] : tm; }
        *out = tm;
    }
}

// FP64 energy of a workgroup's particles into its tile's partial (fixed
// shuffle order, so the sum is the same in every run). Every thread calls it.
template <int THREADS>
__device__ __forceinline__ void block_energy(double ke, double pe, double2* out) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ke += __shfl_xor(ke, off);
        pe += __shfl_xor(pe, off);
    }
    __shared__ double ske[THREADS / 64], spe[THREADS / 64];
    if (__lane_id() == 0) {
        ske[threadIdx.x / 64] = ke;
        spe[threadIdx.x / 64] = pe;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tk = 0.0, tp = 0.0;
        for (int w = 0; w < THREADS / 64; ++w) {
            tk += ske[w];
            tp += spe[w];
        }
        *out = make_double2(tk, tp);
    }
}

// Kinetic and potential energy of one particle in FP64.
__device__ __forceinline__ void particle_energy(StepConsts const& c, V2 x, V2 v, double& ke,
                                                double& pe) {
    double const m = c.mass;
    ke = 0.5 * m * (static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y);
    pe = m * static_cast<double>(c.gravity) * static_cast<double>(x.y);
}

//------------------------------------------------------------------------------
// Keys: sph_common/common.cpp:43-76 (grid coordinate, z key) and
// Float-to-integer conversions as the x86-64 host code of the frame loop
// performs them (cvttsd2si): truncation, and the "integer indefinite" value
// (the most negative integer) for NaN and out-of-range inputs, which the host
// meets for velocities near rest (adaptive_time_step.h:73-78).
__device__ __forceinline__ int64_t host_double_to_i64(double x) {
    if (!(x > -9223372036854775808.0 && x < 9223372036854775808.0)) {
        return static_cast<int64_t>(0x8000000000000000ull);
    }
    return static_cast<int64_t>(x);
}
__device__ __forceinline__ int32_t host_double_to_i32(double x) {
    if (!(x > -2147483649.0 && x < 2147483648.0)) { return static_cast<int32_t>(0x80000000u); }
    return static_cast<int32_t>(x);
}
__device__ __forceinline__ double flicks_to_seconds(int64_t f) {
    return static_cast<double>(f) / 705600000.0;  // util/flicks.h:88-116
}

// adaptive_time_step.h:70-86 for one sub-step: cfl_maximum_time_step
// (cfl.cpp:12-21: factor h / sqrt(max |v|^2) in double, 1 s when unsafe, to
// flicks), then ceil(cfl step / quantum) clamped to [1, divisor] and to the
// quanta left; dt = quanta * quantum.
__device__ __forceinline__ void frame_begin(SceneDev* sd, float vmax2) {
    FrameDev& fr = sd->frame;
    double const numer = static_cast<double>(fr.cfl * sd->c.kc.h);
    double const denom = sqrt(static_cast<double>(vmax2));
    bool safe = false;
    if (denom != 0.0) {
        safe = fabs(denom) >= 1.0 || fabs(denom) / 2.2250738585072014e-308 > fabs(numer);
    }
    int64_t const cfl_step = host_double_to_i64((safe ? numer / denom : 1.0) * 705600000.0);
    int32_t ns = host_double_to_i32(ceil(flicks_to_seconds(cfl_step) / flicks_to_seconds(fr.per)));
    ns = ns < 1 ? 1 : (ns > fr.divisor ? fr.divisor : ns);
    ns = ns < fr.remaining ? ns : fr.remaining;
    int64_t const sub = static_cast<int64_t>(ns) * fr.per;
    sd->c.dt = static_cast<float>(flicks_to_seconds(sub));
    fr.next_remaining = fr.remaining - ns;
    fr.next_time = fr.time + sub;
}

// Start of a sub-step for one scene: the CFL dt and cleared statistics.
__device__ __forceinline__ void step_begin_scene(SceneDev* sd, SceneStats* stats) {
    if (sd->frame.on) {
        frame_begin(sd, __uint_as_float(stats->vmax2_bits));
    } else if (sd->cfl > 0.0f) {
        // cfl.cpp:12-21: factor * h / sqrt(max |v|^2) in double, over the
        // velocities the previous sub-step left (dt_max while at rest).
        float const vmax2 = __uint_as_float(stats->vmax2_bits);
        float dt = sd->dt_max;
        if (vmax2 > 0.0f) {
            double const d = static_cast<double>(sd->cfl) * sd->c.kc.h /
                             sqrt(static_cast<double>(vmax2));
            if (d < static_cast<double>(dt)) { dt = static_cast<float>(d); }
        }
        sd->c.dt = dt;
    }
    // max |v|^2 stays until the end of this sub-step replaces it: a sub-step
    // redone after a short sort recomputes the same dt.
    SceneStats st{};
    st.dt = sd->c.dt;
    st.vmax2_bits = stats->vmax2_bits;
    *stats = st;
}

// Key range of the index from the per-workgroup partials (one workgroup of
// 256 threads), clearing the short-sort flag and the neighbour-build counters
// of this indexing run; with IndexArgs::step_stats also the start of the
// sub-step for every scene (no kernel reads a scene's dt or statistics before
// the sort is done).
__device__ __forceinline__ void fold_key_range(IndexArgs const& a, uint32_t parts) {
    if (a.step_stats) {
        for (uint32_t s = threadIdx.x; s < a.step_nscenes; s += blockDim.x) {
            step_begin_scene(a.step_scenes + s, a.step_stats + s);
        }
    }
    unsigned long long key_min = ~0ull;
    unsigned long long key_max = 0ull;
    for (uint32_t k = threadIdx.x; k < parts; k += blockDim.x) {
        unsigned long long const lo = a.key_partials[2 * k];
        unsigned long long const hi = a.key_partials[2 * k + 1];
        key_min = lo < key_min ? lo : key_min;
        key_max = hi > key_max ? hi : key_max;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long const lo = __shfl_xor(key_min, off);
        unsigned long long const hi = __shfl_xor(key_max, off);
        key_min = lo < key_min ? lo : key_min;
        key_max = hi > key_max ? hi : key_max;
    }
    __shared__ unsigned long long smin[4], smax[4];
    if (__lane_id() == 0) {
        smin[threadIdx.x / 64] = key_min;
        smax[threadIdx.x / 64] = key_max;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            key_min = smin[w] < key_min ? smin[w] : key_min;
            key_max = smax[w] > key_max ? smax[w] : key_max;
        }
        a.keystate->kmin = key_min;
        a.keystate->kmax = key_max;
        a.keystate->frame_remaining =
          (a.step_stats && a.step_scenes[0].frame.on) ? a.step_scenes[0].frame.next_remaining : -1;
        // A dense index whose tables do not hold every key is short: redone.
        a.keystate->sort_short = (a.dense_bins && a.keystate->dense_out) ? 1u : 0u;
        a.keystate->dense_out = 0u;
        a.keystate->dense_crowded = 0u;
        a.keystate->overflow_count = 0u;
        a.keystate->max_count = 0u;
        a.keystate->long_lists = 0u;
        a.keystate->skin_flags = 0u;
        a.keystate->skin_disp = 0u;
    }
}

// neighborhood.cpp:10-16 (index pairs), with the key range reduced on the fly.
// Each workgroup stores its points' key range in its own partial slot (no
// atomics: one address taken by every wave serialises at the device-coherent
// level); k_key_range folds the partials.
__global__ __launch_bounds__(256) void k_keys(IndexArgs a) {
    unsigned long long key_min = ~0ull;
    unsigned long long key_max = 0ull;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += gridDim.x * blockDim.x) {
        float const cell = a.scenes[a.scene_of[i]].c.cell;
        int2 const g = grid_coord(a.pos[i], cell);
        uint64_t const key = morton2(g.x, g.y);
        a.keys[i] = key;
        if (!a.dense_bins) { a.vals[i] = i; }  // the dense scatter writes the indices itself
        if (a.grid) { a.grid[i] = g; }
        if (a.dense_bins) {
            // Dense index: count the point into its bin (keys outside their
            // scene's table go to the extra bin dense_bins, and the run is
            // short); the count before it is its rank. Runs of one bin across
            // adjacent lanes (neighbouring points often share a cell) take one
            // atomic for the whole run.
            uint64_t const bin = dense_bin(a.dscene[a.scene_of[i]], key, a.dense_bins);
            if (__ballot(bin == a.dense_bins) && __lane_id() == 0) {
                atomicOr(&a.keystate->dense_out, 1u);
            }
            uint32_t const lane = __lane_id();
            uint64_t const prev = __shfl_up(bin, 1);
            unsigned long long const heads = __ballot(lane == 0 || prev != bin);
            unsigned long long const active = __ballot(1);
            unsigned long long const le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
            uint32_t const head = 63u - static_cast<uint32_t>(__clzll(heads & le));
            unsigned long long const after = heads & ~le;
            uint32_t const end = after ? static_cast<uint32_t>(__ffsll(after) - 1) : 64u;
            uint32_t const last = 63u - static_cast<uint32_t>(__clzll(active));
            uint32_t base = 0;
            if (lane == head) {
                uint32_t const run = (end <= last ? end : last + 1u) - head;
                base = atomicAdd(a.dcount + bin, run);
            }
            base = __shfl(base, static_cast<int>(head));
            a.drank[i] = base + (lane - head);
        }
        key_min = key < key_min ? key : key_min;
        key_max = key > key_max ? key : key_max;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long const lo = __shfl_xor(key_min, off);
        unsigned long long const hi = __shfl_xor(key_max, off);
        key_min = lo < key_min ? lo : key_min;
        key_max = hi > key_max ? hi : key_max;
    }
    __shared__ unsigned long long smin[4], smax[4];
    if (__lane_id() == 0) {
        smin[threadIdx.x / 64] = key_min;
        smax[threadIdx.x / 64] = key_max;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            key_min = smin[w] < key_min ? smin[w] : key_min;
            key_max = smax[w] > key_max ? smax[w] : key_max;
        }
        a.key_partials[2 * blockIdx.x] = key_min;
        a.key_partials[2 * blockIdx.x + 1] = key_max;
    }
}

// The key range fold as its own one-workgroup launch. (Folding in the last
// workgroup of k_keys to finish needs a device-scope release per workgroup,
// which writes back the XCD's L2 each time: 20 -> 65 us at 1M points.)
__global__ __launch_bounds__(256) void k_key_range(IndexArgs a, uint32_t parts) {
    fold_key_range(a, parts);
}

// Workgroups of k_keys over n points, each leaving one key-range partial.
static uint32_t key_parts(uint32_t n) {
    uint32_t blocks = (n + 255) / 256;
    if (blocks > KEY_PARTS) { blocks = KEY_PARTS; }
    return blocks ? blocks : 1u;
}

// Dense index scatter: point i goes to the first slot of its bin plus its
// rank (the ranks of one bin are a permutation of 0 .. count - 1, so every
// slot receives exactly one point). Workgroup 0 first folds the key range
// and starts the sub-step (fold_key_range, for k_keys' `parts` partials):
// the dense path has no k_key_range launch, and nothing before this kernel
// reads what the fold writes.
__global__ __launch_bounds__(256) void k_dense_scatter(IndexArgs a, uint32_t parts) {
    if (blockIdx.x == 0) { fold_key_range(a, parts); }
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) { return; }
    uint64_t const key = a.keys[i];
    uint64_t const bin = dense_bin(a.dscene[a.scene_of[i]], key, a.dense_bins);
    uint32_t const t = a.doffset[bin] + a.drank[i];
    const_cast<uint64_t*>(a.sorted_keys)[t] = key;
    const_cast<uint32_t*>(a.sorted_vals)[t] = i;
}

void launch_dense_scatter(IndexArgs const& a, hipStream_t s) {
    uint32_t const blocks = a.n ? (a.n + 255) / 256 : 1u;
    hipLaunchKernelGGL(k_dense_scatter, dim3(blocks), dim3(256), 0, s, a, key_parts(a.n));
}

// Keys and their range; with the dense index (a.dense_bins) the range is
// folded by the dense scatter's first workgroup instead of a launch of its own.
void launch_keys(IndexArgs const& a, hipStream_t s) {
    uint32_t const blocks = key_parts(a.n);
    hipLaunchKernelGGL(k_keys, dim3(blocks), dim3(256), 0, s, a);
    if (!a.dense_bins) { hipLaunchKernelGGL(k_key_range, dim3(1), dim3(256), 0, s, a, blocks); }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        p[i] = v;
    }
}

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (!n) { return; }
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) { blocks = 4096; }
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, v, n);
}

//------------------------------------------------------------------------------
// Cells: neighborhood.cpp:33-124 (block indices, block ranges, key map). A
// block starts where the (scene, key) pair changes.
__global__ __launch_bounds__(256) void k_cell_heads(IndexArgs a) {
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0 && a.check_keys) {
        // The LSD sort ordered only the low 8*passes key bits; every key must
        // share the bits above them.
        unsigned long long const diff = a.keystate->kmin ^ a.keystate->kmax;
        bool const covered = (a.key_passes >= 8) || ((diff >> (8 * a.key_passes)) == 0ull);
        if (!covered) { a.keystate->sort_short = 1u; }
    }
    if (t >= a.n) { return; }
    bool head = (t == 0);
    if (!head) {
        head = a.sorted_keys[t] != a.sorted_keys[t - 1] ||
               a.scene_of[a.sorted_vals[t]] != a.scene_of[a.sorted_vals[t - 1]];
    }
    a.heads[t] = head ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_cells_finalize(IndexArgs a) {
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) { return; }
    uint64_t const key = a.sorted_keys[t];
    uint32_t const o = a.sorted_vals[t];
    uint32_t const tag = a.scene_of[o];
    uint32_t const head = a.heads[t];
    uint32_t const b = a.head_scan[t] + head - 1u;
    a.block_index[t] = b;
    if (head) {
        a.cell_begin[b] = t;
        a.cell_key[b] = key;
        a.cell_tag[b] = tag;
    }
    bool const tail = (t + 1 == a.n) || a.sorted_keys[t + 1] != key ||
                      a.scene_of[a.sorted_vals[t + 1]] != tag;
    if (tail) { a.cell_end[b] = t + 1; }
    if (a.scene_kmax && (t + 1 == a.n || a.scene_of[a.sorted_vals[t + 1]] != tag)) {
        a.scene_kmax[tag] = key;  // sorted by (scene, key): the scene's largest
    }
    a.orig[t] = o;
    a.rank[o] = t;
    if (a.pos_out) { a.pos_out[t] = a.pos[o]; }
    if (a.vel_out) { a.vel_out[t] = a.vel_in[o]; }
    if (a.kv_out) { a.kv_out[t] = a.kv_in[o]; }
    if (a.kd_out) { a.kd_out[t] = a.kd_in[o]; }
}

// Dense index: the scatter left each cell's points in the order
// of the counting atomics; point t of a cell goes to the cell's first slot
// plus the number of its points with a smaller original index, which is the
// order of the stable sort by key, and the cell's data and the gathers are
// written there. A cell above DENSE_CELL_MAX points flags the indexing run
// as short (redone by the sort) and keeps the scatter's order, still one
// point per slot.
__global__ __launch_bounds__(256) void k_cells_finalize_dense(IndexArgs a) {
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) { return; }
    uint64_t const key = a.sorted_keys[t];
    uint32_t const* __restrict__ v = a.sorted_vals;
    uint32_t const o = v[t];
    uint32_t const tag = a.scene_of[o];
    uint64_t const bin = dense_bin(a.dscene[tag], key, a.dense_bins);
    uint32_t const b = a.doffset[bin];
    uint32_t const e = a.doffset[bin + 1];
    uint32_t d = t;
    if (e - b <= DENSE_CELL_MAX) {
        uint32_t r = 0;
        for (uint32_t k = b; k < e; ++k) { r += v[k] < o ? 1u : 0u; }
        d = b + r;
    } else if (t == b) {
        atomicOr(&a.keystate->dense_crowded, 1u);
        atomicOr(&a.keystate->sort_short, 1u);
    }
    uint32_t const cb = a.dcell[bin];  // the cell scan's ordinal of the bin
    a.block_index[d] = cb;
    if (t == b) {
        a.cell_begin[cb] = b;
        a.cell_end[cb] = e;
        a.cell_key[cb] = key;
        a.cell_tag[cb] = tag;
    }
    // The scene's last cell holds its largest key (cells never span scenes
    // inside the tables; the extra bin's cell only exists in a short run).
    if (a.scene_kmax && t + 1 == e && (e == a.n || a.scene_of[v[e]] != tag)) {
        a.scene_kmax[tag] = key;
    }
    a.orig[d] = o;
    a.rank[o] = d;
    if (a.pos_out) { a.pos_out[d] = a.pos[o]; }
    if (a.vel_out) { a.vel_out[d] = a.vel_in[o]; }
    if (a.kv_out) { a.kv_out[d] = a.kv_in[o]; }
    if (a.kd_out) { a.kd_out[d] = a.kd_in[o]; }
}

void launch_cell_heads(IndexArgs const& a, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_heads, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}
void launch_cells_finalize(IndexArgs const& a, hipStream_t s) {
    if (a.dense_bins) {
        hipLaunchKernelGGL(k_cells_finalize_dense, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
        return;
    }
    hipLaunchKernelGGL(k_cells_finalize, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}

//------------------------------------------------------------------------------
// Per-sub-step and per-loop bookkeeping.
__global__ __launch_bounds__(256) void k_step_begin(Batch b) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < b.nscenes;
         s += gridDim.x * blockDim.x) {
        step_begin_scene(const_cast<SceneDev*>(b.scenes) + s, b.stats + s);
    }
}

void launch_step_begin(Batch const& b, hipStream_t s) {
    uint32_t blocks = (b.nscenes + 255) / 256;
    hipLaunchKernelGGL(k_step_begin, dim3(blocks ? blocks : 1), dim3(256), 0, s, b);
}

// Every scene with particles enters the solver loop (one workgroup), except
// a pool scene already removed (its results are discarded at the commit).
template <int THREADS>
__device__ __forceinline__ void loop_begin_block(Batch const& b) {
    uint32_t cnt = 0;
    for (uint32_t s = threadIdx.x; s < b.nscenes; s += blockDim.x) {
        unsigned int const act = (b.scenes[s].n > 0 && !b.scenes[s].retired) ? 1u : 0u;
        b.stats[s].active = act;
        cnt += act;
    }
    uint32_t const words = (b.nscenes < ACTIVE_BIT_SCENES ? b.nscenes : ACTIVE_BIT_SCENES) + 31u;
    for (uint32_t w = threadIdx.x; w < words / 32u; w += blockDim.x) {
        unsigned int bits = 0u;
        for (uint32_t i = 0; i < 32u && 32u * w + i < b.nscenes; ++i) {
            SceneDev const& sc = b.scenes[32u * w + i];
            bits |= (sc.n > 0 && !sc.retired) ? (1u << i) : 0u;
        }
        b.keystate->active_bits[w] = bits;
    }
    for (int off = 32; off > 0; off >>= 1) { cnt += __shfl_xor(cnt, off); }
    __shared__ uint32_t ws[THREADS / 64];
    if (__lane_id() == 0) { ws[threadIdx.x / 64] = cnt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < THREADS / 64; ++w) { tot += ws[w]; }
        b.keystate->active_count = tot;
    }
}

__global__ __launch_bounds__(1024) void k_loop_begin(Batch b) { loop_begin_block<1024>(b); }

void launch_loop_begin(Batch const& b, hipStream_t s) {
    hipLaunchKernelGGL(k_loop_begin, dim3(1), dim3(1024), 0, s, b);
}

// Exit tests of the reference's solver loops, evaluated per scene on the
// device: IISPH iisph.cpp:126-134 (from the third iteration), DFSPH divergence
// dfsph_p.cpp:186-210 (from the second), DFSPH density dfsph_p.cpp:112-132
// (from the third), PCISPH pcisph.cpp:144-148 (from the third). The average is
// the reference's (double) isum / (1000 n) narrowed to float.
__device__ __forceinline__ void loop_decide_scene(Batch const& b, int kind, int it, int slot,
                                                  uint32_t s) {
    SceneStats& st = b.stats[s];
    if (!st.active) { return; }
    SceneDev const& sc = b.scenes[s];
    Reduction const r = b.red[s * MAX_RED_SLOTS + slot];
    double const num = static_cast<double>(r.isum);
    double const den = static_cast<double>(1000ull * sc.n);
    float const avg = static_cast<float>(num / den);
    float const mx = __uint_as_float(r.max_bits);
    bool exit = false;
    int max_iters = 0;
    switch (kind) {
    case LOOP_IISPH:
        max_iters = sc.thr.iisph_max_iters;
        exit = it > 1 && avg <= sc.thr.iisph_avg && mx <= sc.thr.iisph_max;
        st.error_average = avg;
        st.error_max = mx;
        break;
    case LOOP_DIVERGENCE:
        max_iters = sc.thr.dfsph_max_iters;
        if (it > 0) {
            exit = avg <= sc.thr.div_avg && mx <= sc.thr.div_max;
            st.divergence_error_average = avg;
            st.divergence_error_max = mx;
        }
        break;
    case LOOP_DENSITY:
        max_iters = sc.thr.dfsph_max_iters;
        if (it > 1) {
            exit = avg <= sc.thr.dens_avg && mx <= sc.thr.dens_max;
            st.error_average = avg;
            st.error_max = mx;
        }
        break;
    default: {  // LOOP_PCISPH
        max_iters = 6;
        if (it > 1) {
            float const err = mx / sc.c.rho0;
            exit = err <= 0.01f;
            st.error_average = 0.0f;
            st.error_max = err;
        }
        break;
    }
    }
    if (exit || it + 1 >= max_iters) {
        st.active = 0u;
        if (kind == LOOP_DIVERGENCE) {
            st.divergence_iterations = it + 1;
        } else {
            st.iterations = it + 1;
        }
        atomicSub(&b.keystate->active_count, 1u);
        if (s < ACTIVE_BIT_SCENES) { atomicAnd(&b.keystate->active_bits[s >> 5], ~(1u << (s & 31u))); }
    }
}

// One workgroup per active scene: the sweep's tile partials of the scene
// (integer sum and max, so the order does not matter) into the scene's slot,
// then the exit test.
constexpr int DECIDE_THREADS = 1024;

__global__ __launch_bounds__(DECIDE_THREADS) void k_loop_decide(Batch b, int kind, int it,
                                                                 int slot) {
    for (uint32_t s = blockIdx.x; s < b.nscenes; s += gridDim.x) {
        if (!b.stats[s].active) { continue; }  // the same for the whole workgroup
        SceneDev const& sd = b.scenes[s];
        unsigned long long v = 0;
        unsigned int m = 0;
        for (uint32_t k = threadIdx.x; k < sd.ntiles; k += blockDim.x) {
            Reduction const r = b.tile_red[sd.first_tile + k];
            v += r.isum;
            m = r.max_bits > m ? r.max_bits : m;
        }
        __syncthreads();  // the previous scene's use of the block's LDS is over
        block_residual<DECIDE_THREADS>(v, m, &b.red[s * MAX_RED_SLOTS + slot]);
        if (threadIdx.x == 0) { loop_decide_scene(b, kind, it, slot, s); }
    }
}

void launch_loop_decide(Batch const& b, int kind, int it, int slot, hipStream_t s) {
    uint32_t const blocks = b.nscenes < 4096 ? b.nscenes : 4096;
    hipLaunchKernelGGL(k_loop_decide, dim3(blocks ? blocks : 1), dim3(DECIDE_THREADS), 0, s, b,
                       kind, it, slot);
}

//------------------------------------------------------------------------------
// Cell windows. One thread per (cell, neighbour cell): the cell's key gives
// back its grid coordinate (the Morton code inverts exactly), and each of the
// 3x3 neighbour cells is looked up once per cell instead of once per particle.
// Cells are sorted by (scene, key), so a lookup among the same cells is a
// search of the cell list, galloping from a hint: the cell's own ordinal moved
// by the key difference, which is exact where every cell in between is
// occupied (the interior of a fluid) and otherwise a nearby start, so the
// probes stay on a few cache lines. The other (solid) cells are looked up in
// their static hash set.
__device__ __forceinline__ bool cell_before(CellList const& cl, uint32_t i, uint32_t tag,
                                            uint64_t key) {
    uint32_t const ti = cl.tag[i];
    return ti < tag || (ti == tag && cl.key[i] < key);
}

// First cell i in [lo, hi) that is not before (tag, key); hi when none.
__device__ __forceinline__ uint32_t cell_bisect(CellList const& cl, uint32_t lo, uint32_t hi,
                                                uint32_t tag, uint64_t key) {
    while (lo < hi) {
        uint32_t const mid = lo + (hi - lo) / 2;
        if (cell_before(cl, mid, tag, key)) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return lo;
}

// Same over [0, n), probing at hint +- 1, 2, 4, ... first: O(log |i - hint|).
__device__ __forceinline__ uint32_t cell_gallop(CellList const& cl, uint32_t n, uint32_t hint,
                                                uint32_t tag, uint64_t key) {
    uint32_t lo, hi;
    if (cell_before(cl, hint, tag, key)) {
        lo = hint + 1;
        hi = n;
        for (uint32_t step = 1;; step <<= 1) {
            uint32_t const probe = hint + step;
            if (probe >= n) { break; }
            if (!cell_before(cl, probe, tag, key)) {
                hi = probe;
                break;
            }
            lo = probe + 1;
        }
    } else {
        lo = 0;
        hi = hint;
        for (uint32_t step = 1; step <= hint; step <<= 1) {
            uint32_t const probe = hint - step;
            if (cell_before(cl, probe, tag, key)) {
                lo = probe + 1;
                break;
            }
            hi = probe;
        }
    }
    return cell_bisect(cl, lo, hi, tag, key);
}

__device__ __forceinline__ uint2 cell_range(CellList const& cl, uint32_t i, uint32_t n,
                                            uint32_t tag, uint64_t key) {
    if (i < n && cl.tag[i] == tag && cl.key[i] == key) {
        return make_uint2(cl.begin[i], cl.end[i]);
    }
    return make_uint2(0u, 0u);
}

__device__ __forceinline__ uint2 cell_set_range(CellSet const& cs, CellList const& cl,
                                                uint64_t key, uint32_t tag) {
    uint32_t const lo = static_cast<uint32_t>(key);
    uint32_t const hi = static_cast<uint32_t>(key >> 32);
    uint32_t s = cell_set_hash(key, tag) & cs.mask;
    for (;;) {
        uint4 const e = cs.slots[s];
        if (e.w == EMPTY_SLOT) { return make_uint2(0u, 0u); }
        if (e.x == lo && e.y == hi && e.z == tag) { return make_uint2(cl.begin[e.w], cl.end[e.w]); }
        s = (s + 1) & cs.mask;
    }
}

__global__ __launch_bounds__(256) void k_cell_windows(WindowArgs w) {
    uint32_t const n = *w.self.count;
    size_t const total = static_cast<size_t>(n) * 9;
    size_t const first = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
    if (first == 0 && w.report) { *w.report = n; }
    for (size_t item = first; item < total; item += size_t(gridDim.x) * blockDim.x) {
        uint32_t const c = static_cast<uint32_t>(item / 9);
        uint32_t const k = static_cast<uint32_t>(item - static_cast<size_t>(c) * 9);
        uint64_t const key = w.self.key[c];
        uint32_t const tag = w.self.tag[c];
        // neighborhood.h:233-243: j outer, i inner, each from g - 1 to g + 1
        // (int32 arithmetic; the key takes the coordinates as uint32).
        uint32_t const gx = compact_bits_2(key) + (k % 3u) - 1u;
        uint32_t const gy = compact_bits_2(key >> 1) + (k / 3u) - 1u;
        uint64_t const nk = morton2(static_cast<int32_t>(gx), static_cast<int32_t>(gy));
        // Hint: ordinals advance at most as fast as keys.
        uint32_t hint = c;
        if (nk > key) {
            uint64_t const d = nk - key;
            hint = d < static_cast<uint64_t>(n - 1u - c) ? c + static_cast<uint32_t>(d) : n - 1u;
        } else {
            uint64_t const d = key - nk;
            hint = d < static_cast<uint64_t>(c) ? c - static_cast<uint32_t>(d) : 0u;
        }
        uint32_t const i = (w.self.key[hint] == nk && w.self.tag[hint] == tag)
                             ? hint
                             : cell_gallop(w.self, n, hint, tag, nk);
        uint2 r = cell_range(w.self, i, n, tag, nk);
        if (w.owin) {
            uint2 const o = w.other_set.slots ? cell_set_range(w.other_set, w.other, nk, tag)
                                              : make_uint2(0u, 0u);
            w.owin[item] = o;
            if (o.y > o.x) { r.y |= WIN_SOLID_NEAR; }
        }
        w.win[item] = r;
    }
}

// Windows over the dense index, one thread per cell: a neighbour cell's
// range is read from the bin offsets of the cell's scene table; the solid
// cells are looked up only where the solid bitmap (over the same bins) has
// the bin, or the key is outside the table, and a cell's solid windows are
// written only when some neighbour cell holds solids (the build reads them
// only then: WIN_SOLID_NEAR).
__global__ __launch_bounds__(256) void k_cell_windows_dense(WindowArgs w) {
    // Grid-stride over blocks of 256 cells; each block's 256 x 9 ranges are
    // staged in LDS and written out as one contiguous run.
    __shared__ uint2 sw[256 * 9];
    uint32_t const n = *w.self.count;
    if (blockIdx.x == 0 && threadIdx.x == 0 && w.report) { *w.report = n; }
    for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
        uint32_t const c = base + threadIdx.x;
        uint2 r[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) { r[k] = make_uint2(0u, 0u); }
        if (c < n) {
            uint64_t const key = w.self.key[c];
            uint32_t const tag = w.self.tag[c];
            DenseScene const ds = w.dscene[tag];
            uint32_t const gx0 = compact_bits_2(key);
            uint32_t const gy0 = compact_bits_2(key >> 1);
            uint64_t nks[9];
            uint64_t bins[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                // neighborhood.h:233-243: j outer, i inner, each from g - 1 to g + 1.
                uint32_t const gx = gx0 + static_cast<uint32_t>(k % 3) - 1u;
                uint32_t const gy = gy0 + static_cast<uint32_t>(k / 3) - 1u;
                uint64_t const nk = morton2(static_cast<int32_t>(gx), static_cast<int32_t>(gy));
                uint64_t const bin = dense_bin(ds, nk, w.dense_bins);
                nks[k] = nk;
                bins[k] = bin;
                if (bin < w.dense_bins) {
                    uint32_t const b = w.doffset[bin];
                    uint32_t const e = w.doffset[bin + 1];
                    if (e > b) { r[k] = make_uint2(b, e); }
                }
            }
            if (w.owin) {
                uint2 o[9];
                bool any = false;
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    uint64_t const nk = nks[k];
                    uint64_t const bin = bins[k];
                    bool const maybe = w.other_set.slots &&
                                       (bin >= w.dense_bins || !w.other_bits ||
                                        ((w.other_bits[bin >> 5] >> (bin & 31u)) & 1u));
                    o[k] = maybe ? cell_set_range(w.other_set, w.other, nk, tag) : make_uint2(0u, 0u);
                    if (o[k].y > o[k].x) {
                        r[k].y |= WIN_SOLID_NEAR;
                        any = true;
                    }
                }
                if (any) {
#pragma unroll
                    for (int k = 0; k < 9; ++k) { w.owin[static_cast<size_t>(c) * 9 + k] = o[k]; }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) { sw[threadIdx.x * 9 + k] = r[k]; }
        __syncthreads();
        uint32_t const cells = (n - base) < 256u ? (n - base) : 256u;
        uint2* const out = w.win + static_cast<size_t>(base) * 9;
        for (uint32_t i = threadIdx.x; i < cells * 9u; i += 256u) { out[i] = sw[i]; }
        __syncthreads();
    }
}

void launch_cell_windows(WindowArgs const& w, hipStream_t s) {
    if (!w.max_cells) { return; }
    if (w.dense_bins) {
        uint32_t blocks = (w.max_cells + 255) / 256;
        if (blocks > 1024u) { blocks = 1024u; }
        hipLaunchKernelGGL(k_cell_windows_dense, dim3(blocks), dim3(256), 0, s, w);
        return;
    }
    size_t blocks = (static_cast<size_t>(w.max_cells) * 9 + 255) / 256;
    if (blocks > 4096) { blocks = 4096; }
    hipLaunchKernelGGL(k_cell_windows, dim3(blocks), dim3(256), 0, s, w);
}

//------------------------------------------------------------------------------
// Neighbour build: sph_common/neighborhood.h:198-282.
//
// Candidates are scanned in the reference's order (cells j-outer, i-inner; in a
// cell, sorted slots, i.e. increasing original index) and accepted while
// distance < cut-off, up to 64 in scan order. Each thread keeps its accepted
// list in LDS as (distance bits << 32 | slot), sorted by insertion in the
// reference's neighbor_order (neighborhood.h:172-186): by distance, whose
// non-negative bit patterns order like the floats, then by original index.

// True when entry a sorts after entry b.
__device__ __forceinline__ bool entry_after(uint64_t a, uint64_t b,
                                            uint32_t const* __restrict__ oorig) {
    uint32_t const da = static_cast<uint32_t>(a >> 32);
    uint32_t const db = static_cast<uint32_t>(b >> 32);
    if (da != db) { return da > db; }
    return oorig[static_cast<uint32_t>(a)] > oorig[static_cast<uint32_t>(b)];
}

template <int N>
__device__ constexpr auto const& net_of() {
    if constexpr (N == 8) {
        return NET8;
    } else if constexpr (N == 12) {
        return NET12;
    } else if constexpr (N == 16) {
        return NET16;
    } else if constexpr (N == 32) {
        return NET32;
    } else {
        return NET64;
    }
}

// Largest value of v over the lanes of the wave.
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        int const o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}

// Sorts one lane's cnt <= N collected entries into the reference order (by
// distance, then by original index) and leaves them as (distance bits << 32 |
// slot). They arrive as (|v|^2 bits << 32 | slot), v = candidate - particle,
// and the network orders them by (|v|^2, slot). The distances are then taken
// as imath_length takes them, sqrt(|v|^2), which is monotone in |v|^2, so the
// order is by distance except among equal distances (rare), which an
// insertion pass with the exact comparator puts in original index order. A
// |v|^2 below 2 FLT_MIN, where imath_length scales instead, sends the list
// to that pass with every distance taken again from the positions.
//
// The entries are compared as doubles: an entry's sign bit is clear and its
// top 11 bits (the double's exponent field) come from a finite float's
// exponent, so it is never a NaN or infinity, and non-negative doubles order
// exactly like their bit patterns. A compare-exchange is then one v_min_f64
// and one v_max_f64 (the kernels keep f64 denormals, so an entry with a tiny
// |v|^2 keeps its bits). Unused inputs hold the largest finite double, which
// sorts after every entry.
template <int N, int W>
__device__ __forceinline__ void net_sort(uint64_t (*ent)[W], uint32_t lane, int cnt,
                                         uint32_t const* __restrict__ oorig,
                                         V2 const* __restrict__ opos, V2 p) {
    constexpr auto const& net = net_of<N>();
    static_assert(net.inputs == N, "network size");
    constexpr uint64_t PAD = 0x7FEFFFFFFFFFFFFFull;
    double d64[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        uint64_t const v = ent[k][lane];
        d64[k] = __builtin_bit_cast(double, (k < cnt) ? v : PAD);
    }
#pragma unroll
    for (int c = 0; c < net.count; ++c) {
        double const x = d64[net.ce[c].a];
        double const y = d64[net.ce[c].b];
        d64[net.ce[c].a] = __builtin_fmin(x, y);
        d64[net.ce[c].b] = __builtin_fmax(x, y);
    }
    uint64_t e[N];
#pragma unroll
    for (int k = 0; k < N; ++k) { e[k] = __builtin_bit_cast(uint64_t, d64[k]); }
    bool tiny = false;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (k < cnt) {
            float const l2 = __uint_as_float(static_cast<uint32_t>(e[k] >> 32));
            tiny = tiny || !(l2 >= 2.0f * FLT_MIN);
            float const d = __builtin_sqrtf(l2);
            e[k] = (static_cast<uint64_t>(__float_as_uint(d)) << 32) | static_cast<uint32_t>(e[k]);
        }
    }
    bool tie = false;
#pragma unroll
    for (int k = 1; k < N; ++k) {
        tie = tie || (k < cnt && (e[k - 1] >> 32) == (e[k] >> 32));
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (k < cnt) { ent[k][lane] = e[k]; }
    }
    if (tiny) {
        for (int k = 0; k < cnt; ++k) {
            uint32_t const s = static_cast<uint32_t>(ent[k][lane]);
            float const d = imath_length(sub(opos[s], p));
            ent[k][lane] = (static_cast<uint64_t>(__float_as_uint(d)) << 32) | s;
        }
    }
    if (tie || tiny) {
        for (int k = 1; k < cnt; ++k) {
            uint64_t const key = ent[k][lane];
            int kk = k;
            while (kk > 0) {
                uint64_t const prev = ent[kk - 1][lane];
                if (!entry_after(prev, key, oorig)) { break; }
                ent[kk][lane] = prev;
                --kk;
            }
            ent[kk][lane] = key;
        }
    }
}

// Positions of candidates s .. s + C - 1 (s inside the range), two per 16-byte
// load; entries past the range's end are loaded but not used (position
// buffers carry tail padding for the last range).
struct alignas(8) V2Pair {
    V2 a;
    V2 b;
};
template <int C>
__device__ __forceinline__ void load_chunk(V2 (&q)[C], V2 const* __restrict__ opos, uint32_t s) {
    static_assert(C % 2 == 0, "chunks are loaded in pairs");
#pragma unroll
    for (int u = 0; u < C; u += 2) {
        V2Pair v;
        __builtin_memcpy(&v, opos + s + u, sizeof(V2Pair));
        q[u] = v.a;
        q[u + 1] = v.b;
    }
}

// Acceptance of one loaded chunk, in scan order: appended unsorted (sorted
// afterwards by net_sort). Returns false when the scan stops: a capped list
// overflowed (cnt = -1) or a full-capacity list is full. CUT2: acceptance as
// |v|^2 < cut2 (the scene's cut2 > 0, always for a normal cut-off), else as
// Imath length < maxd; the choice is made once per particle, not per
// candidate.
//
template <int CAP, int C, int W, bool CUT2>
__device__ __forceinline__ bool take_chunk(uint64_t (*ent)[W], uint32_t lane, V2 p,
                                           uint32_t self, bool same, V2 const (&q)[C],
                                           uint32_t s0, uint32_t e, float maxd, float cut2,
                                           int& cnt) {
    bool in[C];
    float l2[C];
#pragma unroll
    for (int u = 0; u < C; ++u) {
        V2 const v = sub(q[u], p);
        l2[u] = v.x * v.x + v.y * v.y;  // as in imath_length
        if constexpr (CUT2) {
            in[u] = l2[u] < cut2;
        } else {
            in[u] = imath_length(v) < maxd;
        }
    }
    // Unsorted append of (|v|^2 bits << 32 | slot); the distances are taken
    // after the sort (net_sort), once per accepted entry. A capped list
    // (CAP < 64) that overflows is deferred (cnt = -1); a full 64-entry list
    // keeps the first 64 accepted in scan order (neighborhood.h:233-250) and
    // ends the scan.
#pragma unroll
    for (int u = 0; u < C; ++u) {
        uint32_t const s = s0 + u;
        bool const acc = s < e && !(same && s == self) && in[u];
        uint64_t const entry = (static_cast<uint64_t>(__float_as_uint(l2[u])) << 32) | s;
        if (acc && cnt < CAP) { ent[cnt][lane] = entry; }
        if constexpr (CAP < MAXN) {
            cnt += acc ? 1 : 0;
        } else {
            cnt += (acc && cnt < CAP) ? 1 : 0;
        }
    }
    if (CAP < MAXN && cnt > CAP) {
        cnt = -1;
        return false;
    }
    return !(CAP == MAXN && cnt == CAP);
}

// The scan of one particle's 3x3 cell window (nine slot ranges in scan order)
// in chunks of C candidates. Returns the count (-1: a capped list overflowed)
// and ORs the ranges' end words (solid flags) into flags; a full 64-entry list
// stops early and the remaining ranges' flags are taken from the window.
template <int CAP, int W, bool CUT2>
__device__ __forceinline__ int scan_window(uint64_t (*ent)[W], uint32_t lane, V2 p,
                                           uint32_t self, bool same,
                                           uint2 const* __restrict__ win,
                                           V2 const* __restrict__ opos, float maxd, float cut2,
                                           uint32_t& flags) {
    constexpr int C = 4;
    int cnt = 0;
    uint2 cur = win[0];
    if constexpr (CAP <= 16) {
        // The 16-entry build (the common case): a list either fits or is
        // deferred. The first chunk of each range is loaded while the range
        // before it is scanned (ranges of a few candidates are the rule).
        V2 qn[C];
        load_chunk<C>(qn, opos, cur.x);
        uint2 nxt = win[1];
        for (int k = 0; k < 9; ++k) {
            V2 q[C];
#pragma unroll
            for (int u = 0; u < C; ++u) { q[u] = qn[u]; }
            uint2 const nn = win[k < 7 ? k + 2 : 8];
            if (k < 8) { load_chunk<C>(qn, opos, nxt.x); }
            flags |= cur.y;
            uint32_t const e = cur.y & ~WIN_SOLID_NEAR;
            for (uint32_t s0 = cur.x; s0 < e; s0 += C) {
                if (s0 != cur.x) { load_chunk<C>(q, opos, s0); }
                if (!take_chunk<CAP, C, W, CUT2>(ent, lane, p, self, same, q, s0, e, maxd, cut2,
                                                 cnt)) {
                    return cnt;  // overflow (-1)
                }
            }
            cur = nxt;
            nxt = nn;
        }
        return cnt;
    }
    bool go = true;
    for (int k = 0; k < 9 && go; ++k) {
        uint2 const nxt = win[k < 8 ? k + 1 : 8];  // next range in flight
        flags |= cur.y;
        uint32_t const e = cur.y & ~WIN_SOLID_NEAR;
        for (uint32_t s0 = cur.x; s0 < e; s0 += C) {
            V2 q[C];
            load_chunk<C>(q, opos, s0);
            if (!take_chunk<CAP, C, W, CUT2>(ent, lane, p, self, same, q, s0, e, maxd, cut2,
                                             cnt)) {
                go = false;  // overflow (-1) or a full 64-entry list
                break;
            }
        }
        cur = nxt;
    }
    // A full list stops before the last ranges: their solid flags are taken
    // from the window directly.
    for (int k = 0; k < 9 && !go; ++k) { flags |= win[k].y; }
    return cnt;
}

// Sorts each lane's collected entries (net_sort) with the smallest network that
// covers the wave's longest list (a deferred, overflowed list counts as empty
// here); this also turns a single entry's |v|^2 into its distance.
template <int CAP, int W>
__device__ __forceinline__ void sort_collected(uint64_t (*ent)[W], uint32_t lane, int cnt,
                                               uint32_t const* __restrict__ oorig,
                                               V2 const* __restrict__ opos, V2 p) {
    int const wmax = wave_max(cnt);
    if (wmax > 0) {
        if (wmax <= 8) {
            net_sort<8, W>(ent, lane, cnt, oorig, opos, p);
        } else if (wmax <= 12) {
            net_sort<12, W>(ent, lane, cnt, oorig, opos, p);
        } else if (wmax <= 16) {
            net_sort<16, W>(ent, lane, cnt, oorig, opos, p);
        } else if (CAP <= 32 || wmax <= 32) {
            if constexpr (CAP >= 32) { net_sort<32, W>(ent, lane, cnt, oorig, opos, p); }
        } else {
            if constexpr (CAP >= 64) { net_sort<64, W>(ent, lane, cnt, oorig, opos, p); }
        }
    }
}

// Accepted candidates of one particle over its 3x3 cell window. Returns the
// count, or -1 when CAP < 64 and more than CAP candidates are accepted.
// A candidate is accepted when |v| < maxd; with cut2 > 0 (StepConsts::cut2)
// that is decided as |v|^2 < cut2, the same answer without the square root,
// which is then taken for accepted candidates only.
// Window ranges carry WIN_SOLID_NEAR in their end's top bit when the same
// neighbour cell holds solids (k_cell_windows); *near receives the OR.
template <int CAP, int W>
__device__ __forceinline__ int collect_window(uint64_t (*ent)[W], uint32_t lane,
                                              V2 p, uint32_t self, bool same,
                                              uint2 const* __restrict__ win,
                                              V2 const* __restrict__ opos,
                                              uint32_t const* __restrict__ oorig, float maxd,
                                              float cut2, uint32_t* near = nullptr) {
    uint32_t flags = 0;
    int const cnt = (cut2 > 0.0f)
                      ? scan_window<CAP, W, true>(ent, lane, p, self, same, win, opos, maxd, cut2,
                                                  flags)
                      : scan_window<CAP, W, false>(ent, lane, p, self, same, win, opos, maxd,
                                                   cut2, flags);
    if (near) { *near = flags & WIN_SOLID_NEAR; }
    sort_collected<CAP, W>(ent, lane, cnt, oorig, opos, p);
    return cnt;
}

// Skin reuse (SkinView): the accepted candidates of slot t, appended unsorted
// in candidate order as in take_chunk, then sorted. A candidate is accepted
// by the reference's test: its cell (this sub-step's, in pg) within one of
// the particle's in both axes — the 3x3 window of neighborhood.h:233-243,
// in the wrapping 32-bit cell arithmetic of its keys — and |v|^2 < cut2.
// Returns the count, or -1 when CAP < 64 and the list does not fit; a list
// above 64 entries, which the reference truncates in its scan order, sets
// SKIN_FULL and a short sub-step (redone by the full path).
template <int CAP, int W, int ROWS>
__device__ __forceinline__ int collect_skin(uint64_t (*ent)[W], uint32_t lane, uint32_t t,
                                            uint32_t const* __restrict__ cand, int n,
                                            uint32_t stride, V2 const* __restrict__ cpos,
                                            V2 p, float cell, float cut2, uint32_t self,
                                            uint32_t const* __restrict__ oorig, KeyState* ks) {
    // The 3x3 window test matters only for a pair whose separation in an axis
    // comes within the rounding of floor(x / cell) of a whole cell: cells are
    // assigned from x / cell rounded to float, off by at most 2^-24 |x| / cell
    // each, so points less than cell - 2^-21 (|x| + |y| + 4 cell) apart in both
    // axes (x, y: this particle's coordinates; a candidate lies within 2 cells
    // of them) are in neighbouring or equal cells. Other accepted pairs take
    // both cells as the reference assigns them.
    float const band =
      cell - 0x1p-21f * (__builtin_fabsf(p.x) + __builtin_fabsf(p.y) + 4.0f * cell);
    int cnt = 0;
    // Candidate indices are read two chunks ahead of their gathers, without
    // branches: a row past the count reads the list's last row again (a cache
    // hit, and inside the ROWS rows), and an index past the count is replaced
    // by a valid slot (self) only when it is used, so its load need not have
    // arrived before then. The gathers of a chunk are issued before the next
    // indices, so waiting for them does not wait for those.
    int const last = n > 0 ? (n < ROWS ? n - 1 : ROWS - 1) : 0;
    auto row = [&](int k) -> uint32_t {
        int const r = k < last ? k : last;
        return cand[static_cast<size_t>(r) * stride + t];
    };
    uint32_t ia[4], ib[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        ia[u] = row(u);
        ib[u] = row(4 + u);
    }
    for (int k0 = 0; k0 < n; k0 += 4) {
        uint32_t s[4];
        V2 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            s[u] = (k0 + u < n) ? ia[u] : self;
            q[u] = cpos[s[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ia[u] = ib[u];
            ib[u] = row(k0 + 8 + u);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            V2 const v = sub(q[u], p);
            float const l2 = v.x * v.x + v.y * v.y;  // as in imath_length
            bool acc = k0 + u < n && l2 < cut2;
            if (acc && !(__builtin_fabsf(v.x) < band && __builtin_fabsf(v.y) < band)) {
                int2 const gi = grid_coord(p, cell);
                int2 const gj = grid_coord(q[u], cell);
                uint32_t const dx = static_cast<uint32_t>(gj.x) - static_cast<uint32_t>(gi.x) + 1u;
                uint32_t const dy = static_cast<uint32_t>(gj.y) - static_cast<uint32_t>(gi.y) + 1u;
                acc = dx <= 2u && dy <= 2u;
            }
            uint64_t const entry = (static_cast<uint64_t>(__float_as_uint(l2)) << 32) | s[u];
            if (acc && cnt < CAP) { ent[cnt][lane] = entry; }
            cnt += acc ? 1 : 0;
        }
        if (CAP < MAXN && cnt > CAP) {
            cnt = -1;
            break;
        }
    }
    if (CAP == MAXN && cnt > MAXN) {
        atomicOr(&ks->skin_flags, SKIN_FULL);
        atomicOr(&ks->sort_short, 1u);
        cnt = MAXN;
    }
    sort_collected<CAP, W>(ent, lane, cnt, oorig, cpos, p);
    return cnt;
}

// Lists and the fused per-particle sums of slot t. Returns false (nothing of
// the particle's state changed) when a list would exceed CAP.
template <int MODE, int CAP, int W, bool SKIN>
__device__ __forceinline__ bool build_one(Batch const& a, uint64_t (*ent)[W],
                                          uint32_t lane, uint32_t t, uint32_t scene,
                                          int& max_count) {
    StepConsts const c = a.scenes[scene].c;
    V2 const p = a.pos[t];
    uint32_t const cell = SKIN ? 0u : a.cell_of[t];  // windows of the full path
    float const mf = c.rho0 * c.fluid_volume;  // rho0 * V_j (uniform fluid volume)

    // Fluid neighbours: from the cell window, or from the skin candidates.
    uint32_t solid_near = 0;
    int fc;
    if constexpr (SKIN) {
        fc = collect_skin<CAP, W, SKIN_FLUID>(ent, lane, t, a.skin.cf, a.skin.nf[t], a.stride,
                                              a.pos, p, c.cell, c.cut2, t, a.orig, a.keystate);
        solid_near = a.skin.ns[t] != 0 ? 1u : 0u;
    } else {
        fc = collect_window<CAP, W>(ent, lane, p, t, true, a.win + static_cast<size_t>(cell) * 9,
                                    a.pos, a.orig, c.cell, c.cut2, &solid_near);
    }
    if (fc < 0) {
        max_count = CAP + 1;
        return false;
    }
    float dens = c.fluid_volume * c.rho0 * c.kc.w0;  // density.cpp:57-59
    V2 msum_f{0.0f, 0.0f};
    V2 sum_mg{0.0f, 0.0f};    // DFSPH: sum of m grad W over fluid and solid
    float sum_mg2 = 0.0f;     // DFSPH: sum of |m grad W|^2 over fluid
    // Entries are taken four at a time (their position loads in flight
    // together) and summed in list order.
    for (int k0 = 0; k0 < fc; k0 += 4) {
      uint64_t e4[4];
      V2 q4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k0 + u < fc) { e4[u] = ent[k0 + u][lane]; }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k0 + u < fc) { q4[u] = a.pos[static_cast<uint32_t>(e4[u])]; }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k0 + u >= fc) { break; }
        int const k2 = k0 + u;
        uint64_t const e = e4[u];
        uint32_t const s = static_cast<uint32_t>(e);
        float const d = __uint_as_float(static_cast<uint32_t>(e >> 32));
        V2 const vec = sub(q4[u], p);
        size_t const o = static_cast<size_t>(k2) * a.stride + t;
        a.fl.idx[o] = s;
        if (a.store_geometry) {
            a.fl.dist[o] = d;
            a.fl.vec[o] = vec;
        }
        if (MODE != BUILD_LISTS) {
            float w;
            V2 gw;
            kernel_W_GradW_r(neg(vec), d, c.kc, w, gw);
            a.fl.gw[o] = gw;
            dens += w * mf;                                  // density.cpp:69-74
            if (MODE == BUILD_IISPH) { msum_f = add(msum_f, scale(mf, gw)); }  // iisph_ops.cpp:42-47
            if (MODE == BUILD_DFSPH) {                        // dfsph_p_ops.cpp:47-57
                V2 const mg = scale(mf, gw);
                sum_mg = add(sum_mg, mg);
                sum_mg2 += dot(mg, mg);
            }
        }
      }
    }

    // Solid neighbours (no self exclusion: different arrays). A batch without
    // solid lists (the solid-solid build) stops here.
    if (!a.sl.count) {
        a.fl.count[t] = static_cast<uint8_t>(fc);
        max_count = fc;
        return true;
    }
    SolidView const& sv = a.solids;
    int sc = 0;
    if (SKIN && sv.pos && solid_near) {
        sc = collect_skin<CAP, W, SKIN_SOLID>(ent, lane, t, a.skin.cs, a.skin.ns[t], a.stride,
                                              sv.pos, p, c.cell, c.cut2, 0u, sv.orig,
                                              a.keystate);
        if (sc < 0) {
            max_count = CAP + 1;
            return false;
        }
    } else if (a.swin && sv.pos && solid_near) {
        sc = collect_window<CAP, W>(ent, lane, p, 0xFFFFFFFFu, false,
                                 a.swin + static_cast<size_t>(cell) * 9, sv.pos, sv.orig, c.cell,
                                 c.cut2);
        if (sc < 0) {
            max_count = CAP + 1;
            return false;
        }
    }
    a.fl.count[t] = static_cast<uint8_t>(fc);
    a.sl.count[t] = static_cast<uint8_t>(sc);
    max_count = fc > sc ? fc : sc;
    float dens_s = 0.0f;
    V2 msum_s{0.0f, 0.0f};
    for (int k0 = 0; k0 < sc; k0 += 4) {
      uint64_t e4[4];
      V2 q4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k0 + u < sc) { e4[u] = ent[k0 + u][lane]; }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k0 + u < sc) { q4[u] = sv.pos[static_cast<uint32_t>(e4[u])]; }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k0 + u >= sc) { break; }
        int const k2 = k0 + u;
        uint64_t const e = e4[u];
        uint32_t const b = static_cast<uint32_t>(e);
        float const d = __uint_as_float(static_cast<uint32_t>(e >> 32));
        V2 const vec = sub(q4[u], p);
        size_t const o = static_cast<size_t>(k2) * a.stride + t;
        a.sl.idx[o] = b;
        if (a.store_geometry) {
            a.sl.dist[o] = d;
            a.sl.vec[o] = vec;
        }
        if (MODE != BUILD_LISTS) {
            float w;
            V2 gw;
            kernel_W_GradW_r(neg(vec), d, c.kc, w, gw);
            a.sl.gw[o] = gw;
            float const ms = c.rho0 * sv.vol[b];
            dens_s += w * ms;
            if (MODE == BUILD_IISPH) { msum_s = add(msum_s, scale(ms, gw)); }
            if (MODE == BUILD_DFSPH) { sum_mg = add(sum_mg, scale(ms, gw)); }
        }
      }
    }
    if (MODE == BUILD_LISTS) { return true; }

    float rho = dens;
    if (sc) { rho += dens_s; }  // density.cpp:76-80
    a.rho[t] = rho;

    if (MODE == BUILD_IISPH) {
        // iisph_ops.cpp:49-57 d_ii, fluid pass then solid pass.
        float const neg_dt2 = -(c.dt * c.dt);
        float const denom = rho * rho;
        V2 dii{0.0f, 0.0f};
        if (fc) {
            V2 const numer = scale(neg_dt2, msum_f);
            dii = is_safe_div(numer, denom) ? divs(numer, denom) : V2{0.0f, 0.0f};
        }
        if (sc) {
            V2 const numer = scale(neg_dt2, msum_s);
            V2 const ds = is_safe_div(numer, denom) ? divs(numer, denom) : V2{0.0f, 0.0f};
            dii = add(dii, ds);
        }
        a.dii[t] = dii;
    }
    if ((MODE == BUILD_IISPH || MODE == BUILD_AP) && !a.fext) {
        // forces.cpp:11-83 (zeroed force, y -= m g) and dynamics.cpp:24-37;
        // other force models integrate in k_external_forces after the build.
        V2 const f{0.0f, 0.0f - c.mass * c.gravity};
        float const vden = c.rho0 * c.fluid_volume;
        V2 const numer = scale(c.dt, f);
        if (is_safe_div(numer, vden)) { a.vel[t] = add(a.vel[t], divs(numer, vden)); }
    }
    if (MODE == BUILD_DFSPH) {
        // dfsph_p_ops.cpp:78-82
        float alpha = 0.0f;
        if (fc + sc) {
            float const numer = rho * rho;
            float const denom = dot(sum_mg, sum_mg) + sum_mg2;
            alpha = is_safe_div(numer, denom) ? numer / denom : 0.0f;
        }
        a.alpha[t] = alpha;
    }
    return true;
}

// Wave maximum of the list counts into KeyState::max_count (read first: the
// value rarely grows, so few waves take the atomic), and the wave's count of
// lists above 16 entries into KeyState::long_lists (overflowed lists count:
// m is then CAP + 1).
__device__ __forceinline__ void note_max_count(Batch const& a, int m) {
    unsigned long long const longs = __ballot(m > 16);
    if (longs && __lane_id() == 0) {
        atomicAdd(&a.keystate->long_lists, static_cast<unsigned int>(__popcll(longs)));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        int const o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if (__lane_id() == 0) {
        unsigned int const mu = static_cast<unsigned int>(m);
        if (mu > __atomic_load_n(&a.keystate->max_count, __ATOMIC_RELAXED)) {
            atomicMax(&a.keystate->max_count, mu);
        }
    }
}

template <int MODE, int CAP, int WG, bool SKIN>
__global__ __launch_bounds__(WG) void k_build(Batch a) {
    __shared__ uint64_t ent[CAP][WG];
    uint32_t const lane = threadIdx.x;
    // Build workgroups cover TILE / WG parts of a fluid tile each.
    constexpr uint32_t PARTS = TILE / WG;
    uint32_t const k = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    Tile const tl = a.tiles[k / PARTS];
    uint32_t const part_off = (k % PARTS) * WG;
    if (part_off >= tl.count) { return; }  // the whole workgroup
    uint32_t const sub_off = part_off + lane;
    // Every lane stays for the cross-lane reductions below.
    bool const valid = sub_off < tl.count;
    uint32_t const t = tl.start + sub_off;
    int m = 0;
    bool const ok = valid ? build_one<MODE, CAP, WG, SKIN>(a, ent, lane, t, tl.scene, m) : true;
    note_max_count(a, m);
    if (CAP < MAXN) {
        // Deferred particles, appended with one atomic per wave.
        unsigned long long const mask = __ballot(!ok);
        if (mask) {
            int const leader = __ffsll(mask) - 1;
            unsigned int base = 0;
            uint32_t const wl = __lane_id();
            if (static_cast<int>(wl) == leader) {
                base = atomicAdd(&a.keystate->overflow_count,
                                 static_cast<unsigned int>(__popcll(mask)));
            }
            base = __shfl(base, leader);
            if (!ok) {
                unsigned int const r =
                  static_cast<unsigned int>(__popcll(mask & ((1ull << wl) - 1ull)));
                a.overflow[base + r] = make_uint2(t, tl.scene);
            }
        }
    }
}

// Full-capacity pass over the particles the capped build deferred.
template <int MODE, bool SKIN>
__global__ __launch_bounds__(BUILD_THREADS) void k_build_deferred(Batch a) {
    __shared__ uint64_t ent[MAXN][BUILD_THREADS];
    uint32_t const n = a.keystate->overflow_count;
    for (uint32_t i = blockIdx.x * BUILD_THREADS + threadIdx.x; i < n;
         i += gridDim.x * BUILD_THREADS) {
        uint2 const e = a.overflow[i];
        int m = 0;
        (void)build_one<MODE, MAXN, BUILD_THREADS, SKIN>(a, ent, threadIdx.x, e.x, e.y, m);
        if (m > 0) { atomicMax(&a.keystate->max_count, static_cast<unsigned int>(m)); }
    }
}

// ev (timing only, else null): start/stop events of the build launch and of
// the deferred pass, stamped by the dispatches themselves.
template <int MODE, bool SKIN>
void launch_build_mode(int cap, Batch const& a, hipStream_t s, hipEvent_t const* ev) {
    hipEvent_t const e0 = ev ? ev[0] : nullptr;
    hipEvent_t const e1 = ev ? ev[1] : nullptr;
    uint32_t const blocks = a.ntiles * (TILE / BUILD_WG);
    if (cap <= 16) {
        hipExtLaunchKernelGGL((k_build<MODE, 16, BUILD_WG, SKIN>), dim3(blocks), dim3(BUILD_WG), 0,
                              s, e0, e1, 0, a);
    } else if (cap <= 32) {
        hipExtLaunchKernelGGL((k_build<MODE, 32, BUILD_WG, SKIN>), dim3(blocks), dim3(BUILD_WG), 0,
                              s, e0, e1, 0, a);
    } else {
        // 64 entries of 64 lanes: 32 KB of LDS per workgroup.
        hipExtLaunchKernelGGL((k_build<MODE, MAXN, BUILD_THREADS, SKIN>),
                              dim3(a.ntiles * (TILE / BUILD_THREADS)), dim3(BUILD_THREADS), 0, s,
                              e0, e1, 0, a);
        if (ev) {  // no deferred pass: an empty interval
            (void)hipEventRecord(ev[2], s);
            (void)hipEventRecord(ev[3], s);
        }
        return;
    }
    // Deferred lists are rare (under 1/64 of the particles while the 16-entry
    // build is chosen): one workgroup per CU at most, most of them empty.
    uint32_t const dblocks = a.ntiles < 256 ? a.ntiles : 256;
    hipExtLaunchKernelGGL((k_build_deferred<MODE, SKIN>), dim3(dblocks), dim3(BUILD_THREADS), 0,
                          s, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr, 0, a);
}

// The lists from the cell windows, or (b.skin.reuse) from the skin candidates.
template <int MODE>
void launch_build_mode(int cap, Batch const& a, hipStream_t s, hipEvent_t const* ev) {
    if (a.skin.reuse) {
        launch_build_mode<MODE, true>(cap, a, s, ev);
    } else {
        launch_build_mode<MODE, false>(cap, a, s, ev);
    }
}

void launch_build(int mode, int cap, Batch const& a, hipStream_t s, hipEvent_t const* ev) {
    if (!a.ntiles) {
        if (ev) {
            for (int k = 0; k < 4; ++k) { (void)hipEventRecord(ev[k], s); }
        }
        return;
    }
    switch (mode) {
    case BUILD_IISPH: launch_build_mode<BUILD_IISPH>(cap, a, s, ev); break;
    case BUILD_DFSPH: launch_build_mode<BUILD_DFSPH>(cap, a, s, ev); break;
    case BUILD_AP: launch_build_mode<BUILD_AP>(cap, a, s, ev); break;
    default: launch_build_mode<BUILD_LISTS>(cap, a, s, ev); break;
    }
}

//------------------------------------------------------------------------------
// Skin candidate lists (SkinView in kernels.hpp).

// The columns (or rows) d in [-2, 2] around cell g of a point at coordinate x
// (in cell g) that may hold points within r (cell <= r < 2 cell) of it: -1, 0
// and 1 always, -2 when x - r falls below the low edge of cell g - 1 plus a
// margin, 2 when x + r reaches the low edge of cell g + 2 less the margin. The
// margin covers the float rounding of floor(x / cell) that assigns points to
// cells (relative 2^-24 of the coordinate) and of the edges here.
__device__ __forceinline__ void skin_span(float x, int32_t g, float cell, float r, int& lo,
                                          int& hi) {
    double const dx = static_cast<double>(x);
    double const dr = static_cast<double>(r);
    double const m = 1e-6 * (fabs(dx) + dr);
    double const edge_lo = (static_cast<double>(g) - 1.0) * cell;
    double const edge_hi = (static_cast<double>(g) + 2.0) * cell;
    lo = (dx - dr < edge_lo + m) ? -2 : -1;
    hi = (dx + dr > edge_hi - m) ? 2 : 1;
}

// One thread per fluid slot (tiles): the fluid and solid points with
// |v|^2 < skin_r2 in the reachable cells of the 5x5 window, in window order.
__global__ __launch_bounds__(TILE) void k_skin_candidates(Batch a, WindowArgs w) {
    TileRef const r = tile_ref(a, TILE);
    if (!r.valid) { return; }
    uint32_t const t = r.t;
    StepConsts const c = a.scenes[r.scene].c;
    DenseScene const ds = w.dscene[r.scene];
    V2 const p = a.pos[t];
    int2 const g = grid_coord(p, c.cell);
    float const rr = __builtin_sqrtf(c.skin_r2);
    int xlo, xhi, ylo, yhi;
    skin_span(p.x, g.x, c.cell, rr, xlo, xhi);
    skin_span(p.y, g.y, c.cell, rr, ylo, yhi);
    SolidView const& sv = a.solids;
    bool const solids = sv.pos && w.other_set.slots;
    int nf = 0;
    int ns = 0;
    bool over = false;
    for (int dy = ylo; dy <= yhi; ++dy) {
        // The row's cells: bins, then every range bound requested together.
        uint64_t nk[5];
        uint64_t bin[5];
        uint2 rg[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            int const dx = xlo + k;
            nk[k] = morton2(static_cast<int32_t>(static_cast<uint32_t>(g.x) + dx),
                            static_cast<int32_t>(static_cast<uint32_t>(g.y) + dy));
            bin[k] = (dx <= xhi) ? dense_bin(ds, nk[k], w.dense_bins) : w.dense_bins;
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            rg[k] = (bin[k] < w.dense_bins) ? make_uint2(w.doffset[bin[k]], w.doffset[bin[k] + 1])
                                            : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            for (uint32_t s0 = rg[k].x; s0 < rg[k].y; s0 += 4) {
                V2 q[4];
                load_chunk<4>(q, a.pos, s0);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    uint32_t const s = s0 + u;
                    V2 const v = sub(q[u], p);
                    float const l2 = v.x * v.x + v.y * v.y;
                    if (s < rg[k].y && s != t && l2 < c.skin_r2) {
                        if (nf < SKIN_FLUID) {
                            a.skin.cf[static_cast<size_t>(nf) * a.stride + t] = s;
                            ++nf;
                        } else {
                            over = true;
                        }
                    }
                }
            }
        }
        if (solids) {
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                int const dx = xlo + k;
                bool const maybe =
                  dx <= xhi && (bin[k] >= w.dense_bins || !w.other_bits ||
                                ((w.other_bits[bin[k] >> 5] >> (bin[k] & 31u)) & 1u));
                if (!maybe) { continue; }
                uint2 const rs = cell_set_range(w.other_set, w.other, nk[k], r.scene);
                for (uint32_t s = rs.x; s < rs.y; ++s) {
                    V2 const v = sub(sv.pos[s], p);
                    float const l2 = v.x * v.x + v.y * v.y;
                    if (l2 < c.skin_r2) {
                        if (ns < SKIN_SOLID) {
                            a.skin.cs[static_cast<size_t>(ns) * a.stride + t] = s;
                            ++ns;
                        } else {
                            over = true;
                        }
                    }
                }
            }
        }
    }
    a.skin.nf[t] = static_cast<uint8_t>(nf);
    a.skin.ns[t] = static_cast<uint8_t>(ns);
    a.skin.xref[t] = p;
    if (over) { atomicOr(&a.keystate->skin_flags, SKIN_OVERFLOW); }
}

void launch_skin_candidates(Batch const& b, WindowArgs const& w, hipStream_t s) {
    if (!b.ntiles) { return; }
    hipLaunchKernelGGL(k_skin_candidates, dim3(b.ntiles), dim3(TILE), 0, s, b, w);
}

// One workgroup: the sub-step start of every scene (as fold_key_range does)
// and the counters of this sub-step's build and reuse test cleared.
__global__ __launch_bounds__(256) void k_skin_begin(Batch b) {
    for (uint32_t s = threadIdx.x; s < b.nscenes; s += blockDim.x) {
        step_begin_scene(const_cast<SceneDev*>(b.scenes) + s, b.stats + s);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        KeyState* const ks = b.keystate;
        ks->frame_remaining = b.scenes[0].frame.on ? b.scenes[0].frame.next_remaining : -1;
        ks->sort_short = 0u;
        ks->dense_out = 0u;
        ks->dense_crowded = 0u;
        ks->overflow_count = 0u;
        ks->max_count = 0u;
        ks->long_lists = 0u;
        ks->skin_flags = 0u;
        ks->skin_disp = 0u;
    }
}

void launch_skin_begin(Batch const& b, hipStream_t s) {
    hipLaunchKernelGGL(k_skin_begin, dim3(1), dim3(256), 0, s, b);
}

// One thread per fluid slot (tiles): the state into the rebuild's sorted
// order (as the cells stage gathers it) and |x - x_rebuild|^2 against
// skin_lim2: a point at or past it makes the sub-step short (SKIN_STALE). The
// largest ratio per tile, folded into KeyState::skin_disp, guides the host's
// choice of the next rebuild.
__global__ __launch_bounds__(TILE) void k_skin_gather(Batch a, float const* kv_in,
                                                      float const* kd_in) {
    TileRef const r = tile_ref(a, TILE);
    float ratio = 0.0f;
    bool stale = false;
    if (r.valid) {
        uint32_t const t = r.t;
        StepConsts const c = a.scenes[r.scene].c;
        uint32_t const o = a.orig[t];
        V2 const x = a.pos_orig[o];
        a.pos[t] = x;
        a.vel[t] = a.vel_orig[o];
        if (kv_in) {
            a.kvs[t] = kv_in[o];
            a.kds[t] = kd_in[o];
        }
        V2 const d = sub(x, a.skin.xref[t]);
        float const d2 = d.x * d.x + d.y * d.y;
        stale = !(d2 < c.skin_lim2);  // the exact test; the ratio only guides the host
        ratio = stale ? 1.0f : d2 * (1.0f / c.skin_lim2);
    }
    if (__ballot(stale) && __lane_id() == 0) {
        atomicOr(&a.keystate->skin_flags, SKIN_STALE);
        atomicOr(&a.keystate->sort_short, 1u);
    }
    block_max<TILE>(__float_as_uint(ratio), a.skin.tmax + r.tile);
}

// One workgroup: the tiles' largest displacement ratios into KeyState::skin_disp.
__global__ __launch_bounds__(256) void k_skin_fold(Batch b) {
    unsigned int m = 0;
    for (uint32_t k = threadIdx.x; k < b.ntiles; k += 256u) {
        unsigned int const v = b.skin.tmax[k];
        m = v > m ? v : m;
    }
    block_max<256>(m, &b.keystate->skin_disp);
}

void launch_skin_gather(Batch const& b, float const* kv_in, float const* kd_in, hipStream_t s) {
    if (!b.ntiles) { return; }
    hipLaunchKernelGGL(k_skin_gather, dim3(b.ntiles), dim3(TILE), 0, s, b, kv_in, kd_in);
    hipLaunchKernelGGL(k_skin_fold, dim3(1), dim3(256), 0, s, b);
}

// solids.cpp:52-74 compute_volumes: 1.7 / (W0 + sum W) over solid-solid lists
// (the batch's "fluid" arrays are the solid arrays here).
__global__ __launch_bounds__(TILE) void k_solid_volumes(Batch b) {
    TileRef const r = tile_ref(b, TILE);
    if (!r.valid) { return; }
    KernelConsts const kc = b.scenes[r.scene].c.kc;
    uint32_t const t = r.t;
    int const cnt = b.fl.count[t];
    if (!cnt) {
        b.rho[t] = 1.0f / kc.w0;
        return;
    }
    float m = kc.w0;
    for (int k = 0; k < cnt; ++k) { m += kernel_W(b.fl.dist[static_cast<size_t>(k) * b.stride + t], kc); }
    b.rho[t] = 1.7f / m;
}

void launch_solid_volumes(Batch const& b, uint32_t n, hipStream_t s) {
    if (!n || !b.ntiles) { return; }
    hipLaunchKernelGGL(k_solid_volumes, dim3(b.ntiles), dim3(TILE), 0, s, b);
}

//------------------------------------------------------------------------------
// Neighbour loops of the sweeps walk a list in batches of NB entries:
// the batch's indices and gradients are loaded together, then the gathers of
// the neighbours' fields, then the entries are summed in list order. This
// keeps several dependent load chains in flight per thread without changing
// the order of any floating-point sum.
// The batch depth NB is a tuning choice (2, 4 or 8; Batch::nb).
template <int NB>
struct NbrBatch {
    uint32_t j[NB];
    V2 g[NB];
};

template <int NB>
__device__ __forceinline__ void load_batch(Lists const& l, uint32_t stride, uint32_t t, int k0,
                                           int cnt, NbrBatch<NB>& nb) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        nb.j[u] = 0u;
        nb.g[u] = V2{0.0f, 0.0f};
        if (k0 + u < cnt) {
            size_t const o = static_cast<size_t>(k0 + u) * stride + t;
            nb.j[u] = l.idx[o];
            nb.g[u] = l.gw[o];
        }
    }
}

//------------------------------------------------------------------------------
// IISPH sweeps: sph2d_box/iisph_ops.cpp.

// :94-221 a_ii and density stars fused with the solver loop's first
// iteration (:261-425 beta, new pressure and residual). The first iteration's
// inputs are all known here: the pressures start at +0 (iisph.cpp:83), so
// every sum_j d_ij p_j is exactly (+0, +0) (each term is (kf * +0) grad W /
// rho_j^2 = +-0, or an unsafe divide that is skipped, and +0 + +-0 = +0).
// beta is then exactly +0 as well: each of its terms is
// m dot((+0 - 0 d_jj) - (+0 - 0 d_ji), grad W), and d_jj, d_ji and grad W
// are finite (quotients that is_safe_div admits are below 2^126 in
// magnitude, the others are 0), so 0 d = +-0, +0 - +-0 = +0, the dot product
// is +-0 and +0 + +-0 = +0 for every term, fluid or solid. So beta is not
// summed (no d_jj gather), and rho0 - rho* - beta is rho0 - rho* (x - +0 = x).
// One pass over the list computes a_ii and rho*, and the first sum sweep is
// skipped. Workgroup 0 also starts the loop. Writes a_ii, rho*,
// {sum = (+0, +0), d_ii} and the first pressures {p', rho} into pr_out.
template <int NB>
__global__ __launch_bounds__(TILE) void k_iisph_aii_beta0(Batch a) {
    if (blockIdx.x == 0) { loop_begin_block<TILE>(a); }
    uint32_t const tile = tile_index(a);
    Tile const tl = a.tiles[tile];
    unsigned long long qsum = 0;
    unsigned int emax = 0;
    if (threadIdx.x < tl.count) {
        uint32_t const t = tl.start + threadIdx.x;
        StepConsts const c = a.scenes[tl.scene].c;
        float const mf = c.fluid_volume * c.rho0;  // = rho0 * V (commutative)
        float const neg_dt2 = -(c.dt * c.dt);
        float const kf = neg_dt2 * mf;
        float const dtm = c.dt * mf;
        int const fc = a.fl.count[t];
        V2 const d = a.dii[t];
        V2 const vi = a.vel[t];
        float const ri = a.rho[t];
        float const pi = 0.0f;
        V2 const si{0.0f, 0.0f};
        a.sd[t] = SdDii{si, d};
        float aii;
        float rs;
        if (!fc) {
            rs = ri;
            aii = 0.0f;
        } else {
            float acc = 0.0f;
            float dd = 0.0f;
            for (int k0 = 0; k0 < fc; k0 += NB) {
                NbrBatch<NB> nb;
                load_batch<NB>(a.fl, a.stride, t, k0, fc, nb);
                float rj[NB];
                V2 vj[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    rj[u] = 1.0f;
                    vj[u] = V2{0.0f, 0.0f};
                    if (k0 + u < fc) {
                        uint32_t const j = nb.j[u];
                        rj[u] = a.rho[j];
                        vj[u] = a.vel[j];
                    }
                }
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    if (k0 + u < fc) {
                        V2 const gw = nb.g[u];
                        V2 const numer = scale(kf, gw);
                        float const denom = rj[u] * rj[u];
                        bool const safe = is_safe_div(numer, denom);
                        V2 const dij = safe ? divs(numer, denom) : V2{0.0f, 0.0f};
                        V2 const dji = neg(dij);
                        if (safe) { acc += mf * dot(sub(d, dji), gw); }  // a_ii
                        dd += dtm * dot(sub(vi, vj[u]), gw);             // rho*
                    }
                }
            }
            aii = acc;
            rs = ri + dd;
        }
        int const sc = a.sl.count[t];
        if (sc) {
            float acc = 0.0f;
            float dd = 0.0f;
            for (int k = 0; k < sc; ++k) {
                size_t const o = static_cast<size_t>(k) * a.stride + t;
                uint32_t const b = a.sl.idx[o];
                V2 const gw = a.sl.gw[o];
                float const ms = a.solids.vol[b] * c.rho0;
                acc += ms * dot(d, gw);
                dd += c.dt * ms * dot(sub(vi, a.solids.vel[b]), gw);
            }
            aii += acc;
            rs += dd;
        }
        a.aii[t] = aii;
        a.rho_star[t] = rs;
        // iisph_compute_new_pressures for p = +0
        float pn = pi;
        if (aii != 0.0f) {
            float const numer = c.rho0 - rs;  // - beta, beta = +0
            if (is_safe_div(numer, aii)) {
                float err = -(numer - (aii * pi)) / c.rho0;
                err = (0.0f < err) ? err : 0.0f;  // std::max(0.0f, err)
                emax = __float_as_uint(err);
                qsum = static_cast<unsigned long long>(err * 1000.0f);
                float const np = numer / aii;
                float const mixed = (pi * (1.0f - c.omega)) + (np * c.omega);
                pn = (0.0f < mixed) ? mixed : 0.0f;
            }
        }
        a.pr_out[t] = V2{pn, ri};
    }
    block_residual<TILE>(qsum, emax, &a.tile_red[tile]);
}

// :225-258 sum_j d_ij p_j (fluid only) of one tile's particles.
template <int NB>
__device__ __forceinline__ void iisph_sdp_tile(Batch const& a, Tile const& tl) {
    if (threadIdx.x >= tl.count) { return; }
    uint32_t const t = tl.start + threadIdx.x;
    StepConsts const c = a.scenes[tl.scene].c;
    float const kf = -(c.dt * c.dt) * (c.fluid_volume * c.rho0);
    int const fc = a.fl.count[t];
    V2 s{0.0f, 0.0f};
    for (int k0 = 0; k0 < fc; k0 += NB) {
        NbrBatch<NB> nb;
        load_batch<NB>(a.fl, a.stride, t, k0, fc, nb);
        V2 prj[NB];  // {p_j, rho_j}
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            prj[u] = V2{0.0f, 1.0f};
            if (k0 + u < fc) { prj[u] = a.pr_in[nb.j[u]]; }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            if (k0 + u < fc) {
                V2 const numer = scale(kf * prj[u].x, nb.g[u]);
                float const denom = prj[u].y * prj[u].y;
                if (is_safe_div(numer, denom)) { s = add(s, divs(numer, denom)); }
            }
        }
    }
    a.sd[t] = SdDii{s, a.dii[t]};
}

// Converged scenes skip the sweep.
template <int NB>
__global__ __launch_bounds__(TILE) void k_iisph_sdp(Batch a, ActiveTiles at) {
    Tile const tl = a.tiles[active_tile(a, at)];
    if (!a.stats[tl.scene].active) { return; }
    iisph_sdp_tile<NB>(a, tl);
}

// :261-353 beta (into the new-pressure buffer) fused with :357-425 new
// pressure, quantised error sum and error max, for one tile (every thread of
// the workgroup calls it).
template <int NB>
__device__ __forceinline__ void iisph_beta_p_tile(Batch const& a, Tile const& tl, uint32_t tile) {
    unsigned long long qsum = 0;
    unsigned int emax = 0;
    if (threadIdx.x < tl.count) {
        uint32_t const t = tl.start + threadIdx.x;
        StepConsts const c = a.scenes[tl.scene].c;
        float const mf = c.rho0 * c.fluid_volume;
        float const kf = -(c.dt * c.dt) * mf;
        V2 const pri = a.pr_in[t];  // {p_i, rho_i}
        float const pi = pri.x;
        V2 const si = a.sd[t].sdp;
        int const fc = a.fl.count[t];
        float beta = 0.0f;
        for (int k0 = 0; k0 < fc; k0 += NB) {
            NbrBatch<NB> nb;
            load_batch<NB>(a.fl, a.stride, t, k0, fc, nb);
            float rj[NB], pj[NB];
            V2 sj[NB], djj[NB];
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                rj[u] = 1.0f;
                pj[u] = 0.0f;
                sj[u] = V2{0.0f, 0.0f};
                djj[u] = V2{0.0f, 0.0f};
                if (k0 + u < fc) {
                    uint32_t const j = nb.j[u];
                    V2 const prj = a.pr_in[j];
                    SdDii const sdj = a.sd[j];
                    rj[u] = prj.y;
                    pj[u] = prj.x;
                    sj[u] = sdj.sdp;
                    djj[u] = sdj.dii;
                }
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                if (k0 + u < fc) {
                    V2 const gw = nb.g[u];
                    V2 const numer = scale(kf, gw);
                    float const denom = rj[u] * rj[u];
                    V2 const dij =
                      is_safe_div(numer, denom) ? divs(numer, denom) : V2{0.0f, 0.0f};
                    V2 const dji = neg(dij);
                    V2 const sjk = sub(sj[u], scale(pi, dji));
                    V2 const djj_pj = scale(pj[u], djj[u]);
                    beta += mf * dot(sub(sub(si, djj_pj), sjk), gw);
                }
            }
        }
        int const sc = a.sl.count[t];
        if (sc) {
            float bs = 0.0f;
            for (int k = 0; k < sc; ++k) {
                size_t const o = static_cast<size_t>(k) * a.stride + t;
                uint32_t const b = a.sl.idx[o];
                float const ms = c.rho0 * a.solids.vol[b];
                bs += ms * dot(si, a.sl.gw[o]);
            }
            beta += bs;
        }
        // iisph_compute_new_pressures
        float const alpha = a.aii[t];
        float pn = pi;
        if (alpha != 0.0f) {
            float const numer = c.rho0 - a.rho_star[t] - beta;
            if (is_safe_div(numer, alpha)) {
                float err = -(numer - (alpha * pi)) / c.rho0;
                err = (0.0f < err) ? err : 0.0f;  // std::max(0.0f, err)
                emax = __float_as_uint(err);
                qsum = static_cast<unsigned long long>(err * 1000.0f);
                float const np = numer / alpha;
                float const mixed = (pi * (1.0f - c.omega)) + (np * c.omega);
                pn = (0.0f < mixed) ? mixed : 0.0f;
            }
        }
        a.pr_out[t] = V2{pn, pri.y};
    }
    block_residual<TILE>(qsum, emax, &a.tile_red[tile]);
}

// A converged scene skips the sweep; its result stays in the buffer of its
// last iteration.
template <int NB>
__global__ __launch_bounds__(TILE) void k_iisph_beta_p(Batch a, ActiveTiles at) {
    uint32_t const tile = active_tile(a, at);
    Tile const tl = a.tiles[tile];
    if (!a.stats[tl.scene].active) { return; }  // the whole workgroup
    iisph_beta_p_tile<NB>(a, tl, tile);
}

//------------------------------------------------------------------------------
// IISPH_AP (iisph_ap.cpp, iisph_ap_ops.cpp) and IISPH_PSEUDO_AP
// (iisph_pseudo_ap.cpp, iisph_pseudo_ap_ops.cpp; PSEUDO: a fictitious time
// step of 1 in the solve, so no dt^2 in the diagonal or A p). The pressures
// are one buffer updated in place (each particle reads only its own).

// p / rho^2, or 0 when the divide is unsafe (pressure.cpp:104-107, :134-137).
__device__ __forceinline__ float p_over_rho2(float p, float rho) {
    float const d = rho * rho;
    return is_safe_div(p, d) ? p / d : 0.0f;
}

// One iteration's pressure update of a particle with at least one neighbour
// (iisph_ap_ops.cpp:170-190, iisph_pseudo_ap_ops.cpp:281-301): A p in api
// (before the dt^2 factor of IISPH_AP), the old pressure p; returns the new
// one and the residual contribution.
template <int PSEUDO>
__device__ __forceinline__ float ap_update(StepConsts const& c, float api, float p,
                                           float rho_star, float diag,
                                           unsigned long long& qsum, unsigned int& emax) {
    if (!PSEUDO) { api *= c.dt * c.dt; }
    float const source = c.rho0 - rho_star;
    float const q = api - source;
    float err = is_safe_div(q, c.rho0) ? q / c.rho0 : 0.0f;
    err = (0.0f < err) ? err : 0.0f;  // std::max(0.0f, ...)
    emax = __float_as_uint(err);
    qsum = static_cast<unsigned long long>(err * 1000.0f);
    float const numer = c.omega * (source - api);
    float const step = is_safe_div(numer, diag) ? numer / diag : 0.0f;
    float const pn = p + step;
    return (0.0f < pn) ? pn : 0.0f;
}

// iisph_ap_ops.cpp:19-106 / iisph_pseudo_ap_ops.cpp:19-110 diagonals and
// density stars, fused with the loop's first iteration: the pressures start
// at +0 (iisph_ap.cpp:95), so every pressure acceleration is exactly (+0, +0)
// ((-m (+0 + +0)) grad W = +-0 terms added to +0) and so is A p; the first
// update needs only the particle's own diagonal and density star. Workgroup 0
// also starts the loop.
template <int NB, int PSEUDO>
__global__ __launch_bounds__(TILE) void k_ap_diag0(Batch a) {
    if (blockIdx.x == 0) { loop_begin_block<TILE>(a); }
    uint32_t const tile = tile_index(a);
    Tile const tl = a.tiles[tile];
    unsigned long long qsum = 0;
    unsigned int emax = 0;
    if (threadIdx.x < tl.count) {
        uint32_t const t = tl.start + threadIdx.x;
        StepConsts const c = a.scenes[tl.scene].c;
        int const fc = a.fl.count[t];
        int const sc = a.sl.count[t];
        float const rho = a.rho[t];
        float diag = 0.0f;
        float rs = rho;
        float pn = 0.0f;
        if (fc + sc) {
            float const mf = c.fluid_volume * c.rho0;  // fluid mass (uniform volume)
            V2 const vi = a.vel[t];
            V2 sm{0.0f, 0.0f};
            float smgg = 0.0f;
            float div = 0.0f;
            for (int k0 = 0; k0 < fc; k0 += NB) {
                NbrBatch<NB> nb;
                load_batch<NB>(a.fl, a.stride, t, k0, fc, nb);
                V2 vj[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    vj[u] = V2{0.0f, 0.0f};
                    if (k0 + u < fc) { vj[u] = a.vel[nb.j[u]]; }
                }
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    if (k0 + u < fc) {
                        V2 const gw = nb.g[u];
                        sm = add(sm, scale(mf, gw));
                        smgg += mf * dot(gw, gw);
                        div += mf * dot(sub(vi, vj[u]), gw);
                    }
                }
            }
            for (int k = 0; k < sc; ++k) {
                size_t const o = static_cast<size_t>(k) * a.stride + t;
                uint32_t const b = a.sl.idx[o];
                V2 const gw = a.sl.gw[o];
                float const ms = a.solids.vol[b] * c.rho0;
                sm = add(sm, scale(ms, gw));
                div += ms * dot(sub(vi, a.solids.vel[b]), gw);
            }
            float const inner = dot(sm, sm) + mf * smgg;
            float const numer = PSEUDO ? -inner : -(c.dt * c.dt) * inner;
            float const denom = rho * rho;
            diag = is_safe_div(numer, denom) ? numer / denom : 0.0f;
            rs = rho + c.dt * div;
            if (!(PSEUDO && diag == 0.0f)) {
                pn = ap_update<PSEUDO>(c, 0.0f, 0.0f, rs, diag, qsum, emax);
            }
        }
        a.aii[t] = diag;
        a.rho_star[t] = rs;
        a.p_out[t] = pn;
    }
    block_residual<TILE>(qsum, emax, &a.tile_red[tile]);
}

// Pressure acceleration of slot t (pressure.cpp:83-178; also the pseudo
// displacements, iisph_pseudo_ap_ops.cpp:113-225): the fluid pass sets it, the
// boundary pass adds its own sum.
template <int NB>
__device__ __forceinline__ V2 ap_acceleration(Batch const& a, StepConsts const& c, uint32_t t) {
    int const fc = a.fl.count[t];
    int const sc = a.sl.count[t];
    float const* P = a.p_in;
    float const pi = p_over_rho2(P[t], a.rho[t]);
    float const mf = c.rho0 * c.fluid_volume;
    V2 acc{0.0f, 0.0f};
    for (int k0 = 0; k0 < fc; k0 += NB) {
        NbrBatch<NB> nb;
        load_batch<NB>(a.fl, a.stride, t, k0, fc, nb);
        float pj[NB], rj[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            pj[u] = 0.0f;
            rj[u] = 1.0f;
            if (k0 + u < fc) {
                pj[u] = P[nb.j[u]];
                rj[u] = a.rho[nb.j[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            if (k0 + u < fc) {
                float const coeff = pi + p_over_rho2(pj[u], rj[u]);
                acc = add(acc, scale(-mf * coeff, nb.g[u]));
            }
        }
    }
    if (sc) {
        V2 ab{0.0f, 0.0f};
        for (int k = 0; k < sc; ++k) {
            size_t const o = static_cast<size_t>(k) * a.stride + t;
            float const ms = c.rho0 * a.solids.vol[a.sl.idx[o]];
            ab = add(ab, scale(-ms * pi, a.sl.gw[o]));
        }
        acc = add(acc, ab);
    }
    return acc;
}

// Accelerations from the current pressures, for the scenes still iterating.
template <int NB>
__global__ __launch_bounds__(TILE) void k_ap_accel(Batch a, ActiveTiles at) {
    TileRef const r = tile_ref_of(a, active_tile(a, at));
    if (!a.stats[r.scene].active) { return; }  // the whole workgroup
    if (!r.valid) { return; }
    StepConsts const c = a.scenes[r.scene].c;
    a.pacc[r.t] = ap_acceleration<NB>(a, c, r.t);
}

// iisph_ap_ops.cpp:108-194 / iisph_pseudo_ap_ops.cpp:227-305: A p from the
// accelerations, the in-place pressure update and its residual.
template <int NB, int PSEUDO>
__global__ __launch_bounds__(TILE) void k_ap_iterate(Batch a, ActiveTiles at) {
    uint32_t const tile = active_tile(a, at);
    Tile const tl = a.tiles[tile];
    if (!a.stats[tl.scene].active) { return; }  // the whole workgroup
    unsigned long long qsum = 0;
    unsigned int emax = 0;
    if (threadIdx.x < tl.count) {
        uint32_t const t = tl.start + threadIdx.x;
        StepConsts const c = a.scenes[tl.scene].c;
        int const fc = a.fl.count[t];
        int const sc = a.sl.count[t];
        float* const P = a.p_out;
        float const diag = a.aii[t];
        if (!fc && !sc) {
            P[t] = 0.0f;
        } else if (PSEUDO && diag == 0.0f) {
            P[t] = 0.0f;
        } else {
            float const mf = c.rho0 * c.fluid_volume;
            V2 const ai = a.pacc[t];
            float api = 0.0f;
            for (int k0 = 0; k0 < fc; k0 += NB) {
                NbrBatch<NB> nb;
                load_batch<NB>(a.fl, a.stride, t, k0, fc, nb);
                V2 aj[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    aj[u] = V2{0.0f, 0.0f};
                    if (k0 + u < fc) { aj[u] = a.pacc[nb.j[u]]; }
                }
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    if (k0 + u < fc) { api += mf * dot(sub(ai, aj[u]), nb.g[u]); }
                }
            }
            for (int k = 0; k < sc; ++k) {
                size_t const o = static_cast<size_t>(k) * a.stride + t;
                float const ms = c.rho0 * a.solids.vol[a.sl.idx[o]];
                api += ms * dot(ai, a.sl.gw[o]);
            }
            P[t] = ap_update<PSEUDO>(c, api, P[t], a.rho_star[t], diag, qsum, emax);
        }
    }
    block_residual<TILE>(qsum, emax, &a.tile_red[tile]);
}

// The final accelerations (every scene) and the integration: IISPH_AP
// dynamics.cpp:53-62 (v += dt a, x += dt v); IISPH_PSEUDO_AP
// iisph_pseudo_ap_ops.cpp:307-341 (x += dt v + d, v += d / dt when safe);
// state written back in original order, max |v|^2 and energies per tile.
template <int NB, int PSEUDO>
__global__ __launch_bounds__(TILE) void k_ap_final(Batch a) {
    TileRef const r = tile_ref(a, TILE);
    unsigned int vm = 0;
    double ke = 0.0, pe = 0.0;
    if (r.valid) {
        uint32_t const t = r.t;
        StepConsts const c = a.scenes[r.scene].c;
        V2 const acc = ap_acceleration<NB>(a, c, t);
        a.pacc[t] = acc;
        V2 v = a.vel[t];
        V2 x = a.pos[t];
        if (PSEUDO) {
            V2 const vs = v;
            x = add(add(x, scale(c.dt, vs)), acc);
            if (is_safe_div(vs, c.dt)) { v = add(vs, divs(acc, c.dt)); }
        } else {
            v = add(v, scale(c.dt, acc));
            x = add(x, scale(c.dt, v));
        }
        uint32_t const o = a.orig[t];
        a.pos_orig[o] = x;
        a.vel_orig[o] = v;
        vm = __float_as_uint(dot(v, v));
        particle_energy(c, x, v, ke, pe);
    }
    block_max<TILE>(vm, a.tile_vmax + r.tile);
    block_energy<TILE>(ke, pe, a.tile_energy + r.tile);
}

//------------------------------------------------------------------------------
// pressure.cpp:9-80 apply_pressures_in_place (fluid pass, then solid pass).
template <int GRAVITY, int FINAL, int NB>
__global__ __launch_bounds__(TILE) void k_apply(Batch a, int all_scenes, ActiveTiles at) {
    TileRef const r = tile_ref_of(a, active_tile(a, at));
    if (!all_scenes && !a.stats[r.scene].active) { return; }  // whole workgroup
    unsigned int vm = 0;
    double ke = 0.0, pe = 0.0;
    if (r.valid) {
        uint32_t const t = r.t;
        StepConsts const c = a.scenes[r.scene].c;
        V2 v = a.vel[t];
        if (GRAVITY) {
            V2 const f = a.fext ? a.fext[t] : V2{0.0f, 0.0f - c.mass * c.gravity};
            float const vden = c.rho0 * c.fluid_volume;
            V2 const numer = scale(c.dt, f);
            if (is_safe_div(numer, vden)) { v = add(v, divs(numer, vden)); }
        }
        int const fc = a.fl.count[t];
        int const sc = a.sl.count[t];
        // FINAL (the IISPH end of the sub-step): pressures and densities from
        // the scene's last {p, rho} buffer (odd iteration count: pr_alt); the
        // pressure is also written to p_out for the field read-back.
        // Otherwise (DFSPH): the stiffnesses p_in and densities rho.
        float const* P = a.p_in;
        V2 const* PR = nullptr;
        if (FINAL) {
            PR = (a.pr_alt && (a.stats[r.scene].iterations & 1)) ? a.pr_alt : a.pr_in;
            a.p_out[t] = PR[t].x;
        }
        if (fc || sc) {
            float const ri = FINAL ? PR[t].y : a.rho[t];
            float const ri2 = ri * ri;
            float const pi = FINAL ? PR[t].x : P[t];
            float const pori = is_safe_div(pi, ri2) ? pi / ri2 : 0.0f;
            if (fc) {
                float const mf = c.rho0 * c.fluid_volume;
                V2 acc{0.0f, 0.0f};
                for (int k0 = 0; k0 < fc; k0 += NB) {
                    NbrBatch<NB> nb;
                    load_batch<NB>(a.fl, a.stride, t, k0, fc, nb);
                    float rj[NB], pj[NB];
#pragma unroll
                    for (int u = 0; u < NB; ++u) {
                        rj[u] = 1.0f;
                        pj[u] = 0.0f;
                        if (k0 + u < fc) {
                            if (FINAL) {
                                V2 const prj = PR[nb.j[u]];
                                rj[u] = prj.y;
                                pj[u] = prj.x;
                            } else {
                                rj[u] = a.rho[nb.j[u]];
                                pj[u] = P[nb.j[u]];
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < NB; ++u) {
                        if (k0 + u < fc) {
                            float const rj2 = rj[u] * rj[u];
                            float const porj = is_safe_div(pj[u], rj2) ? pj[u] / rj2 : 0.0f;
                            acc = add(acc, scale(-mf * (pori + porj), nb.g[u]));
                        }
                    }
                }
                v = add(v, scale(c.dt, acc));
            }
            if (sc) {
                V2 acc{0.0f, 0.0f};
                for (int k = 0; k < sc; ++k) {
                    size_t const o = static_cast<size_t>(k) * a.stride + t;
                    uint32_t const b = a.sl.idx[o];
                    float const ms = c.rho0 * a.solids.vol[b];
                    acc = add(acc, scale(-ms * pori, a.sl.gw[o]));
                }
                v = add(v, scale(c.dt, acc));
            }
        }
        if (FINAL) {
            // dynamics.cpp:46-51: x += dt * v, written back in original order.
            V2 const x = add(a.pos[t], scale(c.dt, v));
            uint32_t const o = a.orig[t];
            a.pos_orig[o] = x;
            a.vel_orig[o] = v;
            vm = __float_as_uint(dot(v, v));
            particle_energy(c, x, v, ke, pe);
        } else {
            a.vel[t] = v;
        }
    }
    if (FINAL) {
        block_max<TILE>(vm, a.tile_vmax + r.tile);
        block_energy<TILE>(ke, pe, a.tile_energy + r.tile);
    }
}

// x += dt v and scatter the sorted state back to original order
// (dfsph_p.cpp:368-369), including the DFSPH warm-start stiffnesses.
template <int DFSPH>
__global__ __launch_bounds__(TILE) void k_finalize(Batch a) {
    TileRef const r = tile_ref(a, TILE);
    unsigned int vm = 0;
    double ke = 0.0, pe = 0.0;
    if (r.valid) {
        uint32_t const t = r.t;
        StepConsts const c = a.scenes[r.scene].c;
        V2 const v = a.vel[t];
        V2 const x = add(a.pos[t], scale(c.dt, v));
        uint32_t const o = a.orig[t];
        a.pos_orig[o] = x;
        a.vel_orig[o] = v;
        if (DFSPH) {
            a.kv_orig[o] = a.kvs[t];
            a.kd_orig[o] = a.kds[t];
        }
        vm = __float_as_uint(dot(v, v));
        particle_energy(c, x, v, ke, pe);
    }
    block_max<TILE>(vm, a.tile_vmax + r.tile);
    block_energy<TILE>(ke, pe, a.tile_energy + r.tile);
}

//------------------------------------------------------------------------------
// DFSPH_P sweeps: divergence.cpp:5-74 and dfsph_p_ops.cpp:87-126, with the
// residual of common.h:209-232 fused in.
template <int NB>
__device__ __forceinline__ float divergence_of(Batch const& a, StepConsts const& c, uint32_t t,
                                               V2 vi) {
    float const mf = c.rho0 * c.fluid_volume;
    int const fc = a.fl.count[t];
    float div = 0.0f;
    for (int k0 = 0; k0 < fc; k0 += NB) {
        NbrBatch<NB> nb;
        load_batch<NB>(a.fl, a.stride, t, k0, fc, nb);
        V2 vj[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            vj[u] = V2{0.0f, 0.0f};
            if (k0 + u < fc) { vj[u] = a.vel[nb.j[u]]; }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            if (k0 + u < fc) { div += mf * dot(sub(vi, vj[u]), nb.g[u]); }
        }
    }
    int const sc = a.sl.count[t];
    if (sc) {
        float ds = 0.0f;
        for (int k = 0; k < sc; ++k) {
            size_t const o = static_cast<size_t>(k) * a.stride + t;
            uint32_t const b = a.sl.idx[o];
            float const ms = c.rho0 * a.solids.vol[b];
            ds += ms * dot(sub(vi, a.solids.vel[b]), a.sl.gw[o]);
        }
        div += ds;
    }
    return div;
}

// One divergence-solve iteration: divergence, kappa_v, residual of
// max(0, dt * div / rho0) (dfsph_p.cpp:157-197).
template <int NB>
__global__ __launch_bounds__(TILE) void k_dfsph_div(Batch a, ActiveTiles at) {
    TileRef const r = tile_ref_of(a, active_tile(a, at));
    if (!a.stats[r.scene].active) { return; }
    unsigned long long qsum = 0;
    unsigned int emax = 0;
    if (r.valid) {
        uint32_t const t = r.t;
        StepConsts const c = a.scenes[r.scene].c;
        float const div = divergence_of<NB>(a, c, t, a.vel[t]);
        a.rho_star[t] = div;
        float kv = 0.0f;
        if (!(div <= 0.0f)) {
            float const num = div * a.alpha[t];
            kv = is_safe_div(num, c.dt) ? num / c.dt : 0.0f;
        }
        a.kappa[t] = kv;
        float const val = (div <= 0.0f) ? 0.0f : (c.dt * div / c.rho0);
        emax = __float_as_uint(val);
        qsum = static_cast<unsigned long long>(static_cast<long long>(val * 1000.0f));
    }
    block_residual<TILE>(qsum, emax, &a.tile_red[r.tile]);
}

// One density-solve iteration: density star (density.cpp:111-145), kappa_rho,
// residual of max(0, (rho* - rho0) / rho0) (dfsph_p.cpp:82-122).
template <int NB>
__global__ __launch_bounds__(TILE) void k_dfsph_dens(Batch a, ActiveTiles at) {
    TileRef const r = tile_ref_of(a, active_tile(a, at));
    if (!a.stats[r.scene].active) { return; }
    unsigned long long qsum = 0;
    unsigned int emax = 0;
    if (r.valid) {
        uint32_t const t = r.t;
        StepConsts const c = a.scenes[r.scene].c;
        float const div = divergence_of<NB>(a, c, t, a.vel[t]);
        float rs = a.rho[t] + (c.dt * div);
        rs = (0.0f < rs) ? rs : 0.0f;  // std::max(0.0f, ...)
        a.rho_star[t] = rs;
        float const dd = rs - c.rho0;
        float kd = 0.0f;
        if (!(dd <= 0.0f)) {
            float const num = dd * a.alpha[t];
            float const den = c.dt * c.dt;
            kd = is_safe_div(num, den) ? num / den : 0.0f;
        }
        a.kappa[t] = kd;
        float const err = rs - c.rho0;
        float const val = (err <= 0.0f) ? 0.0f : (err / c.rho0);
        emax = __float_as_uint(val);
        qsum = static_cast<unsigned long long>(static_cast<long long>(val * 1000.0f));
    }
    block_residual<TILE>(qsum, emax, &a.tile_red[r.tile]);
}

//------------------------------------------------------------------------------
// External forces (sph2d_box/forces.cpp:11-149).

// util/random.h:112-134 Lehmer_rand_gen_64 seeded with i, first draw: the
// 128-bit state i is multiplied by 0xda942042e4dd58b5 three times (twice at
// seeding, once per draw) and the draw is the state's high 64 bits. The state
// is kept as two 64-bit halves: (hi, lo) * m = (hi m + mulhi(lo, m), lo m)
// modulo 2^128.
__device__ __forceinline__ uint64_t lehmer64_first_draw(uint64_t seed) {
    constexpr uint64_t MUL = 0xda942042e4dd58b5ull;
    uint64_t lo = seed;
    uint64_t hi = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint64_t const h = hi * MUL + __umul64hi(lo, MUL);
        lo = lo * MUL;
        hi = h;
    }
    return hi;
}

// forces.cpp:134-141: std::uniform_real_distribution<float>{-pi, pi} over that
// generator, as the host C++ library (libstdc++ bits/random.tcc) computes it:
// generate_canonical<float, 24> takes one draw u of the 2^64-wide range,
// s = float(u) / 2^64, clamped below 1 (to the largest float below 1), and the
// distribution returns s * (b - a) + a in float.
__device__ __forceinline__ float anti_coupling_angle(uint64_t i) {
    float const a = -3.14159274101257324219f;  // float(-M_PI)
    float const b = 3.14159274101257324219f;   // float(M_PI)
    float s = static_cast<float>(lehmer64_first_draw(i)) * 5.42101086242752217004e-20f;  // 2^-64
    if (s >= 1.0f) { s = 0.99999994039535522461f; }  // nextafter(1, 0)
    return s * (b - a) + a;
}

// cos and sin of the angle rounded from double precision (the host library's
// float cos/sin are accurate to well under one ulp, so the two agree but for
// rare last-bit differences).
__device__ __forceinline__ V2 unit_direction(float angle) {
    double const d = static_cast<double>(angle);
    return V2{static_cast<float>(cos(d)), static_cast<float>(sin(d))};
}

template <int APPLY>
__global__ __launch_bounds__(TILE) void k_external_forces(Batch a, int host_forces) {
    TileRef const r = tile_ref(a, TILE);
    if (!r.valid) { return; }
    uint32_t const t = r.t;
    SceneDev const& sd = a.scenes[r.scene];
    StepConsts const c = sd.c;
    ForceDev const fm = sd.force;
    V2 f = host_forces ? a.fext[t] : V2{0.0f, 0.0f};
    V2 const x = a.pos[t];
    V2 v = a.vel[t];
    if (fm.flags & FORCE_GRAVITY) {
        // :62-69 accumulate_gravity_forces
        f.y -= c.mass * c.gravity;
    }
    if (fm.flags & FORCE_POLE) {
        // :85-114 accumulate_constant_pole_attraction_forces: r = x - pole,
        // Imath normalized() (zero for a zero length), forces -= magnitude rN,
        // then += the x, y components of the cross product of the torque
        // (0, 0, 0.125 magnitude) with (rN, 0).
        V2 const rv = sub(x, V2{fm.pole_x, fm.pole_y});
        float const l = imath_length(rv);
        V2 const rn = (l == 0.0f) ? V2{0.0f, 0.0f} : V2{rv.x / l, rv.y / l};
        f = sub(f, scale(fm.pole_mag, rn));
        float const tz = 0.125f * fm.pole_mag;
        V2 const turn{0.0f * 0.0f - tz * rn.y, tz * rn.x - 0.0f * 0.0f};
        f = add(f, turn);
    }
    if (fm.flags & FORCE_DRAG) {
        // :116-124 accumulate_simple_drag_forces: forces += (-magnitude d) v
        f = add(f, scale(fm.drag_mult, v));
    }
    if (fm.flags & FORCE_ANTI_COUPLING) {
        // :126-149: one kick per fluid neighbour closer than the limit; the
        // list is in (distance, index) order, so the walk stops at the first
        // neighbour that is not. Every kick of particle i has the same angle
        // (its generator is re-seeded with i for each neighbour).
        int const cnt = a.fl.count[t];
        bool have = false;
        V2 kick{0.0f, 0.0f};
        for (int k = 0; k < cnt; ++k) {
            uint32_t const o = a.fl.idx[static_cast<size_t>(k) * a.stride + t];
            float const dist = imath_length(sub(a.pos[o], x));
            if (!(dist < fm.anti_maxd)) { break; }
            if (!have) {
                uint64_t const i = a.orig[t] - sd.offset;
                kick = scale(fm.anti_mag, unit_direction(anti_coupling_angle(i)));
                have = true;
            }
            f = add(f, kick);
        }
    }
    a.fext[t] = f;
    if (APPLY) {
        // dynamics.cpp:24-37 integrate_velocities_in_place
        float const vden = c.rho0 * c.fluid_volume;
        V2 const numer = scale(c.dt, f);
        if (is_safe_div(numer, vden)) { v = add(v, divs(numer, vden)); }
        a.vel[t] = v;
    }
}

void launch_external_forces(Batch const& b, int apply, int host_forces, hipStream_t s) {
    if (!b.ntiles || !b.fext) { return; }
    if (apply) {
        hipLaunchKernelGGL(k_external_forces<1>, dim3(b.ntiles), dim3(TILE), 0, s, b, host_forces);
    } else {
        hipLaunchKernelGGL(k_external_forces<0>, dim3(b.ntiles), dim3(TILE), 0, s, b, host_forces);
    }
}

//------------------------------------------------------------------------------
// Colors (sph2d_box/colors.cpp:9-70).

// std::clamp(v, lo, hi) as the host library writes it (NaN passes through).
__device__ __forceinline__ float clamp01(float v) {
    return (v < 0.0f) ? 0.0f : ((1.0f < v) ? 1.0f : v);
}

// :9-20 c = clamp(1 - rho / (1.1 rho0), 0, 1); RGBA (c, c, 1, 1).
__global__ __launch_bounds__(TILE) void k_colors(Batch a, float4* out) {
    TileRef const r = tile_ref(a, TILE);
    if (!r.valid) { return; }
    uint32_t const t = r.t;
    float const rho0 = a.scenes[r.scene].c.rho0;
    float const norm_d = a.rho[t] / (1.1f * rho0);
    float const c = clamp01(1.0f - norm_d);
    out[a.orig[t]] = make_float4(c, c, 1.0f, 1.0f);
}

// :57-70 each channel static_cast<uint8_t>(255 * clamp(f, 0, 1)).
__global__ void k_convert_colors(float4 const* in, uchar4* out, uint32_t n) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    float4 const c = in[i];
    out[i] = make_uchar4(static_cast<unsigned char>(255.0f * clamp01(c.x)),
                         static_cast<unsigned char>(255.0f * clamp01(c.y)),
                         static_cast<unsigned char>(255.0f * clamp01(c.z)),
                         static_cast<unsigned char>(255.0f * clamp01(c.w)));
}

void launch_colors(Batch const& b, float4* out, hipStream_t s) {
    if (!b.ntiles) { return; }
    hipLaunchKernelGGL(k_colors, dim3(b.ntiles), dim3(TILE), 0, s, b, out);
}

void launch_convert_colors(float4 const* in, uchar4* out, uint32_t n, hipStream_t s) {
    if (!n) { return; }
    hipLaunchKernelGGL(k_convert_colors, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n);
}

//------------------------------------------------------------------------------
// PCISPH: sph2d_box/pcisph.cpp:36-164 and pcisph_ops.cpp.

// :54-88 first predictor: v_ext = v + (dt/m) F_ext; v* = v_ext; x* = x + dt v*;
// p = 0.
__global__ __launch_bounds__(TILE) void k_pci_init(Batch a) {
    TileRef const r = tile_ref(a, TILE);
    if (!r.valid) { return; }
    uint32_t const t = r.t;
    StepConsts const c = a.scenes[r.scene].c;
    V2 const f = a.fext ? a.fext[t] : V2{0.0f, 0.0f - c.mass * c.gravity};
    float const dtm = c.dt / c.mass;
    V2 const vext = add(a.vel[t], scale(dtm, f));
    a.vext[t] = vext;
    a.vstar_out[t] = vext;
    a.xstar[t] = add(a.pos[t], scale(c.dt, vext));
    a.p_out[t] = 0.0f;
}

// :90-115 geometry at x* (neighborhoods_from_state.cpp:103-128), densities
// (pcisph_ops.cpp:18-63), pressure update (:65-76) and the max density error
// (common.cpp:78-85). PCISPH densities are kept in rho. Of the recomputed
// pair geometry only the kernel gradients are stored (the force sweep takes
// the vectors and distances again from x*, the same arithmetic); the vectors
// and distances are written as well when they are kept for a read-back.
// Neighbour entries are taken NB at a time (their loads in flight together)
// and processed in list order.
template <int NB>
__global__ __launch_bounds__(TILE) void k_pci_density(Batch a, ActiveTiles at) {
    TileRef const r = tile_ref_of(a, active_tile(a, at));
    if (!a.stats[r.scene].active) { return; }
    unsigned int emax = 0;
    if (r.valid) {
        uint32_t const t = r.t;
        StepConsts const c = a.scenes[r.scene].c;
        V2 const xs = a.xstar_in[t];
        int const fc = a.fl.count[t];
        float w = c.kc.w0;
        for (int k0 = 0; k0 < fc; k0 += NB) {
            uint32_t j[NB];
            V2 xj[NB];
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                j[u] = 0u;
                if (k0 + u < fc) { j[u] = a.fl.idx[static_cast<size_t>(k0 + u) * a.stride + t]; }
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                xj[u] = xs;
                if (k0 + u < fc) { xj[u] = a.xstar_in[j[u]]; }
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                if (k0 + u < fc) {
                    size_t const o = static_cast<size_t>(k0 + u) * a.stride + t;
                    V2 const vec = sub(xj[u], xs);
                    float const d = imath_length(vec);
                    if (a.store_geometry) {
                        a.fl.vec[o] = vec;
                        a.fl.dist[o] = d;
                    }
                    float wk;
                    V2 gw;
                    kernel_W_GradW_r(neg(vec), d, c.kc, wk, gw);
                    a.fl.gw[o] = gw;
                    w += wk;
                }
            }
        }
        float rho = c.mass * w;
        int const sc = a.sl.count[t];
        if (sc) {
            float vf = 0.0f;
            for (int k = 0; k < sc; ++k) {
                size_t const o = static_cast<size_t>(k) * a.stride + t;
                uint32_t const b = a.sl.idx[o];
                V2 const vec = sub(a.solids.pos[b], xs);
                float const d = imath_length(vec);
                float wk;
                V2 gw;
                kernel_W_GradW_r(neg(vec), d, c.kc, wk, gw);
                a.sl.gw[o] = gw;
                vf += a.solids.vol[b] * wk;
            }
            rho += c.rho0 * vf;
        }
        a.rho[t] = rho;
        float e = rho - c.rho0;
        e = (e < 0.0f) ? 0.0f : e;  // std::max(rho - rho0, 0.0f)
        a.p_out[t] = a.p_out[t] + e / c.pci_denom;
        emax = __float_as_uint(e);
    }
    block_residual<TILE>(0ull, emax, &a.tile_red[r.tile]);
}

// :117-142 pressure + artificial viscosity forces, solid forces, then the next
// predictor (:73-88) from the new force. A converged scene carries its last
// predictor forward. Neighbour entries are taken NB at a time.
template <int NB>
__global__ __launch_bounds__(TILE) void k_pci_force(Batch a, ActiveTiles at) {
    TileRef const r = tile_ref_of(a, active_tile(a, at));
    if (!r.valid) { return; }
    uint32_t const t = r.t;
    if (!a.stats[r.scene].active) { return; }  // last predictor stays in place
    StepConsts const c = a.scenes[r.scene].c;
    float const H = c.kc.h;
    float const M = c.mass;
    float const N2 = 0.01f * H * H;
    float const* rho = a.rho;
    float const* p = a.p_out;
    V2 F{0.0f, 0.0f};
    float const density = rho[t];
    float const por = p[t] / (density * density);
    int const fc = a.fl.count[t];
    if (fc) {
        V2 const v = a.vstar_in[t];
        V2 const xs = a.xstar_in[t];
        V2 f{0.0f, 0.0f};
        for (int k0 = 0; k0 < fc; k0 += NB) {
            uint32_t j[NB];
            V2 gw[NB], vj[NB], xj[NB];
            float od[NB], pj[NB];
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                j[u] = 0u;
                gw[u] = V2{0.0f, 0.0f};
                if (k0 + u < fc) {
                    size_t const o = static_cast<size_t>(k0 + u) * a.stride + t;
                    j[u] = a.fl.idx[o];
                    gw[u] = a.fl.gw[o];
                }
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                od[u] = 1.0f;
                pj[u] = 0.0f;
                vj[u] = v;
                xj[u] = xs;
                if (k0 + u < fc) {
                    od[u] = rho[j[u]];
                    pj[u] = p[j[u]];
                    vj[u] = a.vstar_in[j[u]];
                    xj[u] = a.xstar_in[j[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                if (k0 + u < fc) {
                    f = add(f, scale(por + (pj[u] / (od[u] * od[u])), gw[u]));
                    V2 const dv = sub(vj[u], v);
                    // The pair's vector and distance at x*, as the density
                    // sweep took them (neighborhoods_from_state.cpp:103-128).
                    V2 const rv = sub(xj[u], xs);
                    float const dvr = dot(rv, dv);
                    if (dvr < 0.0f) {
                        float const dist = imath_length(rv);
                        float const mu = H * dvr / (dist + N2);
                        float const den_avg = (density + od[u]) / 2.0f;
                        float const pi_ab =
                          ((-1.0f * c.viscosity * mu) + (2.0f * (mu * mu))) / den_avg;
                        f = add(f, scale(pi_ab, gw[u]));
                    }
                }
            }
        }
        F = add(F, scale(-(M * M), f));
    }
    int const sc = a.sl.count[t];
    if (sc) {
        V2 g{0.0f, 0.0f};
        for (int k = 0; k < sc; ++k) {
            size_t const o = static_cast<size_t>(k) * a.stride + t;
            uint32_t const b = a.sl.idx[o];
            g = add(g, scale(a.solids.vol[b], a.sl.gw[o]));
        }
        F = add(F, scale(-M * c.rho0 * por, g));
    }
    a.force[t] = F;
    float const dtm = c.dt / c.mass;
    V2 const vn = add(a.vext[t], scale(dtm, F));
    a.vstar_out[t] = vn;
    a.xstar[t] = add(a.pos[t], scale(c.dt, vn));
}

// :151-163 final update: v = v_ext + (dt/m) F; x += dt v. Both equal the last
// predictor written by k_pci_force (in vstar_alt and xstar_alt after an odd
// number of iterations).
__global__ __launch_bounds__(TILE) void k_pci_finalize(Batch a) {
    TileRef const r = tile_ref(a, TILE);
    unsigned int vm = 0;
    double ke = 0.0, pe = 0.0;
    if (r.valid) {
        uint32_t const t = r.t;
        bool const odd = (a.stats[r.scene].iterations & 1) != 0;
        V2 const v = (odd ? a.vstar_alt : a.vstar_in)[t];
        V2 const x = (odd ? a.xstar_alt : a.xstar_in)[t];
        uint32_t const o = a.orig[t];
        a.pos_orig[o] = x;
        a.vel_orig[o] = v;
        vm = __float_as_uint(dot(v, v));
        particle_energy(a.scenes[r.scene].c, x, v, ke, pe);
    }
    block_max<TILE>(vm, a.tile_vmax + r.tile);
    block_energy<TILE>(ke, pe, a.tile_energy + r.tile);
}

// Per-scene totals from the scene's tile partials: max |v|^2 and, when the
// final kernel computed them, the FP64 energies. One workgroup of
// TOTALS_THREADS per scene; thread i takes tiles i, i + TOTALS_THREADS, ...
// of the scene in order, then a fixed tree over the waves. The order depends
// only on the scene's own tiles, so the result is the same alone and in any
// batch.
constexpr int TOTALS_THREADS = 1024;
__global__ __launch_bounds__(TOTALS_THREADS) void k_scene_totals(Batch a, int energy) {
    constexpr int WAVES = TOTALS_THREADS / 64;
    __shared__ double ske[WAVES], spe[WAVES];
    __shared__ unsigned int svm[WAVES];
    for (uint32_t s = blockIdx.x; s < a.nscenes; s += gridDim.x) {
        SceneDev const& sd = a.scenes[s];
        double ke = 0.0, pe = 0.0;
        unsigned int vm = 0;
        for (uint32_t k = threadIdx.x; k < sd.ntiles; k += TOTALS_THREADS) {
            unsigned int const m = a.tile_vmax[sd.first_tile + k];
            vm = m > vm ? m : vm;
            if (energy) {
                double2 const e = a.tile_energy[sd.first_tile + k];
                ke += e.x;
                pe += e.y;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            ke += __shfl_xor(ke, off);
            pe += __shfl_xor(pe, off);
            unsigned int const o = __shfl_xor(vm, off);
            vm = o > vm ? o : vm;
        }
        if (__lane_id() == 0) {
            ske[threadIdx.x / 64] = ke;
            spe[threadIdx.x / 64] = pe;
            svm[threadIdx.x / 64] = vm;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned int m = svm[0];
            for (int w = 1; w < WAVES; ++w) { m = svm[w] > m ? svm[w] : m; }
            a.stats[s].vmax2_bits = m;
            if (energy) {
                double k2[WAVES / 2], p2[WAVES / 2];
                for (int w = 0; w < WAVES / 2; ++w) {
                    k2[w] = ske[2 * w] + ske[2 * w + 1];
                    p2[w] = spe[2 * w] + spe[2 * w + 1];
                }
                for (int h = WAVES / 4; h >= 1; h >>= 1) {
                    for (int w = 0; w < h; ++w) {
                        k2[w] = k2[2 * w] + k2[2 * w + 1];
                        p2[w] = p2[2 * w] + p2[2 * w + 1];
                    }
                }
                a.stats[s].kinetic_energy = k2[0];
                a.stats[s].potential_energy = p2[0];
                if (sd.frame.on) {
                    // End of the sub-step: commit the frame loop's progress.
                    FrameDev& fr = const_cast<SceneDev&>(sd).frame;
                    fr.remaining = fr.next_remaining;
                    fr.time = fr.next_time;
                }
            }
        }
        __syncthreads();
    }
}

//------------------------------------------------------------------------------
// max |v|^2 of each scene over its original-order velocities (cfl.cpp:12-21);
// the tiles cover original indices with the same scene ranges.
__global__ __launch_bounds__(TILE) void k_vmax2(Batch a) {
    TileRef const r = tile_ref(a, TILE);
    unsigned int vm = 0;
    if (r.valid) { vm = __float_as_uint(dot(a.vel_orig[r.t], a.vel_orig[r.t])); }
    block_max<TILE>(vm, a.tile_vmax + r.tile);
}

__global__ __launch_bounds__(256) void k_scatter_f32(float const* src, uint32_t const* orig,
                                                     float* dst, uint32_t n) {
    uint32_t const t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) { dst[orig[t]] = src[t]; }
}
__global__ __launch_bounds__(256) void k_scatter_v2(V2 const* src, uint32_t src_stride,
                                                    uint32_t const* orig, V2* dst, uint32_t n) {
    uint32_t const t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) { dst[orig[t]] = src[static_cast<size_t>(t) * src_stride]; }
}

__global__ __launch_bounds__(256) void k_export_lists(Lists l, uint32_t n, uint32_t stride,
                                                      uint32_t const* orig_self,
                                                      uint32_t const* orig_other,
                                                      uint32_t const* other_scene_of,
                                                      SceneDev const* scenes, int solid_other,
                                                      uint8_t* counts, uint32_t* idx,
                                                      float* dist, V2* vec) {
    uint32_t const t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) { return; }
    uint32_t const o = orig_self[t];
    int const cnt = l.count[t];
    counts[o] = static_cast<uint8_t>(cnt);
    for (int k = 0; k < MAXN; ++k) {
        size_t const src = static_cast<size_t>(k) * stride + t;
        size_t const dst = static_cast<size_t>(o) * MAXN + k;
        if (k < cnt) {
            uint32_t const og = orig_other[l.idx[src]];
            SceneDev const& sc = scenes[other_scene_of[og]];
            idx[dst] = og - (solid_other ? sc.soffset : sc.offset);
            if (dist) { dist[dst] = l.dist[src]; }
            if (vec) { vec[dst] = l.vec[src]; }
        } else {
            idx[dst] = 0;
            if (dist) { dist[dst] = 0.0f; }
            if (vec) { vec[dst] = V2{0.0f, 0.0f}; }
        }
    }
}

//------------------------------------------------------------------------------
static inline dim3 tile_grid(Batch const& b) { return dim3(b.ntiles ? b.ntiles : 1); }
// Over the tiles of `at` (its total) when at.n > 0, else every tile.
static inline dim3 active_grid(Batch const& b, ActiveTiles const& at) {
    return at.n ? dim3(at.start[at.n] ? at.start[at.n] : 1) : tile_grid(b);
}

// Kernel instantiation for the batch depth Batch::nb (2, 4 or 8).
#define NB_DISPATCH(nb, KERNEL, ...)                                                  \
    do {                                                                              \
        if ((nb) <= 2) {                                                              \
            hipLaunchKernelGGL((KERNEL<2>), __VA_ARGS__);                             \
        } else if ((nb) <= 4) {                                                       \
            hipLaunchKernelGGL((KERNEL<4>), __VA_ARGS__);                             \
        } else {                                                                      \
            hipLaunchKernelGGL((KERNEL<8>), __VA_ARGS__);                             \
        }                                                                             \
    } while (0)

void launch_iisph_aii_beta0(Batch const& a, hipStream_t s) {
    NB_DISPATCH(a.nb, k_iisph_aii_beta0, tile_grid(a), dim3(TILE), 0, s, a);
}
void launch_iisph_sdp(Batch const& a, ActiveTiles const& at, hipStream_t s) {
    NB_DISPATCH(a.nb, k_iisph_sdp, active_grid(a, at), dim3(TILE), 0, s, a, at);
}
void launch_iisph_beta_p(Batch const& a, ActiveTiles const& at, hipStream_t s) {
    NB_DISPATCH(a.nb, k_iisph_beta_p, active_grid(a, at), dim3(TILE), 0, s, a, at);
}

#define AP_DISPATCH(KERNEL, PSEUDO)                                                    \
    do {                                                                              \
        if (a.nb <= 2) {                                                              \
            hipLaunchKernelGGL((KERNEL<2, PSEUDO>), tile_grid(a), dim3(TILE), 0, s, a);   \
        } else if (a.nb <= 4) {                                                       \
            hipLaunchKernelGGL((KERNEL<4, PSEUDO>), tile_grid(a), dim3(TILE), 0, s, a);   \
        } else {                                                                      \
            hipLaunchKernelGGL((KERNEL<8, PSEUDO>), tile_grid(a), dim3(TILE), 0, s, a);   \
        }                                                                             \
    } while (0)

// As AP_DISPATCH, over `grid` with the ActiveTiles `at` as the second argument.
#define AP_DISPATCH_AT(KERNEL, PSEUDO)                                                 \
    do {                                                                              \
        if (a.nb <= 2) {                                                              \
            hipLaunchKernelGGL((KERNEL<2, PSEUDO>), grid, dim3(TILE), 0, s, a, at);       \
        } else if (a.nb <= 4) {                                                       \
            hipLaunchKernelGGL((KERNEL<4, PSEUDO>), grid, dim3(TILE), 0, s, a, at);       \
        } else {                                                                      \
            hipLaunchKernelGGL((KERNEL<8, PSEUDO>), grid, dim3(TILE), 0, s, a, at);       \
        }                                                                             \
    } while (0)

void launch_ap_diag0(Batch const& a, int pseudo, hipStream_t s) {
    if (pseudo) {
        AP_DISPATCH(k_ap_diag0, 1);
    } else {
        AP_DISPATCH(k_ap_diag0, 0);
    }
}
void launch_ap_accel(Batch const& a, ActiveTiles const& at, hipStream_t s) {
    NB_DISPATCH(a.nb, k_ap_accel, active_grid(a, at), dim3(TILE), 0, s, a, at);
}
void launch_ap_iterate(Batch const& a, int pseudo, ActiveTiles const& at, hipStream_t s) {
    dim3 const grid = active_grid(a, at);
    if (pseudo) {
        AP_DISPATCH_AT(k_ap_iterate, 1);
    } else {
        AP_DISPATCH_AT(k_ap_iterate, 0);
    }
}
void launch_ap_final(Batch const& a, int pseudo, hipStream_t s) {
    if (pseudo) {
        AP_DISPATCH(k_ap_final, 1);
    } else {
        AP_DISPATCH(k_ap_final, 0);
    }
}

template <int G, int F>
void launch_apply_nb(Batch const& a, int all_scenes, ActiveTiles const& at, hipStream_t s) {
    dim3 const grid = active_grid(a, at);
    if (a.nb <= 2) {
        hipLaunchKernelGGL((k_apply<G, F, 2>), grid, dim3(TILE), 0, s, a, all_scenes, at);
    } else if (a.nb <= 4) {
        hipLaunchKernelGGL((k_apply<G, F, 4>), grid, dim3(TILE), 0, s, a, all_scenes, at);
    } else {
        hipLaunchKernelGGL((k_apply<G, F, 8>), grid, dim3(TILE), 0, s, a, all_scenes, at);
    }
}

void launch_apply(Batch const& a, int with_gravity, int finalize, int all_scenes,
                  hipStream_t s, ActiveTiles const& at) {
    ActiveTiles const every{};
    ActiveTiles const& use = all_scenes ? every : at;  // all scenes: every tile
    if (with_gravity && finalize) {
        launch_apply_nb<1, 1>(a, all_scenes, use, s);
    } else if (with_gravity) {
        launch_apply_nb<1, 0>(a, all_scenes, use, s);
    } else if (finalize) {
        launch_apply_nb<0, 1>(a, all_scenes, use, s);
    } else {
        launch_apply_nb<0, 0>(a, all_scenes, use, s);
    }
}
void launch_finalize(Batch const& a, int dfsph, hipStream_t s) {
    if (dfsph) {
        hipLaunchKernelGGL(k_finalize<1>, tile_grid(a), dim3(TILE), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_finalize<0>, tile_grid(a), dim3(TILE), 0, s, a);
    }
}
void launch_dfsph_div(Batch const& a, ActiveTiles const& at, hipStream_t s) {
    NB_DISPATCH(a.nb, k_dfsph_div, active_grid(a, at), dim3(TILE), 0, s, a, at);
}
void launch_dfsph_dens(Batch const& a, ActiveTiles const& at, hipStream_t s) {
    NB_DISPATCH(a.nb, k_dfsph_dens, active_grid(a, at), dim3(TILE), 0, s, a, at);
}
void launch_pci_init(Batch const& a, hipStream_t s) {
    hipLaunchKernelGGL(k_pci_init, tile_grid(a), dim3(TILE), 0, s, a);
}
void launch_pci_density(Batch const& a, ActiveTiles const& at, hipStream_t s) {
    NB_DISPATCH(a.nb, k_pci_density, active_grid(a, at), dim3(TILE), 0, s, a, at);
}
void launch_pci_force(Batch const& a, ActiveTiles const& at, hipStream_t s) {
    NB_DISPATCH(a.nb, k_pci_force, active_grid(a, at), dim3(TILE), 0, s, a, at);
}
void launch_pci_finalize(Batch const& a, hipStream_t s) {
    hipLaunchKernelGGL(k_pci_finalize, tile_grid(a), dim3(TILE), 0, s, a);
}
// scene_of[i] = scene of slot i, from the tile table (one workgroup per tile).
__global__ __launch_bounds__(TILE) void k_scene_of(Tile const* tiles, uint32_t* scene_of) {
    Tile const tl = tiles[blockIdx.x];
    if (threadIdx.x < tl.count) { scene_of[tl.start + threadIdx.x] = tl.scene; }
}

void launch_scene_of(Tile const* tiles, uint32_t ntiles, uint32_t* scene_of, hipStream_t s) {
    if (!ntiles) { return; }
    hipLaunchKernelGGL(k_scene_of, dim3(ntiles), dim3(TILE), 0, s, tiles, scene_of);
}

// Moves the state of repacked scenes: for each segment {src, dst, n}, slots
// [src, src + n) of the four state arrays go to [dst, dst + n) of the scratch
// arrays (segments in increasing dst, contiguous from seg[0].dst).
__global__ __launch_bounds__(256) void k_repack(uint3 const* seg, uint32_t nseg, uint32_t total,
                                                V2 const* pos, V2 const* vel, float const* kv,
                                                float const* kd, V2* pos_o, V2* vel_o,
                                                float* kv_o, float* kd_o) {
    uint32_t const base = seg[0].y;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += gridDim.x * blockDim.x) {
        uint32_t const d = base + i;
        uint32_t lo = 0, hi = nseg;  // last segment with dst <= d
        while (hi - lo > 1) {
            uint32_t const mid = (lo + hi) / 2;
            if (seg[mid].y <= d) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        uint32_t const src = seg[lo].x + (d - seg[lo].y);
        pos_o[d] = pos[src];
        vel_o[d] = vel[src];
        kv_o[d] = kv[src];
        kd_o[d] = kd[src];
    }
}

void launch_repack(uint3 const* seg, uint32_t nseg, uint32_t total, V2 const* pos, V2 const* vel,
                   float const* kv, float const* kd, V2* pos_o, V2* vel_o, float* kv_o,
                   float* kd_o, hipStream_t s) {
    if (!total || !nseg) { return; }
    uint32_t blocks = (total + 255) / 256;
    if (blocks > 8192) { blocks = 8192; }
    hipLaunchKernelGGL(k_repack, dim3(blocks), dim3(256), 0, s, seg, nseg, total, pos, vel, kv,
                       kd, pos_o, vel_o, kv_o, kd_o);
}

void launch_scene_totals(Batch const& a, int energy, hipStream_t s) {
    uint32_t const blocks = a.nscenes < 4096 ? a.nscenes : 4096;
    hipLaunchKernelGGL(k_scene_totals, dim3(blocks ? blocks : 1), dim3(TOTALS_THREADS), 0, s, a,
                       energy);
}
void launch_vmax2(Batch const& a, hipStream_t s) {
    hipLaunchKernelGGL(k_vmax2, tile_grid(a), dim3(TILE), 0, s, a);
    launch_scene_totals(a, 0, s);
}
void launch_scatter_f32(float const* src, uint32_t const* orig, float* dst, uint32_t n,
                        hipStream_t s) {
    if (!n) { return; }
    hipLaunchKernelGGL(k_scatter_f32, dim3((n + 255) / 256), dim3(256), 0, s, src, orig, dst, n);
}
void launch_scatter_v2(V2 const* src, uint32_t src_stride, uint32_t const* orig, V2* dst,
                       uint32_t n, hipStream_t s) {
    if (!n) { return; }
    hipLaunchKernelGGL(k_scatter_v2, dim3((n + 255) / 256), dim3(256), 0, s, src, src_stride,
                       orig, dst, n);
}
void launch_export_lists(Lists const& l, uint32_t n, uint32_t stride, uint32_t const* orig_self,
                         uint32_t const* orig_other, uint32_t const* other_scene_of,
                         SceneDev const* scenes, int solid_other, uint8_t* counts, uint32_t* idx,
                         float* dist, V2* vec, hipStream_t s) {
    if (!n) { return; }
    hipLaunchKernelGGL(k_export_lists, dim3((n + 255) / 256), dim3(256), 0, s, l, n, stride,
                       orig_self, orig_other, other_scene_of, scenes, solid_other, counts, idx,
                       dist, vec);
}

}  // namespace emsph
