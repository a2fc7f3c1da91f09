// Host orchestration of the 2D SPH sub-step on one gfx950 GPU, for a batch of
// M >= 1 scenes, and the C ABI of include/emerald_sph.h.
//
// One sub-step (the reference's <method>_sub_step, e.g. iisph.cpp:23-151):
//   1. keys      grid coords + 64-bit Morton keys, in original order
//   2. sort      stable LSD radix sort of (key, index) over the significant key
//                digits, then over the scene id (scenes stay contiguous)
//   3. cells     run heads, exclusive scan -> block indices, cell ranges in
//                sorted (scene, key) order; gather of the state into sorted
//                order; per cell, the slot ranges of its 3x3 neighbour cells
//                (fluid and solid), found by searching the sorted cell lists
//   4. build     64-capped neighbour lists in reference order (fluid and solid),
//                with the first per-particle sums of the method fused in
//   5. solver    method sweeps over all scenes at once; after each iteration
//                every scene applies the reference's exit test on the device
//                and drops out of later sweeps when it passes; the host only
//                reads how many scenes are still iterating
//   6. finalize  x += dt v, state scattered back to original order
// Solids are static: their cells, lists and volumes are built once at
// configuration (solids.cpp:76-137) with the same kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/emerald_sph.h"
#include "kernels.hpp"

using namespace emsph;

namespace {

thread_local std::string g_last_error;

emsph_status fail(emsph_status s, std::string const& msg) {
    g_last_error = msg;
    return s;
}

}  // namespace

// Error reporting for the library's host-only translation units.
emsph_status emsph_set_error(emsph_status s, char const* msg) { return fail(s, msg); }
bool emsph_snapshot_read_file(char const* path, std::vector<unsigned char>& out);

namespace {

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t const e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                            \
            return fail(EMSPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                  \
    } while (0)

#define ST_TRY(expr)                       \
    do {                                   \
        emsph_status const s_ = (expr);    \
        if (s_ != EMSPH_OK) { return s_; } \
    } while (0)

constexpr int64_t kFlicksPerSecond = 705600000;
constexpr float kPi = 3.14159274101257324218750f;

double to_seconds(int64_t f) {
    return static_cast<double>(f) / static_cast<double>(kFlicksPerSecond);
}

KernelConsts make_kernel_consts(float h) {
    KernelConsts c;
    c.h = h;
    c.hh = h * h;
    c.hhh = h * h * h;
    c.k = h * h * h * kPi;
    c.four_h = 4.0f * h;
    // W(0, h) = _wa(0) / k (kernels.h:25-35)
    float const q = 0.0f / h;
    c.w0 = (1.0f + q * q * (q * 0.75f - 1.5f)) / c.k;
    // Correctly rounded double reciprocals (the host divide is IEEE).
    c.inv_h = 1.0 / static_cast<double>(c.h);
    c.inv_k = 1.0 / static_cast<double>(c.k);
    c.inv_hh = 1.0 / static_cast<double>(c.hh);
    c.inv_hhh = 1.0 / static_cast<double>(c.hhh);
    c.inv_four_h = 1.0 / static_cast<double>(c.four_h);
    c.r_fast_hi = (h >= 0x1p-20f && h <= 0x1p10f) ? 4.0f * h : -1.0f;
    return c;
}

// StepConsts::cut2 of a cut-off (host sqrtf is the correctly rounded IEEE
// square root, as on the device). For cut >= 1e-18 every |v|^2 below 2 FLT_MIN
// has Imath length < 2.2e-19 < cut and is also below cut2 (>= 1e-36).
float cut_l2(float cut) {
    if (!(cut >= 1e-18f) || !std::isfinite(cut)) { return 0.0f; }
    float x = cut * cut;
    while (x > 0.0f && std::sqrt(x) >= cut) { x = std::nextafter(x, 0.0f); }
    while (std::sqrt(x) < cut) { x = std::nextafter(x, std::numeric_limits<float>::infinity()); }
    return x;
}

StepConsts step_consts(emsph_config const& cfg, float dt) {
    StepConsts k{};
    emsph_parameters const& p = cfg.params;
    k.kc = make_kernel_consts(p.support);
    k.cell = p.support * 2.0f;
    k.cut2 = cut_l2(k.cell);
    k.rho0 = p.target_density;
    k.fluid_volume = cfg.mass_per_particle / p.target_density;
    k.mass = cfg.mass_per_particle;
    k.gravity = p.gravity;
    k.dt = dt;
    k.omega = p.iisph_omega;
    k.pci_denom = cfg.pressure_correction_denom;
    k.viscosity = p.viscosity;
    // Skin lists (SkinView): candidates within cell + h, rounded up with a
    // margin above the float |v|^2 test's rounding (3 roundings of 2^-24);
    // reuse while |x - x_rebuild|^2 stays below (h/2)^2, rounded down with a
    // margin for the rounding of the displacement and of this test. Then a
    // point outside the candidates (exactly farther than cell + h at the
    // rebuild) stays farther than cell + 2^-16 cell, where the reference's
    // |v|^2 < cut2 fails.
    double const rs = static_cast<double>(k.cell) + static_cast<double>(p.support);
    k.skin_r2 = std::nextafter(static_cast<float>(rs * rs * (1.0 + 0x1p-18)), HUGE_VALF);
    double const hl = 0.5 * static_cast<double>(p.support);
    k.skin_lim2 = std::nextafter(static_cast<float>(hl * hl * (1.0 - 0x1p-12)), 0.0f);
    return k;
}

Thresholds thresholds(emsph_parameters const& p) {
    Thresholds t{};
    t.iisph_avg = p.iisph_error_average_threshold;
    t.iisph_max = p.iisph_error_max_threshold;
    t.div_avg = p.dfsph_divergence_error_average_threshold;
    t.div_max = p.dfsph_divergence_error_max_threshold;
    t.dens_avg = p.dfsph_density_error_average_threshold;
    t.dens_max = p.dfsph_density_error_max_threshold;
    t.iisph_max_iters = std::min(p.iisph_max_pressure_iterations, MAX_RED_SLOTS);
    t.dfsph_max_iters = std::min(p.dfsph_max_correction_iterations, MAX_RED_SLOTS / 2);
    return t;
}

uint32_t round_up(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }

// Device bytes held by the library (all DBufs of all engines).
std::atomic<int64_t> g_device_bytes{0};

// A device allocation that frees itself.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(DBuf const&) = delete;
    DBuf& operator=(DBuf const&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            (void)hipFree(p);
            g_device_bytes -= static_cast<int64_t>((n + TAIL_PAD) * sizeof(T));
        }
        p = nullptr;
        n = 0;
    }
    // Every buffer carries TAIL_PAD readable elements past its count: the
    // neighbour build loads candidate positions 16 bytes at a time, up to
    // three elements past the end of a cell's slot range.
    static constexpr size_t TAIL_PAD = 4;
    hipError_t alloc(size_t count) {
        release();
        if (count == 0) { count = 1; }
        hipError_t const e =
          hipMalloc(reinterpret_cast<void**>(&p), (count + TAIL_PAD) * sizeof(T));
        if (e == hipSuccess) {
            n = count;
            g_device_bytes += static_cast<int64_t>((n + TAIL_PAD) * sizeof(T));
        } else {
            p = nullptr;
        }
        return e;
    }
    hipError_t ensure(size_t count) { return (p && n >= count) ? hipSuccess : alloc(count); }
    void swap(DBuf& o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
};

// Keys, sort and cells of one point set (fluid or solid).
struct SpatialIndex {
    uint32_t cap = 0;
    DBuf<uint64_t> keys0, keys1;
    DBuf<uint32_t> vals0, vals1;
    DBuf<uint32_t> counts, scan_scratch;
    DBuf<uint32_t> heads, head_scan, block_index, cell_begin, cell_end, cell_tag;
    DBuf<uint64_t> cell_key;
    DBuf<uint32_t> ncells;  // device cell count of the last indexing run
    DBuf<uint4> set;        // static cell set (solids), see CellSet
    uint32_t set_mask = 0;
    std::vector<uint64_t> set_keys;  // the set's cell keys and scene tags (host copy)
    std::vector<uint32_t> set_tags;
    DBuf<uint32_t> set_bits;         // the set's cells as bits over a dense layout's bins
    uint64_t set_bits_for = 0;       // that layout's generation (0: none)
    DBuf<uint32_t> orig, rank;
    DBuf<uint2> win;        // [cell][9] slot ranges of the 3x3 neighbour cells
    DBuf<int2> grid;
    uint64_t* sorted_keys = nullptr;
    uint32_t* sorted_vals = nullptr;
    int key_passes = 0;
    int scene_passes = 0;
    // Dense cell index (see DenseScene): one table of dense_bins bins in which
    // scene s owns 4^dense_layout[s] bins (dense_host); the counts are zero
    // between runs. dense_gen numbers the layouts.
    DBuf<uint32_t> dcount, doffset, dcell, drank, dscratch;
    DBuf<DenseScene> dscene;
    std::vector<int> dense_layout;
    std::vector<DenseScene> dense_host;
    uint64_t dense_bins = 0;
    size_t dense_cap = 0;  // bins the tables hold
    uint64_t dense_gen = 0;
    uint64_t direct_kmax = 0;  // key maximum of the last run that read the range back
    DBuf<uint64_t> scene_kmax;  // each scene's largest key of the last run

    // Tables for per-scene bits (allocated outside any graph capture).
    // *changed: a new layout (graphs captured with the old one are stale).
    hipError_t set_dense(std::vector<int> const& bits, uint32_t c, hipStream_t s,
                         bool* changed) {
        *changed = false;
        if (dense_bins && bits == dense_layout) { return hipSuccess; }
        std::vector<DenseScene> ds(bits.size());
        uint64_t base = 0;
        for (size_t k = 0; k < bits.size(); ++k) {
            uint64_t const limit = bits[k] > 0 ? (uint64_t{1} << (2 * bits[k])) : 0;
            ds[k] = DenseScene{limit, static_cast<uint32_t>(base), 0u};
            base += limit;
        }
        dense_bins = 0;
        dense_layout.clear();
        if (base == 0 || base + 2 > (uint64_t{1} << 32)) { return hipSuccess; }
        hipError_t e;
        size_t const need = static_cast<size_t>(base) + 2;
        if (need > dense_cap || !drank.p || drank.n < c) {
            dense_cap = 0;
            if ((e = dcount.alloc(need)) || (e = doffset.alloc(need)) || (e = dcell.alloc(need)) ||
                (e = drank.alloc(c)) || (e = dscratch.alloc(2 * scan_scratch_words(need) + 64))) {
                return e;
            }
            if ((e = hipMemsetAsync(dcount.p, 0, need * sizeof(uint32_t), s))) { return e; }
            dense_cap = need;
        }
        if (dscene.n < ds.size() && (e = dscene.alloc(ds.size()))) { return e; }
        if ((e = hipMemcpyAsync(dscene.p, ds.data(), ds.size() * sizeof(DenseScene),
                                hipMemcpyHostToDevice, s))) {
            return e;
        }
        if ((e = hipStreamSynchronize(s))) { return e; }
        dense_host = std::move(ds);
        dense_layout = bits;
        dense_bins = base;
        ++dense_gen;
        *changed = true;
        return hipSuccess;
    }

    // Keeps the arrays when they already hold c points.
    hipError_t ensure(uint32_t c) { return (keys0.p && cap >= c) ? hipSuccess : alloc(c, false); }

    hipError_t alloc(uint32_t c, bool with_grid) {
        cap = c;
        hipError_t e;
        if ((e = keys0.alloc(c)) || (e = keys1.alloc(c)) || (e = vals0.alloc(c)) ||
            (e = vals1.alloc(c)) || (e = counts.alloc(radix_counts_words(c) + 256)) ||
            (e = scan_scratch.alloc(scan_scratch_words(radix_counts_words(c) + c) + 64)) ||
            (e = heads.alloc(c)) || (e = head_scan.alloc(c)) || (e = block_index.alloc(c)) ||
            (e = cell_begin.alloc(c)) || (e = cell_end.alloc(c)) ||
            (e = cell_key.alloc(c)) || (e = cell_tag.alloc(c)) ||
            (e = ncells.alloc(1)) || (e = orig.alloc(c)) || (e = rank.alloc(c)) ||
            (e = win.alloc(static_cast<size_t>(c) * 9))) {
            return e;
        }
        if (with_grid) { return grid.alloc(c); }
        return hipSuccess;
    }

    CellList cells() const {
        return CellList{cell_key.p, cell_tag.p, cell_begin.p, cell_end.p, ncells.p};
    }
};

// Event pairs recorded around each timed launch, resolved only when the pool
// is full or the totals are read, so timing adds no host synchronisation.
// Solid-solid lists of the solid build (for the volumes only).
struct SolidScratch {
    DBuf<uint8_t> count;
    DBuf<uint32_t> idx;
    DBuf<float> dist;
    DBuf<V2> vec;
    void release() {
        count.release();
        idx.release();
        dist.release();
        vec.release();
    }
};

struct KernelTimer {
    static constexpr size_t POOL = 8192;
    bool enabled = false;
    std::map<std::string, std::pair<double, int64_t>> acc;
    std::vector<hipEvent_t> ev;        // 2 per record
    std::vector<char const*> names;    // string literals
    size_t used = 0;
};

// Pinned read-back block.
struct Pinned {
    KeyState keys[2];  // double-buffered read-backs of the solver loops
    KeyState direct;   // synchronous read-back of the key range
};

struct SceneHost {
    emsph_config cfg{};
    uint32_t cap = 0;     // fluid capacity reserved for the scene
    uint32_t n = 0;       // fluid particles
    uint32_t ns = 0;      // solid particles
    uint32_t offset = 0;  // first fluid slot
    uint32_t soffset = 0; // first solid slot
    float dt = 0.0f;      // sub-step dt (the largest one when cfl > 0)
    float cfl = 0.0f;     // > 0: per-sub-step dt from the scene's CFL number
    bool force_set = false;     // force model set explicitly (else gravity of the params)
    emsph_force_model force{};
    bool retired = false;       // a pool scene removed, not yet committed
};

// The scene's force model: the one set explicitly, or the reference's
// default_gravity_forces with the parameters' gravity.
emsph_force_model force_model(SceneHost const& h) {
    if (h.force_set) { return h.force; }
    emsph_force_model m{};
    m.flags = EMSPH_FORCE_GRAVITY;
    m.gravity = h.cfg.params.gravity;
    return m;
}

int scene_pass_count(uint32_t m) {
    if (m <= 1) { return 0; }
    int p = 0;
    uint32_t v = m - 1;
    while (v) {
        ++p;
        v >>= 8;
    }
    return p;
}

}  // namespace

// One batch of scenes resident on the device.
struct Engine {
    hipStream_t stream = nullptr;
    int device = 0;
    std::vector<SceneHost> sh;
    uint32_t M = 0;
    uint32_t cap = 0;      // total fluid capacity
    uint32_t stride = 0;
    uint32_t N = 0;        // total fluid particles (scenes packed at their offsets)
    uint32_t scap = 0;
    uint32_t NS = 0;
    bool store_geometry = false;
    bool have_lists = false;
    int last_method = -1;

    DBuf<SceneDev> scenes;
    DBuf<Tile> tiles;
    uint32_t ntiles = 0;
    DBuf<uint32_t> scene_of;   // original fluid index -> scene
    DBuf<SceneStats> stats;
    std::vector<SceneStats> host_stats;
    DBuf<Reduction> red;
    DBuf<KeyState> keystate;
    DBuf<unsigned long long> key_partials;

    DBuf<V2> pos, vel;
    DBuf<float> kv, kd;
    SpatialIndex fidx;
    DBuf<V2> spos, svel, dii, vext, xstar, xstar1, vstar0, vstar1, force;
    DBuf<V2> pr0, pr1;  // IISPH {p, rho} ping-pong
    DBuf<V2> pacc;      // IISPH_AP / IISPH_PSEUDO_AP pressure accelerations
    DBuf<SdDii> sd;     // IISPH {sum_j d_ij p_j, d_ii}
    DBuf<float> rho, rho_star, aii, p0, p1, alpha, kvs, kds;
    float* cur_p = nullptr;
    DBuf<uint8_t> fcount, scount;
    DBuf<uint32_t> fl_idx, sl_idx;
    DBuf<V2> fl_gw, sl_gw, fl_vec, sl_vec;
    DBuf<float> fl_dist, sl_dist;
    DBuf<uint2> swin;       // [fluid cell][9] solid slot ranges
    SolidScratch solid_scratch;
    bool keep_solid_scratch = false;
    DBuf<uint3> repack_segments;
    DBuf<uint2> overflow;   // particles deferred by the capped neighbour build
    DBuf<double2> tile_energy;
    DBuf<Reduction> tile_red;
    DBuf<unsigned int> tile_vmax;
    int build_cap = 32;     // LDS list capacity of the next build (from the last max count)

    SpatialIndex sidx;
    DBuf<uint32_t> s_scene_of;
    DBuf<Tile> stiles;
    uint32_t nstiles = 0;
    DBuf<V2> s_pos, s_vel, s_pos_orig;
    DBuf<float> s_vol;

    Pinned* host = nullptr;
    int passes_hint = -1;
    // Dense cell index of the fluid: one scene's table bits for the next
    // sub-step (0: radix sort); a batch's per-scene bits, taken after each
    // exact run (dense_stale); whether the current sub-step's layout is ready
    // (dense_on); a sticky switch-off after a crowded cell.
    int dense_bits = 0;
    std::vector<int> dense_scene_bits;
    bool dense_scene_ok = false;
    bool dense_stale = true;
    bool dense_on = false;
    bool dense_disabled = false;
    // The indexing choices of the current sub-step attempt (pass count, dense
    // bits), fixed when it begins: enqueue, graph key and read-back agree.
    int prep_passes = -1;
    int prep_dense = 0;
    bool keys_checked = false;
    bool sort_short = false;
    KernelTimer timer;
    // Solver iterations launched by the last attempt of the current sub-step
    // (the maximum over the batch's scenes), and running totals.
    int last_iterations = 0;
    int last_div_iterations = 0;
    // sub-steps, pressure/density iterations, divergence iterations, sub-steps
    // redone after a short sort
    int64_t totals[4] = {0, 0, 0, 0};

    hipEvent_t read_ev[2] = {nullptr, nullptr};
    KeyState keys{};  // the last key state read back
    int loop_expect[4] = {64, 64, 64, 64};  // iterations each solver loop needed last time
    bool loop_known = false;  // keys holds a read-back of the running solver loop
    // Captured graphs of the fixed segments of a sub-step, by what their launch
    // arguments depend on; layout_gen changes whenever buffers or sizes do.
    bool use_graphs = true;
    int nb = 4;              // neighbour batch depth of the sweeps
    bool xcd_remap = true;   // XCD-aware workgroup -> tile order
    uint64_t layout_gen = 1;
    std::map<std::array<uint64_t, 10>, hipGraphExec_t> graphs;
    // External forces: a host callback (contexts only) and the per-slot force
    // buffer used whenever some scene's model is not the default gravity.
    emsph_user_forces_fn user_fn = nullptr;
    void* user_data = nullptr;
    bool use_fext = false;
    DBuf<V2> fext;
    // Skin candidate lists (SkinView): enabled by the caller; valid after a
    // rebuild sub-step built them without overflow; a reuse sub-step takes its
    // lists from them. skin_rho: sqrt of the largest displacement ratio seen
    // by the last reuse (0 at a rebuild), skin_rho_prev the one before, from
    // which the next sub-step's ratio is extrapolated.
    bool skin_on = false;
    bool skin_valid = false;
    bool skin_reuse = false;    // this attempt reuses the candidates
    bool skin_built = false;    // this attempt builds them
    float skin_rho = 0.0f, skin_rho_prev = 0.0f;
    int64_t skin_counts[2] = {0, 0};  // sub-steps that rebuilt / reused
    int skin_run = 0;        // reuses since the last candidate build
    int skin_pause = 0;      // sub-steps left without candidates (motion too fast to pay)
    DBuf<uint32_t> skin_cf, skin_cs;
    DBuf<uint8_t> skin_nf, skin_ns;
    DBuf<V2> skin_xref;
    DBuf<unsigned int> skin_tmax;
    float step_time = 0.0f;  // time of the current sub-step (for the callback)
    bool frame_on = false;   // a device-side frame loop is running (its time is on the device)
    // State::colors of a context: computed at the end of each frame step (or
    // on request), RGBA in original order; a host User_colors_function
    // replaces the default target-density colors.
    emsph_user_colors_fn colors_fn = nullptr;
    void* colors_data = nullptr;
    DBuf<float4> colors;
    DBuf<uchar4> colors8;
    bool colors_valid = false;
    struct {
        std::vector<V2> spos, svel, fs;
        std::vector<uint32_t> orig;
        std::vector<float> pos, vel, forces;
    } fhost;

    ~Engine() {
        // Work still queued may copy into the pinned block or run the graphs:
        // drain the stream before anything it uses is released.
        if (stream) { (void)hipStreamSynchronize(stream); }
        if (host) { (void)hipHostFree(host); }
        for (hipEvent_t e : read_ev) {
            if (e) { (void)hipEventDestroy(e); }
        }
        for (auto& kv : graphs) { (void)hipGraphExecDestroy(kv.second); }
        for (hipEvent_t e : timer.ev) { (void)hipEventDestroy(e); }
        if (stream) { (void)hipStreamDestroy(stream); }
    }
};

struct emsph_context {
    Engine e;
    emsph_config cfg{};
};

struct emsph_batch {
    Engine e;
};

struct emsph_search {
    Engine e;
    float cell = 0.0f;
    bool ran = false;
};

// A scene of a pool: its parameters and static solids (kept on the host for
// the solid rebuild of each commit) and, until the commit that uploads it, its
// initial state.
struct PoolScene {
    int id = 0;
    emsph_config cfg{};
    std::vector<float> solids;
    std::vector<float> pos, vel;
    uint32_t n = 0;
    uint32_t ns = 0;
};

// A batch whose scenes come and go (the served form).
struct emsph_pool {
    Engine e;
    int method = 1;
    uint32_t fluid_cap = 0;
    uint32_t solid_cap = 0;
    uint32_t max_scenes = 0;
    float cfl = 0.0f;
    int next_id = 1;
    std::vector<PoolScene> live;     // in slot order
    std::vector<PoolScene> adds;     // staged
    std::vector<int> removes;        // staged
};

namespace {

hipError_t timer_flush(Engine* g) {
    KernelTimer& tm = g->timer;
    if (!tm.used) { return hipSuccess; }
    hipError_t e = hipEventSynchronize(tm.ev[2 * tm.used - 1]);
    if (e) { return e; }
    for (size_t k = 0; k < tm.used; ++k) {
        float ms = 0.0f;
        if ((e = hipEventElapsedTime(&ms, tm.ev[2 * k], tm.ev[2 * k + 1]))) { return e; }
        auto& slot = tm.acc[tm.names[k]];
        slot.first += ms;
        slot.second += 1;
    }
    tm.used = 0;
    return hipSuccess;
}

template <class F>
hipError_t timed(Engine* g, char const* name, F&& launch) {
    if (!g->timer.enabled) {
        launch();
        return hipGetLastError();
    }
    KernelTimer& tm = g->timer;
    if (tm.used == KernelTimer::POOL) {
        hipError_t const e = timer_flush(g);
        if (e) { return e; }
    }
    size_t const k = tm.used;
    while (tm.ev.size() < 2 * (k + 1)) {
        hipEvent_t e = nullptr;
        hipError_t const r = hipEventCreate(&e);
        if (r) { return r; }
        tm.ev.push_back(e);
    }
    if (tm.names.size() <= k) { tm.names.resize(k + 1); }
    tm.names[k] = name;
    hipError_t e = hipEventRecord(tm.ev[2 * k], g->stream);
    if (e) { return e; }
    launch();
    if ((e = hipGetLastError())) { return e; }
    if ((e = hipEventRecord(tm.ev[2 * k + 1], g->stream))) { return e; }
    tm.used = k + 1;
    return hipSuccess;
}

// The build stage under kernel timing: two records, "build" and
// "build_deferred", whose events the two dispatches stamp themselves
// (hipExtLaunchKernelGGL), so the intervals are the kernels' own durations.
template <class F>
hipError_t timed_build(Engine* g, F&& launch) {
    if (!g->timer.enabled) {
        launch(nullptr);
        return hipGetLastError();
    }
    KernelTimer& tm = g->timer;
    if (tm.used + 2 > KernelTimer::POOL) {
        hipError_t const e = timer_flush(g);
        if (e) { return e; }
    }
    size_t const k = tm.used;
    while (tm.ev.size() < 2 * (k + 2)) {
        hipEvent_t e = nullptr;
        hipError_t const r = hipEventCreate(&e);
        if (r) { return r; }
        tm.ev.push_back(e);
    }
    if (tm.names.size() < k + 2) { tm.names.resize(k + 2); }
    tm.names[k] = "build";
    tm.names[k + 1] = "build_deferred";
    launch(&tm.ev[2 * k]);
    hipError_t const e = hipGetLastError();
    if (e) { return e; }
    tm.used = k + 2;
    return hipSuccess;
}

emsph_status engine_init(Engine* g) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        return fail(EMSPH_ERR_NO_DEVICE, "no HIP device");
    }
    HIP_TRY(hipGetDevice(&g->device));
    HIP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g->host), sizeof(Pinned)));
    HIP_TRY(hipEventCreateWithFlags(&g->read_ev[0], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&g->read_ev[1], hipEventDisableTiming));
    HIP_TRY(g->keystate.alloc(1));
    HIP_TRY(g->key_partials.alloc(2 * KEY_PARTS));
    return EMSPH_OK;
}

Batch fluid_batch(Engine* g, float const* p_in = nullptr, float* p_out = nullptr) {
    Batch b{};
    b.scenes = g->scenes.p;
    b.tiles = g->tiles.p;
    b.ntiles = g->ntiles;
    b.nscenes = g->M;
    b.stride = g->stride;
    b.stats = g->stats.p;
    b.red = g->red.p;
    b.keystate = g->keystate.p;
    b.pos = g->spos.p;
    b.vel = g->svel.p;
    b.orig = g->fidx.orig.p;
    b.rank = g->fidx.rank.p;
    b.cell_of = g->fidx.block_index.p;
    b.win = g->fidx.win.p;
    b.swin = g->swin.p;
    b.overflow = g->overflow.p;
    b.tile_energy = g->tile_energy.p;
    b.tile_red = g->tile_red.p;
    b.tile_vmax = g->tile_vmax.p;
    if (g->NS) {
        b.solids = SolidView{g->s_pos.p, g->s_vel.p, g->s_vol.p, g->sidx.rank.p, g->sidx.orig.p};
    }
    b.fl = Lists{g->fcount.p, g->fl_idx.p, g->fl_gw.p, g->fl_dist.p, g->fl_vec.p};
    b.sl = Lists{g->scount.p, g->sl_idx.p, g->sl_gw.p, g->sl_dist.p, g->sl_vec.p};
    b.rho = g->rho.p;
    b.rho_star = g->rho_star.p;
    b.aii = g->aii.p;
    b.dii = g->dii.p;
    b.p_in = p_in;
    b.p_out = p_out;
    b.sd = g->sd.p;
    b.pacc = g->pacc.p;
    b.alpha = g->alpha.p;
    b.pos_orig = g->pos.p;
    b.vel_orig = g->vel.p;
    b.kv_orig = g->kv.p;
    b.kd_orig = g->kd.p;
    b.kvs = g->kvs.p;
    b.kds = g->kds.p;
    b.vext = g->vext.p;
    b.xstar = g->xstar.p;
    b.force = g->force.p;
    b.fext = g->use_fext ? g->fext.p : nullptr;
    b.store_geometry = g->store_geometry ? 1 : 0;
    b.nb = g->nb;
    b.xcd_remap = g->xcd_remap ? 1 : 0;
    if (g->skin_on) {
        b.skin = SkinView{g->skin_reuse ? 1 : 0, g->skin_cf.p, g->skin_nf.p, g->skin_cs.p,
                          g->skin_ns.p, g->skin_xref.p, g->skin_tmax.p};
    }
    return b;
}

// Device scene entry of one host scene whose tiles start at first_tile.
SceneDev scene_dev(SceneHost const& h, uint32_t first_tile) {
    SceneDev sd{};
    sd.c = step_consts(h.cfg, h.dt);
    sd.thr = thresholds(h.cfg.params);
    sd.offset = h.offset;
    sd.n = h.n;
    sd.soffset = h.soffset;
    sd.ns = h.ns;
    sd.first_tile = first_tile;
    sd.ntiles = (h.n + TILE - 1) / TILE;
    sd.cfl = h.cfl;
    sd.dt_max = h.dt;
    emsph_force_model const fm = force_model(h);
    sd.c.gravity = fm.gravity;
    sd.force.flags = fm.flags;
    sd.force.pole_mag = fm.pole_magnitude;
    sd.force.pole_x = fm.pole_x;
    sd.force.pole_y = fm.pole_y;
    sd.force.drag_mult = -fm.drag_magnitude * fm.drag_diameter;
    sd.force.anti_maxd = fm.anti_coupling_max_distance;
    sd.force.anti_mag = fm.anti_coupling_magnitude;
    sd.retired = h.retired ? 1u : 0u;
    return sd;
}

// Scene table and tiles from the host scene list; fluid slots are packed at
// each scene's offset (offsets are prefix sums of the scene capacities).
emsph_status upload_layout(Engine* g) {
    ++g->layout_gen;
    std::vector<SceneDev> sd(g->M);
    std::vector<Tile> tl;
    for (uint32_t s = 0; s < g->M; ++s) {
        SceneHost const& h = g->sh[s];
        sd[s] = scene_dev(h, static_cast<uint32_t>(tl.size()));
        for (uint32_t k = 0; k < h.n; k += TILE) {
            tl.push_back(Tile{s, h.offset + k, std::min<uint32_t>(TILE, h.n - k), 0});
        }
    }
    g->ntiles = static_cast<uint32_t>(tl.size());
    HIP_TRY(g->scenes.ensure(std::max<uint32_t>(g->M, 1)));
    HIP_TRY(g->tiles.ensure(std::max<size_t>(tl.size(), 1)));
    HIP_TRY(g->tile_energy.ensure(std::max<size_t>(tl.size(), 1)));
    HIP_TRY(g->tile_red.ensure(std::max<size_t>(tl.size(), 1)));
    HIP_TRY(g->tile_vmax.ensure(std::max<size_t>(tl.size(), 1)));
    HIP_TRY(g->scene_of.ensure(std::max<uint32_t>(g->cap, 1)));
    HIP_TRY(g->stats.ensure(std::max<uint32_t>(g->M, 1)));
    HIP_TRY(g->red.ensure(static_cast<size_t>(std::max<uint32_t>(g->M, 1)) * MAX_RED_SLOTS));
    HIP_TRY(hipMemcpyAsync(g->scenes.p, sd.data(), sd.size() * sizeof(SceneDev),
                           hipMemcpyHostToDevice, g->stream));
    if (!tl.empty()) {
        HIP_TRY(hipMemcpyAsync(g->tiles.p, tl.data(), tl.size() * sizeof(Tile),
                               hipMemcpyHostToDevice, g->stream));
    }
    launch_scene_of(g->tiles.p, g->ntiles, g->scene_of.p, g->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->host_stats.assign(g->M, SceneStats{});
    // N is the packed extent: every slot up to the last scene's end. Scenes
    // with n < cap leave holes; holes are outside every tile, and their keys
    // are never produced because the key/sort stage runs over packed slots
    // only when every scene is full (the batch), or over one scene (a context).
    return EMSPH_OK;
}

hipError_t alloc_fluid(Engine* g, uint32_t cap) {
    g->cap = cap;
    g->stride = round_up(std::max<uint32_t>(cap, 1), 256);
    size_t const L = static_cast<size_t>(MAXN) * g->stride;
    hipError_t e;
    if ((e = g->pos.alloc(cap)) || (e = g->vel.alloc(cap)) || (e = g->kv.alloc(cap)) ||
        (e = g->kd.alloc(cap)) || (e = g->fidx.alloc(cap, false)) || (e = g->spos.alloc(cap)) ||
        (e = g->svel.alloc(cap)) || (e = g->dii.alloc(cap)) ||
        (e = g->vext.alloc(cap)) || (e = g->xstar.alloc(cap)) || (e = g->xstar1.alloc(cap)) ||
        (e = g->vstar0.alloc(cap)) ||
        (e = g->vstar1.alloc(cap)) || (e = g->force.alloc(cap)) || (e = g->rho.alloc(cap)) ||
        (e = g->rho_star.alloc(cap)) || (e = g->aii.alloc(cap)) || (e = g->p0.alloc(cap)) ||
        (e = g->p1.alloc(cap)) || (e = g->alpha.alloc(cap)) || (e = g->kvs.alloc(cap)) ||
        (e = g->kds.alloc(cap)) || (e = g->pr0.alloc(cap)) || (e = g->pr1.alloc(cap)) ||
        (e = g->pacc.alloc(cap)) ||
        (e = g->sd.alloc(cap)) || (e = g->fcount.alloc(g->stride)) ||
        (e = g->scount.alloc(g->stride)) || (e = g->fl_idx.alloc(L)) ||
        (e = g->sl_idx.alloc(L)) || (e = g->fl_gw.alloc(L)) || (e = g->sl_gw.alloc(L)) ||
        (e = g->swin.alloc(static_cast<size_t>(cap) * 9)) || (e = g->overflow.alloc(cap))) {
        return e;
    }
    if ((e = hipMemsetAsync(g->kv.p, 0, cap * sizeof(float), g->stream))) { return e; }
    if ((e = hipMemsetAsync(g->kd.p, 0, cap * sizeof(float), g->stream))) { return e; }
    if ((e = hipMemsetAsync(g->vel.p, 0, cap * sizeof(V2), g->stream))) { return e; }
    return hipSuccess;
}

hipError_t ensure_geometry(Engine* g) {
    hipError_t e = hipSuccess;
    size_t const m = static_cast<size_t>(MAXN) * g->stride;
    if (!g->fl_dist.p) { ++g->layout_gen; }
    if (!g->fl_dist.p && (e = g->fl_dist.alloc(m))) { return e; }
    if (!g->fl_vec.p && (e = g->fl_vec.alloc(m))) { return e; }
    if (!g->sl_dist.p && (e = g->sl_dist.alloc(m))) { return e; }
    if (!g->sl_vec.p && (e = g->sl_vec.alloc(m))) { return e; }
    return e;
}

// Keys -> sort -> cells of n points (original order) with per-point scenes.
//
// passes_hint < 0 reads the key range back before sorting (one host sync) and
// sorts exactly the significant key digits. passes_hint >= 0 sorts that many
// digits without waiting; k_cell_heads then flags (KeyState::sort_short) a key
// range those digits do not cover, and the caller redoes the work with the
// exact count before anything it produced is used.
//
// The cells stage ends with the cell windows of the index against itself and,
// when other is given, against other's cells into owin.
hipError_t index_points(Engine* g, SpatialIndex& si, V2 const* pos, uint32_t n,
                        uint32_t const* scene_of, uint32_t nscenes, int passes_hint,
                        IndexArgs gathers, bool time_it, SpatialIndex const* other = nullptr,
                        uint2* owin = nullptr, bool use_dense = false,
                        WindowArgs* wout = nullptr) {
    hipStream_t const s = g->stream;
    auto run = [&](char const* name, auto&& f) -> hipError_t {
        if (time_it) { return timed(g, name, f); }
        f();
        return hipGetLastError();
    };
    hipError_t e;
    IndexArgs a = gathers;
    a.pos = pos;
    a.n = n;
    a.scenes = g->scenes.p;
    a.scene_of = scene_of;
    a.keys = si.keys0.p;
    a.vals = si.vals0.p;
    a.grid = si.grid.p;
    a.keystate = g->keystate.p;
    a.key_partials = g->key_partials.p;
    a.scene_kmax = si.scene_kmax.n >= nscenes ? si.scene_kmax.p : nullptr;
    bool const dense = use_dense && passes_hint >= 0 && si.dense_bins > 0 &&
                       si.dense_layout.size() == nscenes;
    if (dense) {
        a.dense_bins = si.dense_bins;
        a.dscene = si.dscene.p;
        a.dcount = si.dcount.p;
        a.dcell = si.dcell.p;
        a.doffset = si.doffset.p;
        a.drank = si.drank.p;
    }
    if ((e = run("keys", [&] { launch_keys(a, s); }))) { return e; }
    if (dense) {
        // Counting sort over the dense table: each bin's first slot and cell
        // ordinal (the counts are cleared for the next run; the cell count
        // goes where the head scan of the radix path puts it), scatter, then
        // the cells stage below puts each cell in original index order.
        si.key_passes = 0;
        si.scene_passes = 0;
        si.sorted_keys = si.keys1.p;
        si.sorted_vals = si.vals1.p;
        a.sorted_keys = si.sorted_keys;
        a.sorted_vals = si.sorted_vals;
        uint32_t const bins = static_cast<uint32_t>(a.dense_bins) + 1u;
        if ((e = run("dense_sort", [&] {
                 launch_cell_scan_clear(si.dcount.p, si.doffset.p, si.dcell.p, bins,
                                        si.dscratch.p, si.doffset.p + bins, si.ncells.p, s);
                 launch_dense_scatter(a, s);
             }))) {
            return e;
        }
    } else if (passes_hint < 0) {
        if ((e = hipMemcpyAsync(&g->host->direct, g->keystate.p, sizeof(KeyState),
                                hipMemcpyDeviceToHost, s))) {
            return e;
        }
        if ((e = hipStreamSynchronize(s))) { return e; }
        si.key_passes = radix_pass_count(g->host->direct.kmin, g->host->direct.kmax);
        si.direct_kmax = g->host->direct.kmax;
    } else {
        si.key_passes = passes_hint;
    }
    if (!dense) {
        si.scene_passes = scene_pass_count(nscenes);
        int which = 0;
        if ((e = run("radix_sort", [&] {
                 which = launch_radix_sort(si.keys0.p, si.vals0.p, si.keys1.p, si.vals1.p, n,
                                           si.key_passes, scene_of, si.scene_passes, si.counts.p,
                                           si.scan_scratch.p, s);
             }))) {
            return e;
        }
        si.sorted_keys = which ? si.keys1.p : si.keys0.p;
        si.sorted_vals = which ? si.vals1.p : si.vals0.p;
    }
    a.key_passes = si.key_passes;
    a.check_keys = (passes_hint >= 0 && !dense) ? 1 : 0;
    a.sorted_keys = si.sorted_keys;
    a.sorted_vals = si.sorted_vals;
    a.heads = si.heads.p;
    a.head_scan = si.head_scan.p;
    a.block_index = si.block_index.p;
    a.cell_begin = si.cell_begin.p;
    a.cell_end = si.cell_end.p;
    a.cell_key = si.cell_key.p;
    a.cell_tag = si.cell_tag.p;
    a.orig = si.orig.p;
    a.rank = si.rank.p;
    WindowArgs w{};
    w.self = si.cells();
    if (other && other->set.p) {
        w.other = other->cells();
        w.other_set = CellSet{other->set.p, other->set_mask};
    }
    w.max_cells = n;
    if (dense) {
        w.dense_bins = a.dense_bins;
        w.dscene = a.dscene;
        w.doffset = si.doffset.p;
        if (other && other->set_bits_for == si.dense_gen) { w.other_bits = other->set_bits.p; }
    }
    w.win = si.win.p;
    w.owin = owin;
    w.report = &g->keystate.p->cells;
    if (wout) { *wout = w; }
    return run("cells", [&] {
        if (!dense) {  // the dense path's cell scan gave the ordinals and the count
            launch_cell_heads(a, s);
            launch_exclusive_scan(si.heads.p, si.head_scan.p, n, si.scan_scratch.p, si.ncells.p,
                                  s);
        }
        launch_cells_finalize(a, s);
        launch_cell_windows(w, s);
    });
}

// The static hash set of an index's cells (built on the host from the cell
// list; load factor at most 1/4).
emsph_status build_cell_set(Engine* g, SpatialIndex& si) {
    uint32_t nc = 0;
    HIP_TRY(hipMemcpyAsync(&nc, si.ncells.p, sizeof(nc), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    std::vector<uint64_t> keys(nc);
    std::vector<uint32_t> tags(nc);
    if (nc) {
        HIP_TRY(hipMemcpyAsync(keys.data(), si.cell_key.p, nc * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipMemcpyAsync(tags.data(), si.cell_tag.p, nc * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    uint32_t cap = 64;
    while (cap < 4ull * nc) { cap <<= 1; }
    std::vector<uint4> slots(cap, make_uint4(0u, 0u, 0u, EMPTY_SLOT));
    for (uint32_t c = 0; c < nc; ++c) {
        uint32_t s = cell_set_hash(keys[c], tags[c]) & (cap - 1);
        while (slots[s].w != EMPTY_SLOT) { s = (s + 1) & (cap - 1); }
        slots[s] = make_uint4(static_cast<uint32_t>(keys[c]), static_cast<uint32_t>(keys[c] >> 32),
                              tags[c], c);
    }
    HIP_TRY(si.set.ensure(cap));
    HIP_TRY(hipMemcpyAsync(si.set.p, slots.data(), cap * sizeof(uint4), hipMemcpyHostToDevice,
                           g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    si.set_mask = cap - 1;
    si.set_keys = std::move(keys);
    si.set_tags = std::move(tags);
    si.set_bits_for = 0;  // the bitmap (if any) is rebuilt for the new set
    return EMSPH_OK;
}

// Bitmap of the set's cells over the bins of the dense layout of index d (for
// the dense windows; outside any graph capture).
emsph_status build_set_bits(Engine* g, SpatialIndex& si, SpatialIndex const& d) {
    if (si.set_bits_for == d.dense_gen && si.set_bits.p) { return EMSPH_OK; }
    std::vector<uint32_t> words(static_cast<size_t>((d.dense_bins + 31) / 32), 0u);
    for (size_t c = 0; c < si.set_keys.size(); ++c) {
        uint32_t const tag = si.set_tags[c];
        if (tag >= d.dense_host.size()) { continue; }
        uint64_t const bin = dense_bin(d.dense_host[tag], si.set_keys[c], d.dense_bins);
        if (bin < d.dense_bins) { words[bin >> 5] |= 1u << (bin & 31u); }
    }
    HIP_TRY(si.set_bits.alloc(words.size()));
    HIP_TRY(hipMemcpyAsync(si.set_bits.p, words.data(), words.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    si.set_bits_for = d.dense_gen;
    return EMSPH_OK;
}

// Solids of every scene: index, solid-solid lists with geometry, volumes
// (solids.cpp:76-137).
emsph_status build_solids(Engine* g, std::vector<float const*> const& sxy) {
    g->skin_valid = false;
    g->NS = 0;
    for (uint32_t s = 0; s < g->M; ++s) {
        g->sh[s].soffset = g->NS;
        g->NS += g->sh[s].ns;
    }
    if (!g->NS) { return EMSPH_OK; }
    uint32_t const NS = g->NS;
    g->scap = NS;
    uint32_t const sstride = round_up(NS, 256);
    size_t const SL = static_cast<size_t>(MAXN) * sstride;
    // Solid-solid lists, only for the volumes: scratch that pools keep between
    // commits (grow-only) and other engines release after the build.
    SolidScratch& sc = g->solid_scratch;
    hipError_t e;
    if ((e = g->sidx.ensure(NS)) || (e = g->s_pos.ensure(NS)) || (e = g->s_vel.ensure(NS)) ||
        (e = g->s_pos_orig.ensure(NS)) || (e = g->s_vol.ensure(NS)) ||
        (e = g->s_scene_of.ensure(NS)) || (e = sc.count.ensure(sstride)) ||
        (e = sc.idx.ensure(SL)) || (e = sc.dist.ensure(SL)) || (e = sc.vec.ensure(SL))) {
        return fail(EMSPH_ERR_OUT_OF_MEMORY, "solid allocation");
    }
    std::vector<uint32_t> so(NS);
    std::vector<float> staged(2 * static_cast<size_t>(NS));
    std::vector<Tile> tl;
    for (uint32_t s = 0; s < g->M; ++s) {
        SceneHost const& h = g->sh[s];
        if (h.ns) {
            std::memcpy(staged.data() + 2 * static_cast<size_t>(h.soffset), sxy[s],
                        h.ns * sizeof(V2));
        }
        for (uint32_t k = 0; k < h.ns; ++k) { so[h.soffset + k] = s; }
        for (uint32_t k = 0; k < h.ns; k += TILE) {
            tl.push_back(Tile{s, h.soffset + k, std::min<uint32_t>(TILE, h.ns - k), 0});
        }
    }
    HIP_TRY(hipMemcpyAsync(g->s_pos_orig.p, staged.data(), NS * sizeof(V2),
                           hipMemcpyHostToDevice, g->stream));
    g->nstiles = static_cast<uint32_t>(tl.size());
    HIP_TRY(g->stiles.ensure(tl.size()));
    HIP_TRY(hipMemcpyAsync(g->stiles.p, tl.data(), tl.size() * sizeof(Tile),
                           hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemcpyAsync(g->s_scene_of.p, so.data(), NS * sizeof(uint32_t),
                           hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipMemsetAsync(g->s_vel.p, 0, NS * sizeof(V2), g->stream));
    IndexArgs ga{};
    ga.pos_out = g->s_pos.p;
    HIP_TRY(index_points(g, g->sidx, g->s_pos_orig.p, NS, g->s_scene_of.p, g->M, -1, ga, false));
    ST_TRY(build_cell_set(g, g->sidx));
    Batch b{};
    b.scenes = g->scenes.p;
    b.tiles = g->stiles.p;
    b.ntiles = g->nstiles;
    b.nscenes = g->M;
    b.stride = sstride;
    b.pos = g->s_pos.p;
    b.orig = g->sidx.orig.p;
    b.rank = g->sidx.rank.p;
    b.cell_of = g->sidx.block_index.p;
    b.win = g->sidx.win.p;
    b.keystate = g->keystate.p;
    b.fl = Lists{sc.count.p, sc.idx.p, nullptr, sc.dist.p, sc.vec.p};
    b.sl = Lists{};  // solid-solid lists only
    b.rho = g->s_vol.p;
    b.store_geometry = 1;
    launch_build(BUILD_LISTS, MAXN, b, g->stream);
    launch_solid_volumes(b, NS, g->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (!g->keep_solid_scratch) { sc.release(); }
    return EMSPH_OK;
}

// Configure the engine with M scenes; fluid capacity per scene, solids per scene.
emsph_status engine_configure(Engine* g, std::vector<SceneHost> scenes,
                              std::vector<float const*> const& solid_xy) {
    g->skin_valid = false;
    g->sh = std::move(scenes);
    g->M = static_cast<uint32_t>(g->sh.size());
    uint32_t off = 0;
    for (auto& h : g->sh) {
        h.offset = off;
        off += h.cap;
    }
    hipError_t const e = alloc_fluid(g, std::max<uint32_t>(off, 1));
    if (e) {
        return fail(e == hipErrorOutOfMemory ? EMSPH_ERR_OUT_OF_MEMORY : EMSPH_ERR_HIP,
                    std::string("device allocation: ") + hipGetErrorString(e));
    }
    g->N = off;
    ST_TRY(upload_layout(g));
    ST_TRY(build_solids(g, solid_xy));
    ST_TRY(upload_layout(g));  // solid offsets are now known
    g->passes_hint = -1;
    return EMSPH_OK;
}

//------------------------------------------------------------------------------
// KeyState read-backs.
// The smallest LDS list capacity that held the last build's longest list with
// a margin of two; or 16 while the lists above 16 entries are at most 1/64 of
// the particles: those are deferred to the full-capacity pass, which costs
// less than halving the occupancy of the whole build. (A list longer than the
// capacity is always deferred, never lost.)
int pick_build_cap(KeyState const& k, uint32_t n) {
    if (k.max_count + 2u <= 16u || static_cast<uint64_t>(k.long_lists) * 64u <= n) { return 16; }
    for (int c : BUILD_CAPS) {
        if (k.max_count + 2u <= static_cast<unsigned int>(c)) { return c; }
    }
    return MAXN;
}

// Propagates a HIP error out of a function or lambda returning hipError_t.
#define HIP_PASS(x)                     \
    do {                                \
        hipError_t const e_ = (x);      \
        if (e_ != hipSuccess) { return e_; } \
    } while (0)

// Table bits B of a dense cell index holding keys up to kmax (4^B > kmax),
// or 0 when the table would exceed 4^DENSE_MAX_BITS bins or 8 n + 2^16.
int dense_bits_for(uint64_t kmax, uint32_t n) {
    int bits = 1;
    while (bits <= DENSE_MAX_BITS && (kmax >> (2 * bits)) != 0) { ++bits; }
    if (bits > DENSE_MAX_BITS) { return 0; }
    uint64_t const bins = uint64_t{1} << (2 * bits);
    return bins <= 8ull * n + 65536ull ? bits : 0;
}

// Whether the next indexing run uses the dense cell index (1) or the radix
// sort (0); dense_prepare readied its layout at the start of the sub-step.
int dense_for(Engine const* g) {
    bool const ok = g->passes_hint >= 0 && !g->dense_disabled && g->dense_on &&
                    g->fidx.dense_bins > 0 && g->fidx.dense_layout.size() == g->M;
    return ok ? 1 : 0;
}

// Table bits of each scene's dense index. One scene follows the key range of
// its last run. In a batch each scene's table covers its largest key of the
// last exact run and the key of its solids' largest cell coordinates (the
// walls bound the fluid; a Morton key grows with either coordinate), taken
// again only after an exact run. False: no dense index (some scene's table
// would be too large).
bool dense_want(Engine* g, std::vector<int>& bits) {
    if (g->M == 1) {
        if (g->dense_bits <= 0) { return false; }
        bits.assign(1, g->dense_bits);
        return true;
    }
    if (g->dense_stale) {
        std::vector<uint64_t> km(g->M, 0);
        if (hipMemcpyAsync(km.data(), g->fidx.scene_kmax.p, g->M * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
            hipStreamSynchronize(g->stream) != hipSuccess) {
            (void)hipGetLastError();
            g->dense_disabled = true;
            return false;
        }
        // The solids' bound: the key of their largest non-negative cell
        // coordinates (cells at negative coordinates, outside every table,
        // have keys above it and are looked up in the solid cell set).
        SpatialIndex const& si = g->sidx;
        std::vector<uint32_t> mx(g->M, 0), my(g->M, 0);
        std::vector<bool> any(g->M, false);
        for (size_t c = 0; c < si.set_keys.size(); ++c) {
            uint32_t const tag = si.set_tags[c];
            uint32_t const gx = compact_bits_2(si.set_keys[c]);
            uint32_t const gy = compact_bits_2(si.set_keys[c] >> 1);
            if (tag >= g->M || gx >= 0x80000000u || gy >= 0x80000000u) { continue; }
            mx[tag] = std::max(mx[tag], gx);
            my[tag] = std::max(my[tag], gy);
            any[tag] = true;
        }
        for (uint32_t s = 0; s < g->M; ++s) {
            if (any[s]) {
                uint64_t const bound =
                  morton2(static_cast<int32_t>(mx[s]), static_cast<int32_t>(my[s]));
                km[s] = std::max(km[s], bound);
            }
        }
        g->dense_scene_bits.assign(g->M, 0);
        g->dense_scene_ok = true;
        for (uint32_t s = 0; s < g->M; ++s) {
            if (g->sh[s].n == 0) { continue; }  // no points: no table
            int const b = dense_bits_for(km[s], g->sh[s].n);
            if (!b) { g->dense_scene_ok = false; }
            g->dense_scene_bits[s] = b;
        }
        g->dense_stale = false;
    }
    if (!g->dense_scene_ok) { return false; }
    bits = g->dense_scene_bits;
    return true;
}

// The dense layout of the next sub-step, readied outside any graph capture
// (a scene's table only grows); sets dense_on.
void dense_prepare(Engine* g) {
    g->dense_on = false;
    if (g->dense_disabled || g->N == 0) { return; }
    SpatialIndex& f = g->fidx;
    if (g->M > 1 && f.scene_kmax.n < g->M) {
        if (f.scene_kmax.alloc(g->M) != hipSuccess) {
            (void)hipGetLastError();
            g->dense_disabled = true;
            return;
        }
        g->dense_stale = true;
    }
    if (g->passes_hint < 0) { return; }  // an exact run: radix sort
    std::vector<int> bits;
    if (!dense_want(g, bits)) { return; }
    if (f.dense_layout.size() == bits.size()) {
        for (size_t k = 0; k < bits.size(); ++k) { bits[k] = std::max(bits[k], f.dense_layout[k]); }
    }
    bool changed = false;
    if (f.set_dense(bits, g->cap, g->stream, &changed) != hipSuccess) {
        (void)hipGetLastError();
        g->dense_disabled = true;
        return;
    }
    if (!f.dense_bins) { return; }
    if (changed) {
        ++g->layout_gen;
        if (g->NS && build_set_bits(g, g->sidx, f) != EMSPH_OK) {
            (void)hipGetLastError();
            g->dense_disabled = true;
            return;
        }
    }
    g->dense_on = true;
}

// Key-state read-backs into pinned slot k: enqueue, then wait and take it in.
hipError_t read_async(Engine* g, int k) {
    hipError_t e = hipMemcpyAsync(&g->host->keys[k], g->keystate.p, sizeof(KeyState),
                                  hipMemcpyDeviceToHost, g->stream);
    if (e) { return e; }
    return hipEventRecord(g->read_ev[k], g->stream);
}

hipError_t read_wait(Engine* g, int k) {
    hipError_t const e = hipEventSynchronize(g->read_ev[k]);
    if (e) { return e; }
    g->keys = g->host->keys[k];
    g->build_cap = pick_build_cap(g->keys, g->N);
    if (!g->keys_checked) {
        g->keys_checked = true;
        g->sort_short = g->keys.sort_short != 0u;
        g->passes_hint = radix_pass_count(g->keys.kmin, g->keys.kmax);
        g->dense_bits = dense_bits_for(g->keys.kmax, g->N);
        if (g->keys.dense_crowded) { g->dense_disabled = true; }
    }
    return hipSuccess;
}

hipError_t read_keystate(Engine* g) {
    hipError_t const e = read_async(g, 0);
    return e ? e : read_wait(g, 0);
}

// A solver loop: launch(it) enqueues iteration it, whose kernels skip the
// scenes that have converged. From iteration first_check on (and at the last
// one) the key state is read back one iteration late: iteration it + 1 is
// enqueued before the host waits for the read-back of iteration it, so the
// device does not idle while the host decides; an iteration enqueued after
// every scene converged does nothing. Iterations at or beyond the count the
// same loop needed in the previous sub-step (kind indexes that memory) are not
// enqueued ahead: there the loop most likely ends, and an empty iteration
// costs more than the wait. *done receives the iterations until the last
// scene converged; *short_sort is set when the sub-step's sort was short (the
// caller then redoes the sub-step).
template <class F>
hipError_t run_loop(Engine* g, int kind, int max_it, int first_check, F&& launch, int* done,
                    bool* short_sort) {
    *done = max_it;
    *short_sort = false;
    g->loop_known = false;
    int const expect = g->loop_expect[kind];
    int pending = -1;  // iteration whose read-back is in flight
    for (int it = 0; it < max_it; ++it) {
        if (pending >= 0 && it >= expect) {
            hipError_t e = read_wait(g, pending & 1);
            if (e) { return e; }
            g->loop_known = true;
            if (g->sort_short || g->keys.active_count == 0) {
                *short_sort = g->sort_short;
                *done = pending + 1;
                g->loop_expect[kind] = *done;
                return hipSuccess;
            }
            pending = -1;
        }
        hipError_t e = launch(it);
        if (e) { return e; }
        bool const check = it >= first_check || it + 1 == max_it;
        if (check && (e = read_async(g, it & 1))) { return e; }
        if (pending >= 0) {
            if ((e = read_wait(g, pending & 1))) { return e; }
            g->loop_known = true;
            if (g->sort_short || g->keys.active_count == 0) {
                if (check && (e = read_wait(g, it & 1))) { return e; }  // drain the slot
                *short_sort = g->sort_short;
                *done = pending + 1;
                g->loop_expect[kind] = *done;
                return hipSuccess;
            }
        }
        pending = check ? it : -1;
    }
    if (pending >= 0) {
        hipError_t const e = read_wait(g, pending & 1);
        if (e) { return e; }
        *short_sort = g->sort_short;
    }
    g->loop_expect[kind] = *done;
    return hipSuccess;
}

// The tiles of the scenes still iterating by the latest read-back of this
// loop (g->keys.active_bits, a superset of those iterating now), when that
// read-back exists, the engine holds several scenes and at most
// ACTIVE_TILES_MAX of them are left; else n = 0 (every tile).
ActiveTiles active_tiles(Engine const* g) {
    ActiveTiles at{};
    if (!g->loop_known || g->M < 2 || g->M > ACTIVE_BIT_SCENES ||
        g->keys.active_count > static_cast<unsigned int>(ACTIVE_TILES_MAX)) {
        return at;
    }
    uint32_t first = 0;
    uint32_t total = 0;
    uint32_t n = 0;
    for (uint32_t s = 0; s < g->M; ++s) {
        uint32_t const nt = (g->sh[s].n + TILE - 1) / TILE;
        if ((g->keys.active_bits[s >> 5] >> (s & 31u)) & 1u) {
            if (n == static_cast<uint32_t>(ACTIVE_TILES_MAX)) { return ActiveTiles{}; }
            at.first[n] = first;
            at.start[n] = total;
            total += nt;
            ++n;
        }
        first += nt;
    }
    if (n == 0 || total == 0) { return ActiveTiles{}; }
    at.n = n;
    at.start[n] = total;
    return at;
}

// The first iteration of a solver loop whose exit test can stop some scene:
// the method's first exit check, or an earlier last iteration (the scene's
// maximum). Before it the exit tests only keep every scene iterating, so they
// are not launched.
int first_decision(Engine const* g, int kind, int first_check) {
    int least_max = first_check + 1;
    for (auto const& h : g->sh) {
        emsph_parameters const& p = h.cfg.params;
        int const mx = kind == 0   ? std::min(p.iisph_max_pressure_iterations, MAX_RED_SLOTS)
                       : kind == 3 ? 6
                                   : std::min(p.dfsph_max_correction_iterations, MAX_RED_SLOTS / 2);
        least_max = std::min(least_max, mx);
    }
    return std::min(first_check, least_max - 1);
}

// Host-side state at the start of a sub-step attempt.
void begin_attempt(Engine* g) {
    g->prep_passes = g->passes_hint;
    g->prep_dense = dense_for(g);
    // A reuse of the skin candidates has its short flag (a particle moved too
    // far) checked like a sort's; a rebuild builds them when this attempt's
    // index is the dense one (the same test as index_points').
    SpatialIndex const& si = g->fidx;
    g->skin_built = g->skin_on && g->skin_pause == 0 && !g->skin_reuse && g->prep_dense &&
                    g->prep_passes >= 0 && si.dense_bins > 0 && si.dense_layout.size() == g->M;
    g->keys_checked = g->passes_hint < 0 && !g->skin_reuse;
    g->sort_short = false;
    g->last_iterations = 0;
    g->last_div_iterations = 0;
}

// Steps 1-4 of the sub-step. With a known sort pass count this only enqueues
// work (it can be captured into a graph); otherwise the key range is read back
// first.
hipError_t enqueue_prepare(Engine* g, int build_mode, bool dfsph) {
    if (g->skin_reuse) {
        // The lists from the skin candidates: no keys, sort or cells; the
        // state gathered into the last rebuild's order.
        Batch b = fluid_batch(g);
        auto run = [&](char const* name, auto&& f) -> hipError_t {
            if (g->timer.enabled) { return timed(g, name, f); }
            f();
            return hipGetLastError();
        };
        HIP_PASS(run("skin_gather", [&] {
            launch_skin_begin(b, g->stream);
            launch_skin_gather(b, dfsph ? g->kv.p : nullptr, dfsph ? g->kd.p : nullptr,
                               g->stream);
        }));
        return timed_build(g, [&](hipEvent_t const* ev) {
            launch_build(build_mode, g->build_cap, b, g->stream, ev);
        });
    }
    IndexArgs ga{};
    ga.vel_in = g->vel.p;
    ga.pos_out = g->spos.p;
    ga.vel_out = g->svel.p;
    if (dfsph) {
        ga.kv_in = g->kv.p;
        ga.kd_in = g->kd.p;
        ga.kv_out = g->kvs.p;
        ga.kd_out = g->kds.p;
    }
    Batch b = fluid_batch(g);
    ga.step_scenes = g->scenes.p;
    ga.step_stats = g->stats.p;
    ga.step_nscenes = g->M;
    WindowArgs w{};
    HIP_PASS(index_points(g, g->fidx, g->pos.p, g->N, g->scene_of.p, g->M, g->prep_passes, ga,
                          g->timer.enabled, g->NS ? &g->sidx : nullptr, g->swin.p,
                          g->prep_dense, &w));
    HIP_PASS(timed_build(g, [&](hipEvent_t const* ev) {
        launch_build(build_mode, g->build_cap, b, g->stream, ev);
    }));
    if (g->skin_built && w.dense_bins && w.dscene) {
        if (g->timer.enabled) {
            return timed(g, "skin_candidates", [&] { launch_skin_candidates(b, w, g->stream); });
        }
        launch_skin_candidates(b, w, g->stream);
        return hipGetLastError();
    }
    g->skin_built = false;  // (begin_attempt's test and index_points' agree)
    return hipSuccess;
}

// Host-side fields of the fluid index that a sub-step's sort leaves (also
// after a graph replay, which runs none of the host code).
void finish_prepare(Engine* g) {
    if (g->skin_reuse) {  // no indexing run: the last rebuild's index stays
        g->have_lists = true;
        return;
    }
    SpatialIndex& si = g->fidx;
    if (g->prep_passes < 0) {
        // index_points read the key range back and sorted exactly with the
        // pass count it left in si.key_passes.
        if (g->passes_hint < 0) { g->passes_hint = si.key_passes; }
        g->dense_bits = dense_bits_for(si.direct_kmax, g->N);
        g->dense_stale = true;  // a batch takes its scenes' key maxima again
    } else {
        si.key_passes = g->prep_passes;
    }
    if (g->prep_dense) {
        si.key_passes = 0;
        si.scene_passes = 0;
        si.sorted_keys = si.keys1.p;
        si.sorted_vals = si.vals1.p;
    } else {
        si.scene_passes = scene_pass_count(g->M);
        bool const odd = ((si.key_passes + si.scene_passes) & 1) != 0;
        si.sorted_keys = odd ? si.keys1.p : si.keys0.p;
        si.sorted_vals = odd ? si.vals1.p : si.vals0.p;
    }
    g->have_lists = true;
}

// Runs one fixed segment of the sub-step (enqueue() launches its kernels):
// replayed from a captured graph when the launch arguments are known in
// advance (sort pass count known, no kernel timing), else enqueued directly.
template <class F>
hipError_t segment(Engine* g, int method, int seg, F&& enqueue) {
    if (!g->use_graphs || g->timer.enabled || g->prep_passes < 0 || g->user_fn) {
        return enqueue();
    }
    std::array<uint64_t, 10> const key{
      static_cast<uint64_t>(method), static_cast<uint64_t>(seg),
      static_cast<uint64_t>(g->prep_passes) + 256u * static_cast<uint64_t>(g->prep_dense) +
        65536u * static_cast<uint64_t>(g->skin_reuse ? 2 : (g->skin_built ? 1 : 0)),
      static_cast<uint64_t>(g->build_cap), g->layout_gen,
      static_cast<uint64_t>(g->store_geometry), g->N, g->ntiles, static_cast<uint64_t>(g->nb),
      static_cast<uint64_t>(g->xcd_remap)};
    auto it = g->graphs.find(key);
    if (it == g->graphs.end()) {
        // Graphs of an older layout can never be replayed again.
        for (auto jt = g->graphs.begin(); jt != g->graphs.end();) {
            if (jt->first[4] != g->layout_gen) {
                (void)hipGraphExecDestroy(jt->second);
                jt = g->graphs.erase(jt);
            } else {
                ++jt;
            }
        }
        HIP_PASS(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        hipError_t const e = enqueue();
        hipGraph_t graph = nullptr;
        hipError_t const e2 = hipStreamEndCapture(g->stream, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) { (void)hipGraphDestroy(graph); }
            return e != hipSuccess ? e : e2;
        }
        hipGraphExec_t exec = nullptr;
        hipError_t const e3 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_PASS(e3);
        it = g->graphs.emplace(key, exec).first;
    }
    return hipGraphLaunch(it->second, g->stream);
}

enum Segment : int { SEG_PREFIX = 0, SEG_MIDDLE = 1, SEG_SUFFIX = 2 };

// The host callback's forces (forces.cpp:11-61 calling User_forces_function)
// into g->fext in sorted order: the state at this point of the sub-step is
// copied back in original order, the callback accumulates into zeroed forces,
// and they are copied in. after_dt: the callback's time is t + dt (DFSPH_P,
// dfsph_p.cpp:350). A short sort is detected first (the sub-step is then
// redone and the callback is not called for the discarded attempt).
hipError_t host_user_forces(Engine* g, bool after_dt) {
    if (!g->keys_checked) {
        HIP_PASS(read_keystate(g));
        if (g->sort_short) { return hipSuccess; }
    }
    uint32_t const n = g->N;
    auto& H = g->fhost;
    H.spos.resize(n);
    H.svel.resize(n);
    H.orig.resize(n);
    SceneDev sd{};
    HIP_PASS(hipMemcpyAsync(H.spos.data(), g->spos.p, n * sizeof(V2), hipMemcpyDeviceToHost,
                            g->stream));
    HIP_PASS(hipMemcpyAsync(H.svel.data(), g->svel.p, n * sizeof(V2), hipMemcpyDeviceToHost,
                            g->stream));
    HIP_PASS(hipMemcpyAsync(H.orig.data(), g->fidx.orig.p, n * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, g->stream));
    HIP_PASS(hipMemcpyAsync(&sd, g->scenes.p, sizeof(SceneDev), hipMemcpyDeviceToHost,
                            g->stream));
    HIP_PASS(hipStreamSynchronize(g->stream));
    H.pos.resize(2 * static_cast<size_t>(n));
    H.vel.resize(2 * static_cast<size_t>(n));
    H.forces.assign(2 * static_cast<size_t>(n), 0.0f);
    for (uint32_t t = 0; t < n; ++t) {
        size_t const o = H.orig[t];
        H.pos[2 * o] = H.spos[t].x;
        H.pos[2 * o + 1] = H.spos[t].y;
        H.vel[2 * o] = H.svel[t].x;
        H.vel[2 * o + 1] = H.svel[t].y;
    }
    emsph_config const& cfg = g->sh[0].cfg;
    float const dt = sd.c.dt;
    float const t0 = sd.frame.on ? static_cast<float>(to_seconds(sd.frame.time)) : g->step_time;
    float const t = after_dt ? t0 + dt : t0;
    g->user_fn(g->user_data, n, t, dt, cfg.mass_per_particle, cfg.draw_radius, H.forces.data(),
               H.pos.data(), H.vel.data());
    H.fs.resize(n);
    for (uint32_t k = 0; k < n; ++k) {
        size_t const o = H.orig[k];
        H.fs[k] = V2{H.forces[2 * o], H.forces[2 * o + 1]};
    }
    HIP_PASS(hipMemcpyAsync(g->fext.p, H.fs.data(), n * sizeof(V2), hipMemcpyHostToDevice,
                            g->stream));
    return hipStreamSynchronize(g->stream);
}

// External forces of a sub-step when some scene does not use the default
// gravity model (otherwise the build or the apply kernel integrates gravity
// itself): the host callback's forces, then the device models; apply = 1
// integrates them into the velocities, 0 stores them for the kernel that does.
hipError_t enqueue_forces(Engine* g, Batch const& b, int apply, bool after_dt) {
    if (!b.fext) { return hipSuccess; }
    int host = 0;
    if (g->user_fn) {
        HIP_PASS(host_user_forces(g, after_dt));
        if (g->sort_short) { return hipSuccess; }
        host = 1;
    }
    return timed(g, "external_forces", [&] { launch_external_forces(b, apply, host, g->stream); });
}

// iisph.cpp:23-151 for every scene of the batch.
emsph_status iisph_sub_step(Engine* g) {
    begin_attempt(g);
    V2* pa = g->pr0.p;
    V2* pb = g->pr1.p;
    Batch b = fluid_batch(g);
    // The a_ii sweep also starts the solver loop and runs its first
    // iteration (pressures +0, iisph.cpp:83): its pressures go where
    // iteration 0's would.
    b.pr_out = pb;
    HIP_TRY(segment(g, EMSPH_IISPH, SEG_PREFIX, [&]() -> hipError_t {
        HIP_PASS(enqueue_prepare(g, BUILD_IISPH, false));
        HIP_PASS(enqueue_forces(g, b, 1, false));
        return timed(g, "iisph_aii_beta", [&] { launch_iisph_aii_beta0(b, g->stream); });
    }));
    finish_prepare(g);
    int max_it = 0;
    for (auto const& h : g->sh) {
        max_it = std::max(max_it, std::min(h.cfg.params.iisph_max_pressure_iterations,
                                           MAX_RED_SLOTS));
    }
    bool short_sort = false;
    // iisph.cpp:126: no scene can exit before its third iteration.
    int const decide_from = first_decision(g, 0, 2);
    HIP_TRY(run_loop(g, 0, max_it, 2, [&](int it) -> hipError_t {
        b.pr_in = (it & 1) ? pb : pa;
        b.pr_out = (it & 1) ? pa : pb;
        // iisph.cpp:85-119: sum_j d_ij p_j, then beta and the new pressures
        // (iteration 0 ran in the a_ii sweep).
        if (it > 0) {
            ActiveTiles const at = active_tiles(g);
            HIP_PASS(timed(g, "iisph_sdp", [&] { launch_iisph_sdp(b, at, g->stream); }));
            HIP_PASS(timed(g, "iisph_beta_p", [&] { launch_iisph_beta_p(b, at, g->stream); }));
        }
        if (it < decide_from) { return hipSuccess; }
        return timed(g, "loop", [&] { launch_loop_decide(b, LOOP_IISPH, it, it, g->stream); });
    }, &g->last_iterations, &short_sort));
    if (short_sort) { return EMSPH_OK; }
    if (!g->keys_checked) {
        HIP_TRY(read_keystate(g));
        if (g->sort_short) { return EMSPH_OK; }
    }
    // Each scene's last pressure buffer follows its own iteration count; the
    // final pressures are also written to p0 for the field read-back.
    b.pr_in = pa;
    b.pr_out = nullptr;
    b.pr_alt = pb;
    b.p_out = g->p0.p;
    HIP_TRY(segment(g, EMSPH_IISPH, SEG_SUFFIX, [&]() -> hipError_t {
        HIP_PASS(timed(g, "apply_integrate", [&] { launch_apply(b, 0, 1, 1, g->stream); }));
        launch_scene_totals(b, 1, g->stream);
        return hipGetLastError();
    }));
    g->cur_p = g->p0.p;
    return EMSPH_OK;
}

// iisph_ap.cpp:44-138 (pseudo = false) and iisph_pseudo_ap.cpp:63-155
// (pseudo = true) for every scene of the batch: the same loop shape as IISPH
// (exit tests from the third iteration, iisph thresholds), pressures updated in
// place, the first iteration folded into the diagonal sweep.
emsph_status iisph_ap_sub_step(Engine* g, bool pseudo) {
    begin_attempt(g);
    int const method = pseudo ? EMSPH_IISPH_PSEUDO_AP : EMSPH_IISPH_AP;
    Batch b = fluid_batch(g, g->p0.p, g->p0.p);
    HIP_TRY(segment(g, method, SEG_PREFIX, [&]() -> hipError_t {
        HIP_PASS(enqueue_prepare(g, BUILD_AP, false));
        HIP_PASS(enqueue_forces(g, b, 1, false));
        return timed(g, "ap_diag", [&] { launch_ap_diag0(b, pseudo ? 1 : 0, g->stream); });
    }));
    finish_prepare(g);
    int max_it = 0;
    for (auto const& h : g->sh) {
        max_it = std::max(max_it, std::min(h.cfg.params.iisph_max_pressure_iterations,
                                           MAX_RED_SLOTS));
    }
    bool short_sort = false;
    int const decide_from = first_decision(g, 0, 2);
    HIP_TRY(run_loop(g, 0, max_it, 2, [&](int it) -> hipError_t {
        if (it > 0) {
            ActiveTiles const at = active_tiles(g);
            HIP_PASS(timed(g, "ap_accel", [&] { launch_ap_accel(b, at, g->stream); }));
            HIP_PASS(timed(g, "ap_iterate", [&] {
                launch_ap_iterate(b, pseudo ? 1 : 0, at, g->stream);
            }));
        }
        if (it < decide_from) { return hipSuccess; }
        return timed(g, "loop", [&] { launch_loop_decide(b, LOOP_IISPH, it, it, g->stream); });
    }, &g->last_iterations, &short_sort));
    if (short_sort) { return EMSPH_OK; }
    if (!g->keys_checked) {
        HIP_TRY(read_keystate(g));
        if (g->sort_short) { return EMSPH_OK; }
    }
    HIP_TRY(segment(g, method, SEG_SUFFIX, [&]() -> hipError_t {
        HIP_PASS(timed(g, "ap_final", [&] { launch_ap_final(b, pseudo ? 1 : 0, g->stream); }));
        launch_scene_totals(b, 1, g->stream);
        return hipGetLastError();
    }));
    g->cur_p = g->p0.p;
    return EMSPH_OK;
}

// dfsph_p.cpp:265-374 (with the correction loops of :61-212).
emsph_status dfsph_sub_step(Engine* g) {
    begin_attempt(g);
    Batch b = fluid_batch(g, g->kvs.p, nullptr);
    b.kappa = g->kvs.p;
    int max_it = 0;
    for (auto const& h : g->sh) {
        max_it = std::max(max_it, std::min(h.cfg.params.dfsph_max_correction_iterations,
                                           MAX_RED_SLOTS / 2));
    }
    // Divergence solve (dfsph_p.cpp:136-212): warm start, then iterations.
    HIP_TRY(segment(g, EMSPH_DFSPH_P, SEG_PREFIX, [&]() -> hipError_t {
        HIP_PASS(enqueue_prepare(g, BUILD_DFSPH, true));
        HIP_PASS(timed(g, "apply", [&] { launch_apply(b, 0, 0, 1, g->stream); }));
        return timed(g, "loop", [&] { launch_loop_begin(b, g->stream); });
    }));
    finish_prepare(g);
    bool short_sort = false;
    int const div_from = first_decision(g, 1, 1);
    HIP_TRY(run_loop(g, 1, max_it, 1, [&](int it) -> hipError_t {
        ActiveTiles const at = active_tiles(g);
        HIP_PASS(timed(g, "dfsph_div", [&] { launch_dfsph_div(b, at, g->stream); }));
        HIP_PASS(timed(g, "apply", [&] { launch_apply(b, 0, 0, 0, g->stream, at); }));
        if (it < div_from) { return hipSuccess; }
        return timed(g, "loop", [&] {
            launch_loop_decide(b, LOOP_DIVERGENCE, it, it, g->stream);
        });
    }, &g->last_div_iterations, &short_sort));
    if (short_sort) { return EMSPH_OK; }
    // External forces, then the density solve (dfsph_p.cpp:350-366, :61-134).
    int const base = MAX_RED_SLOTS / 2;
    b.p_in = g->kds.p;
    b.kappa = g->kds.p;
    HIP_TRY(segment(g, EMSPH_DFSPH_P, SEG_MIDDLE, [&]() -> hipError_t {
        HIP_PASS(enqueue_forces(g, b, 0, true));
        HIP_PASS(timed(g, "apply_gravity", [&] { launch_apply(b, 1, 0, 1, g->stream); }));
        return timed(g, "loop", [&] { launch_loop_begin(b, g->stream); });
    }));
    int const dens_from = first_decision(g, 2, 2);
    HIP_TRY(run_loop(g, 2, max_it, 2, [&](int it) -> hipError_t {
        ActiveTiles const at = active_tiles(g);
        HIP_PASS(timed(g, "dfsph_dens", [&] { launch_dfsph_dens(b, at, g->stream); }));
        HIP_PASS(timed(g, "apply", [&] { launch_apply(b, 0, 0, 0, g->stream, at); }));
        if (it < dens_from) { return hipSuccess; }
        return timed(g, "loop", [&] {
            launch_loop_decide(b, LOOP_DENSITY, it, base + it, g->stream);
        });
    }, &g->last_iterations, &short_sort));
    if (short_sort) { return EMSPH_OK; }
    if (!g->keys_checked) {
        HIP_TRY(read_keystate(g));
        if (g->sort_short) { return EMSPH_OK; }
    }
    HIP_TRY(segment(g, EMSPH_DFSPH_P, SEG_SUFFIX, [&]() -> hipError_t {
        HIP_PASS(timed(g, "finalize", [&] { launch_finalize(b, 1, g->stream); }));
        launch_scene_totals(b, 1, g->stream);
        return hipGetLastError();
    }));
    g->cur_p = g->kds.p;
    return EMSPH_OK;
}

// pcisph.cpp:36-164 for every scene of the batch.
emsph_status pcisph_sub_step(Engine* g) {
    begin_attempt(g);
    Batch b = fluid_batch(g, nullptr, g->p0.p);
    V2* vin = g->vstar0.p;
    V2* vout = g->vstar1.p;
    V2* xa = g->xstar.p;
    V2* xb = g->xstar1.p;
    b.vstar_out = vin;
    b.xstar = xa;
    HIP_TRY(segment(g, EMSPH_PCISPH, SEG_PREFIX, [&]() -> hipError_t {
        HIP_PASS(enqueue_prepare(g, BUILD_LISTS, false));
        HIP_PASS(enqueue_forces(g, b, 0, false));
        HIP_PASS(timed(g, "pci_init", [&] { launch_pci_init(b, g->stream); }));
        return timed(g, "loop", [&] { launch_loop_begin(b, g->stream); });
    }));
    finish_prepare(g);
    bool short_sort = false;
    int const pci_from = first_decision(g, 3, 2);
    HIP_TRY(run_loop(g, 3, 6, 2, [&](int it) -> hipError_t {
        // x* ping-pongs like v*: density and force read the iteration's x*,
        // force writes the next one.
        b.xstar_in = (it & 1) ? xb : xa;
        b.xstar = (it & 1) ? xa : xb;
        ActiveTiles const at = active_tiles(g);
        HIP_PASS(timed(g, "pci_density", [&] { launch_pci_density(b, at, g->stream); }));
        b.vstar_in = (it & 1) ? vout : vin;
        b.vstar_out = (it & 1) ? vin : vout;
        HIP_PASS(timed(g, "pci_force", [&] { launch_pci_force(b, at, g->stream); }));
        if (it < pci_from) { return hipSuccess; }
        return timed(g, "loop", [&] { launch_loop_decide(b, LOOP_PCISPH, it, it, g->stream); });
    }, &g->last_iterations, &short_sort));
    if (short_sort) { return EMSPH_OK; }
    if (!g->keys_checked) {
        HIP_TRY(read_keystate(g));
        if (g->sort_short) { return EMSPH_OK; }
    }
    // Each scene's last predictor follows its own iteration count.
    b.vstar_in = g->vstar0.p;
    b.vstar_out = g->vstar1.p;
    b.vstar_alt = g->vstar1.p;
    b.xstar_in = xa;
    b.xstar_alt = xb;
    HIP_TRY(segment(g, EMSPH_PCISPH, SEG_SUFFIX, [&]() -> hipError_t {
        HIP_PASS(timed(g, "pci_finalize", [&] { launch_pci_finalize(b, g->stream); }));
        launch_scene_totals(b, 1, g->stream);
        return hipGetLastError();
    }));
    g->cur_p = g->p0.p;
    return EMSPH_OK;
}

emsph_status check_method(int method) {
    if (method == EMSPH_IISPH || method == EMSPH_PCISPH || method == EMSPH_DFSPH_P ||
        method == EMSPH_IISPH_AP || method == EMSPH_IISPH_PSEUDO_AP) {
        return EMSPH_OK;
    }
    return fail(EMSPH_ERR_UNKNOWN_METHOD,
                "method " + std::to_string(method) + " is not implemented on the GPU path");
}

// Skin candidate lists before a sub-step's attempt: the buffers for the
// engine's capacity (allocated here, outside any graph capture; new buffers
// are a new layout and no candidates) and the choice: a
// first attempt reuses valid candidates unless the displacement ratio,
// extrapolated from the last two sub-steps, would reach 0.9 of the limit; a
// redo always rebuilds.
hipError_t skin_prepare(Engine* g, int attempt) {
    g->skin_reuse = false;
    g->skin_built = false;
    if (!g->skin_on) { return hipSuccess; }
    if (g->skin_pause > 0) {
        if (attempt == 0) { --g->skin_pause; }
        g->skin_valid = false;
        return hipSuccess;
    }
    void const* before[3] = {g->skin_cf.p, g->skin_cs.p, g->skin_tmax.p};
    size_t const stride = g->stride;
    HIP_PASS(g->skin_cf.ensure(static_cast<size_t>(SKIN_FLUID) * stride));
    HIP_PASS(g->skin_cs.ensure(static_cast<size_t>(SKIN_SOLID) * stride));
    HIP_PASS(g->skin_nf.ensure(stride));
    HIP_PASS(g->skin_ns.ensure(stride));
    HIP_PASS(g->skin_xref.ensure(g->cap));
    HIP_PASS(g->skin_tmax.ensure(std::max<uint32_t>(g->ntiles, 1)));
    if (before[0] != g->skin_cf.p || before[1] != g->skin_cs.p || before[2] != g->skin_tmax.p) {
        ++g->layout_gen;
        g->skin_valid = false;
    }
    float const next = 2.0f * g->skin_rho - g->skin_rho_prev;
    g->skin_reuse = attempt == 0 && g->skin_valid && next < 0.9f;
    if (attempt == 0 && !g->skin_reuse && g->skin_valid) {
        // The candidates of this interval are spent: an interval of fewer than
        // SKIN_MIN_RUN reuses does not repay the candidate build (one build
        // costs about as much as SKIN_MIN_RUN reuses save), so the next
        // SKIN_PAUSE sub-steps take the full path alone.
        if (g->skin_run < SKIN_MIN_RUN) { g->skin_pause = SKIN_PAUSE; }
        g->skin_valid = false;
    }
    return hipSuccess;
}

// After a sub-step that was not short: the candidates just built are valid
// unless one overflowed (the engine then stops using them: its lists are too
// crowded), a reuse records its displacement ratio, and a sub-step without
// either (the radix path) leaves none.
void skin_after(Engine* g) {
    if (!g->skin_on) { return; }
    if (g->skin_pause > 0 && !g->skin_reuse) {
        g->skin_valid = false;
        return;
    }
    if (g->skin_reuse) {
        ++g->skin_counts[1];
        ++g->skin_run;
        g->skin_rho_prev = g->skin_rho;
        g->skin_rho = std::sqrt(std::max(0.0f, __builtin_bit_cast(float, g->keys.skin_disp)));
    } else if (g->skin_built) {
        ++g->skin_counts[0];
        bool const over = (g->keys.skin_flags & SKIN_OVERFLOW) != 0u;
        g->skin_valid = !over;
        if (over) { g->skin_on = false; }
        g->skin_rho = 0.0f;
        g->skin_rho_prev = 0.0f;
        g->skin_run = 0;
    } else {
        g->skin_valid = false;
    }
}

emsph_status engine_sub_step(Engine* g, int method) {
    ST_TRY(check_method(method));
    if (g->N == 0 || g->ntiles == 0) { return EMSPH_OK; }  // N = 0: the reference reads block_indices.back()
    HIP_TRY(hipSetDevice(g->device));
    // Dense index tables for the bits the last key range asks for (allocated
    // here, outside any graph capture; a new table is a new layout).
    dense_prepare(g);
    // A sort run with the previous sub-step's pass count is checked at the
    // first read-back; when it was short the sub-step is redone with the exact
    // count (the second attempt sorts exactly and cannot be short).
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (attempt == 1) {
            // A reuse of the skin candidates whose list would have needed the
            // reference's truncation at 64 entries turns them off; one whose
            // particles moved too far is redone like a short sort, by the
            // exact full path.
            if (g->skin_reuse && (g->keys.skin_flags & SKIN_FULL)) { g->skin_on = false; }
            g->passes_hint = -1;
            g->totals[3] += 1;
        }
        HIP_TRY(skin_prepare(g, attempt));
        switch (method) {
        case EMSPH_IISPH: ST_TRY(iisph_sub_step(g)); break;
        case EMSPH_IISPH_AP: ST_TRY(iisph_ap_sub_step(g, false)); break;
        case EMSPH_IISPH_PSEUDO_AP: ST_TRY(iisph_ap_sub_step(g, true)); break;
        case EMSPH_DFSPH_P: ST_TRY(dfsph_sub_step(g)); break;
        default: ST_TRY(pcisph_sub_step(g)); break;
        }
        if (!g->sort_short) { break; }
    }
    skin_after(g);
    g->last_method = method;
    g->totals[0] += 1;
    g->totals[1] += g->last_iterations;
    g->totals[2] += g->last_div_iterations;
    return EMSPH_OK;
}

emsph_status read_stats(Engine* g) {
    if (!g->M) { return EMSPH_OK; }
    HIP_TRY(hipMemcpyAsync(g->host_stats.data(), g->stats.p, g->M * sizeof(SceneStats),
                           hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return EMSPH_OK;
}

void to_public(SceneStats const& s, Engine const* g, emsph_step_stats* out) {
    std::memset(out, 0, sizeof(*out));
    out->iterations = s.iterations;
    out->divergence_iterations = s.divergence_iterations;
    out->error_average = s.error_average;
    out->error_max = s.error_max;
    out->divergence_error_average = s.divergence_error_average;
    out->divergence_error_max = s.divergence_error_max;
    std::memcpy(&out->max_velocity_sq, &s.vmax2_bits, 4);
    out->sort_passes = static_cast<uint32_t>(g->fidx.key_passes + g->fidx.scene_passes);
    out->cell_count = g->keys.cells;
    out->max_neighbors = g->keys.max_count;
    out->dt = s.dt;
    out->kinetic_energy = s.kinetic_energy;
    out->potential_energy = s.potential_energy;
}

emsph_status engine_frame_init(Engine* g, int method) {
    ST_TRY(check_method(method));
    if (method == EMSPH_DFSPH_P && g->cap) {
        HIP_TRY(hipMemsetAsync(g->kv.p, 0, g->cap * sizeof(float), g->stream));
        HIP_TRY(hipMemsetAsync(g->kd.p, 0, g->cap * sizeof(float), g->stream));
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return EMSPH_OK;
}

// A per-particle field of the last sub-step, scene s, original order.
emsph_status engine_get_field(Engine* g, uint32_t s, int field, float* out) {
    if (!g->have_lists || !g->N) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no sub-step yet"); }
    float const* f1 = nullptr;
    V2 const* f2 = nullptr;
    switch (field) {
    case EMSPH_FIELD_DENSITY: f1 = g->rho.p; break;
    case EMSPH_FIELD_DENSITY_STAR: f1 = g->rho_star.p; break;
    case EMSPH_FIELD_PRESSURE: f1 = g->cur_p ? g->cur_p : g->p0.p; break;
    case EMSPH_FIELD_AII: f1 = g->aii.p; break;
    case EMSPH_FIELD_DII: f2 = g->dii.p; break;
    case EMSPH_FIELD_SUM_DIJ_PJ: break;  // the sdp half of the packed {sdp, d_ii}
    case EMSPH_FIELD_ALPHA: f1 = g->alpha.p; break;
    case EMSPH_FIELD_KAPPA_DENSITY: f1 = g->kds.p; break;
    case EMSPH_FIELD_KAPPA_DIVERGENCE: f1 = g->kvs.p; break;
    case EMSPH_FIELD_PRESSURE_FORCE: f2 = g->force.p; break;
    case EMSPH_FIELD_PRESSURE_ACCELERATION: f2 = g->pacc.p; break;
    case EMSPH_FIELD_EXTERNAL_FORCE: f2 = g->fext.p; break;
    default: return fail(EMSPH_ERR_INVALID_ARGUMENT, "unknown field");
    }
    SceneHost const& h = g->sh[s];
    if (field == EMSPH_FIELD_EXTERNAL_FORCE && !g->use_fext) {
        // The fused default: gravity only (forces.cpp:62-83).
        float const fy = 0.0f - h.cfg.mass_per_particle * force_model(h).gravity;
        for (uint32_t i = 0; i < h.n; ++i) {
            out[2 * i] = 0.0f;
            out[2 * i + 1] = fy;
        }
        return EMSPH_OK;
    }
    uint32_t f2_stride = 1;
    if (field == EMSPH_FIELD_SUM_DIJ_PJ) {
        f2 = &g->sd.p[0].sdp;
        f2_stride = sizeof(SdDii) / sizeof(V2);
    }
    if (f1) {
        DBuf<float> tmp;
        HIP_TRY(tmp.alloc(g->N));
        launch_scatter_f32(f1, g->fidx.orig.p, tmp.p, g->N, g->stream);
        HIP_TRY(hipMemcpyAsync(out, tmp.p + h.offset, h.n * sizeof(float), hipMemcpyDeviceToHost,
                               g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    } else {
        DBuf<V2> tmp;
        HIP_TRY(tmp.alloc(g->N));
        launch_scatter_v2(f2, f2_stride, g->fidx.orig.p, tmp.p, g->N, g->stream);
        HIP_TRY(hipMemcpyAsync(out, tmp.p + h.offset, h.n * sizeof(V2), hipMemcpyDeviceToHost,
                               g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    return EMSPH_OK;
}

emsph_status engine_get_neighborhood(Engine* g, uint32_t s, int which, uint8_t* counts,
                                     uint32_t* indices, float* distances, float* vectors_xy) {
    if (!g->have_lists || !g->N) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no sub-step yet"); }
    if ((distances || vectors_xy) && !g->fl_dist.p) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "geometry not stored (emsph_set_store_geometry)");
    }
    if (which && !g->NS) {
        std::memset(counts, 0, g->sh[s].n);
        return EMSPH_OK;
    }
    Lists l = which ? Lists{g->scount.p, g->sl_idx.p, g->sl_gw.p, g->sl_dist.p, g->sl_vec.p}
                    : Lists{g->fcount.p, g->fl_idx.p, g->fl_gw.p, g->fl_dist.p, g->fl_vec.p};
    uint32_t const* other = which ? g->sidx.orig.p : g->fidx.orig.p;
    uint32_t const* other_scene = which ? g->s_scene_of.p : g->scene_of.p;
    size_t const m = static_cast<size_t>(g->N) * MAXN;
    DBuf<uint8_t> dc;
    DBuf<uint32_t> di;
    DBuf<float> dd;
    DBuf<V2> dv;
    HIP_TRY(dc.alloc(g->N));
    HIP_TRY(di.alloc(m));
    if (distances) { HIP_TRY(dd.alloc(m)); }
    if (vectors_xy) { HIP_TRY(dv.alloc(m)); }
    launch_export_lists(l, g->N, g->stride, g->fidx.orig.p, other, other_scene, g->scenes.p,
                        which, dc.p, di.p, distances ? dd.p : nullptr,
                        vectors_xy ? dv.p : nullptr, g->stream);
    HIP_TRY(hipGetLastError());
    SceneHost const& h = g->sh[s];
    size_t const o = static_cast<size_t>(h.offset) * MAXN;
    size_t const k = static_cast<size_t>(h.n) * MAXN;
    HIP_TRY(hipMemcpyAsync(counts, dc.p + h.offset, h.n, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipMemcpyAsync(indices, di.p + o, k * 4, hipMemcpyDeviceToHost, g->stream));
    if (distances) {
        HIP_TRY(hipMemcpyAsync(distances, dd.p + o, k * 4, hipMemcpyDeviceToHost, g->stream));
    }
    if (vectors_xy) {
        HIP_TRY(hipMemcpyAsync(vectors_xy, dv.p + o, k * 8, hipMemcpyDeviceToHost, g->stream));
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return EMSPH_OK;
}

emsph_status engine_set_state(Engine* g, uint32_t s, float const* pos_xy, float const* vel_xy,
                              size_t count) {
    SceneHost& h = g->sh[s];
    if (count > h.cap) { return fail(EMSPH_ERR_CAPACITY, "fluid count above capacity"); }
    g->skin_valid = false;  // new positions: the candidates are rebuilt
    if (count) {
        HIP_TRY(hipMemcpyAsync(g->pos.p + h.offset, pos_xy, count * sizeof(V2),
                               hipMemcpyHostToDevice, g->stream));
        if (vel_xy) {
            HIP_TRY(hipMemcpyAsync(g->vel.p + h.offset, vel_xy, count * sizeof(V2),
                                   hipMemcpyHostToDevice, g->stream));
        } else {
            HIP_TRY(hipMemsetAsync(g->vel.p + h.offset, 0, count * sizeof(V2), g->stream));
        }
    }
    if (count != h.n) {
        h.n = static_cast<uint32_t>(count);
        ST_TRY(upload_layout(g));
        // The sort runs over [0, N): the packed extent ends at the last scene's
        // particles; only a single-scene engine may hold fewer than capacity.
        g->N = g->sh.back().offset + g->sh.back().n;
    }
    g->have_lists = false;
    g->passes_hint = -1;
    // max |v|^2 of the new state, for the CFL dt of the next sub-step.
    if (g->N) {
        Batch b = fluid_batch(g);
        launch_step_begin(b, g->stream);
        launch_vmax2(b, g->stream);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return EMSPH_OK;
}

emsph_status engine_set_cfl(Engine* g, float factor) {
    if (!(factor >= 0.0f)) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "CFL factor < 0"); }
    for (auto& h : g->sh) { h.cfl = factor; }
    ST_TRY(upload_layout(g));
    return EMSPH_OK;
}

emsph_status engine_get_state(Engine* g, uint32_t s, float* pos_xy, float* vel_xy) {
    SceneHost const& h = g->sh[s];
    if (h.n) {
        if (pos_xy) {
            HIP_TRY(hipMemcpyAsync(pos_xy, g->pos.p + h.offset, h.n * sizeof(V2),
                                   hipMemcpyDeviceToHost, g->stream));
        }
        if (vel_xy) {
            HIP_TRY(hipMemcpyAsync(vel_xy, g->vel.p + h.offset, h.n * sizeof(V2),
                                   hipMemcpyDeviceToHost, g->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return EMSPH_OK;
}

emsph_status engine_solid_volumes(Engine* g, uint32_t s, float* out) {
    SceneHost const& h = g->sh[s];
    if (!h.ns) { return EMSPH_OK; }
    DBuf<float> tmp;
    HIP_TRY(tmp.alloc(g->NS));
    launch_scatter_f32(g->s_vol.p, g->sidx.orig.p, tmp.p, g->NS, g->stream);
    HIP_TRY(hipMemcpyAsync(out, tmp.p + h.soffset, h.ns * sizeof(float), hipMemcpyDeviceToHost,
                           g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return EMSPH_OK;
}

emsph_status engine_set_dt(Engine* g, uint32_t s, float dt) {
    if (g->sh[s].dt != dt) {
        g->sh[s].dt = dt;
        uint32_t first_tile = 0;
        for (uint32_t k = 0; k < s; ++k) { first_tile += (g->sh[k].n + TILE - 1) / TILE; }
        SceneDev const sd = scene_dev(g->sh[s], first_tile);
        HIP_TRY(hipMemcpyAsync(g->scenes.p + s, &sd, sizeof(SceneDev), hipMemcpyHostToDevice,
                               g->stream));
        // sd lives on this stack frame: the copy must have read it before the
        // frame ends.
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    return EMSPH_OK;
}

// Totals the pool will hold after its staged adds and removes.
void pool_projection(emsph_pool const* P, uint64_t& n, uint64_t& ns, uint64_t& m) {
    n = ns = m = 0;
    for (auto const& sc : P->live) {
        if (std::find(P->removes.begin(), P->removes.end(), sc.id) != P->removes.end()) {
            continue;
        }
        n += sc.n;
        ns += sc.ns;
        ++m;
    }
    for (auto const& sc : P->adds) {
        n += sc.n;
        ns += sc.ns;
        ++m;
    }
}

// Applies the staged adds and removes: survivors keep their order; those after
// the first removed scene move down over it (gathered through the
// sorted-space scratch arrays in one kernel), added scenes are appended, and
// the solids of the live scenes are re-indexed.
emsph_status pool_commit(emsph_pool* P) {
    Engine* g = &P->e;
    if (P->adds.empty() && P->removes.empty()) { return EMSPH_OK; }
    HIP_TRY(hipSetDevice(g->device));
    g->skin_valid = false;
    hipStream_t const st = g->stream;
    std::vector<PoolScene> next;
    std::vector<uint3> seg;  // {src, dst, n} of survivors that move
    uint32_t src = 0, dst = 0;
    for (auto& sc : P->live) {
        uint32_t const n = sc.n;
        bool const gone =
          std::find(P->removes.begin(), P->removes.end(), sc.id) != P->removes.end();
        if (!gone) {
            if (src != dst) { seg.push_back(make_uint3(src, dst, n)); }
            dst += n;
            next.push_back(std::move(sc));
        }
        src += n;
    }
    if (!seg.empty()) {
        uint32_t const first = seg[0].y;
        uint32_t const total = dst - first;
        HIP_TRY(g->repack_segments.ensure(seg.size()));
        HIP_TRY(hipMemcpyAsync(g->repack_segments.p, seg.data(), seg.size() * sizeof(uint3),
                               hipMemcpyHostToDevice, st));
        launch_repack(g->repack_segments.p, static_cast<uint32_t>(seg.size()), total, g->pos.p,
                      g->vel.p, g->kv.p, g->kd.p, g->spos.p, g->svel.p, g->kvs.p, g->kds.p, st);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(g->pos.p + first, g->spos.p + first, total * sizeof(V2),
                               hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(g->vel.p + first, g->svel.p + first, total * sizeof(V2),
                               hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(g->kv.p + first, g->kvs.p + first, total * sizeof(float),
                               hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(g->kd.p + first, g->kds.p + first, total * sizeof(float),
                               hipMemcpyDeviceToDevice, st));
        // The segment table is read by the kernel before the next commit rewrites it.
    }
    uint32_t off = dst;
    for (auto& sc : P->adds) {
        uint32_t const n = sc.n;
        HIP_TRY(hipMemcpyAsync(g->pos.p + off, sc.pos.data(), n * sizeof(V2),
                               hipMemcpyHostToDevice, st));
        if (!sc.vel.empty()) {
            HIP_TRY(hipMemcpyAsync(g->vel.p + off, sc.vel.data(), n * sizeof(V2),
                                   hipMemcpyHostToDevice, st));
        } else {
            HIP_TRY(hipMemsetAsync(g->vel.p + off, 0, n * sizeof(V2), st));
        }
        HIP_TRY(hipMemsetAsync(g->kv.p + off, 0, n * sizeof(float), st));
        HIP_TRY(hipMemsetAsync(g->kd.p + off, 0, n * sizeof(float), st));
        off += n;
    }
    // The host copies of the added states must outlive the copies.
    HIP_TRY(hipStreamSynchronize(st));
    for (auto& sc : P->adds) {
        std::vector<float>().swap(sc.pos);
        std::vector<float>().swap(sc.vel);
        next.push_back(std::move(sc));
    }
    P->adds.clear();
    P->removes.clear();
    P->live = std::move(next);

    std::vector<SceneHost> hs(P->live.size());
    std::vector<float const*> sx(P->live.size(), nullptr);
    uint32_t o = 0;
    for (size_t k = 0; k < P->live.size(); ++k) {
        PoolScene const& sc = P->live[k];
        hs[k].cfg = sc.cfg;
        hs[k].cap = sc.n;
        hs[k].n = sc.n;
        hs[k].ns = sc.ns;
        hs[k].offset = o;
        hs[k].dt = sc.cfg.seconds_per_sub_step;
        hs[k].cfl = P->cfl;
        sx[k] = sc.solids.empty() ? nullptr : sc.solids.data();
        o += sc.n;
    }
    g->sh = std::move(hs);
    g->M = static_cast<uint32_t>(g->sh.size());
    g->N = o;
    // The dense cell index is laid out again for the new scene list: table
    // sizes from the next exact run, and a new layout generation, so the
    // solid-cell bitmap over its bins is rebuilt from the new solids.
    g->fidx.dense_layout.clear();
    g->fidx.dense_bins = 0;
    g->dense_stale = true;
    ST_TRY(upload_layout(g));
    ST_TRY(build_solids(g, sx));
    ST_TRY(upload_layout(g));  // solid offsets are now known
    g->passes_hint = -1;
    g->have_lists = false;
    if (g->N) {
        Batch b = fluid_batch(g);
        launch_step_begin(b, st);
        launch_vmax2(b, st);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(st));
    return EMSPH_OK;
}

// A copy of everything a sub-step changes (original-order state, DFSPH warm
// starts, per-scene statistics, the host's sort and loop hints), so that trial
// sub-steps can be undone.
struct Snapshot {
    DBuf<V2> pos, vel;
    DBuf<float> kv, kd;
    std::vector<SceneStats> stats;
    int passes_hint = -1;
    int build_cap = 32;
    int loop_expect[4] = {64, 64, 64, 64};
    int64_t totals[4] = {0, 0, 0, 0};
};

hipError_t snapshot_take(Engine* g, Snapshot& sn) {
    uint32_t const n = std::max<uint32_t>(g->N, 1);
    HIP_PASS(sn.pos.ensure(n));
    HIP_PASS(sn.vel.ensure(n));
    HIP_PASS(sn.kv.ensure(n));
    HIP_PASS(sn.kd.ensure(n));
    if (g->N) {
        HIP_PASS(hipMemcpyAsync(sn.pos.p, g->pos.p, g->N * sizeof(V2), hipMemcpyDeviceToDevice,
                                g->stream));
        HIP_PASS(hipMemcpyAsync(sn.vel.p, g->vel.p, g->N * sizeof(V2), hipMemcpyDeviceToDevice,
                                g->stream));
        HIP_PASS(hipMemcpyAsync(sn.kv.p, g->kv.p, g->N * sizeof(float), hipMemcpyDeviceToDevice,
                                g->stream));
        HIP_PASS(hipMemcpyAsync(sn.kd.p, g->kd.p, g->N * sizeof(float), hipMemcpyDeviceToDevice,
                                g->stream));
    }
    sn.stats.resize(g->M);
    if (g->M) {
        HIP_PASS(hipMemcpyAsync(sn.stats.data(), g->stats.p, g->M * sizeof(SceneStats),
                                hipMemcpyDeviceToHost, g->stream));
    }
    HIP_PASS(hipStreamSynchronize(g->stream));
    sn.passes_hint = g->passes_hint;
    sn.build_cap = g->build_cap;
    std::copy(g->loop_expect, g->loop_expect + 4, sn.loop_expect);
    std::copy(g->totals, g->totals + 4, sn.totals);
    return hipSuccess;
}

hipError_t snapshot_restore(Engine* g, Snapshot const& sn) {
    g->skin_valid = false;
    if (g->N) {
        HIP_PASS(hipMemcpyAsync(g->pos.p, sn.pos.p, g->N * sizeof(V2), hipMemcpyDeviceToDevice,
                                g->stream));
        HIP_PASS(hipMemcpyAsync(g->vel.p, sn.vel.p, g->N * sizeof(V2), hipMemcpyDeviceToDevice,
                                g->stream));
        HIP_PASS(hipMemcpyAsync(g->kv.p, sn.kv.p, g->N * sizeof(float), hipMemcpyDeviceToDevice,
                                g->stream));
        HIP_PASS(hipMemcpyAsync(g->kd.p, sn.kd.p, g->N * sizeof(float), hipMemcpyDeviceToDevice,
                                g->stream));
    }
    if (g->M) {
        HIP_PASS(hipMemcpyAsync(g->stats.p, sn.stats.data(), g->M * sizeof(SceneStats),
                                hipMemcpyHostToDevice, g->stream));
    }
    HIP_PASS(hipStreamSynchronize(g->stream));
    g->passes_hint = sn.passes_hint;
    g->build_cap = sn.build_cap;
    std::copy(sn.loop_expect, sn.loop_expect + 4, g->loop_expect);
    std::copy(sn.totals, sn.totals + 4, g->totals);
    g->have_lists = false;
    return hipSuccess;
}

void tuning_get(Engine const* g, emsph_tuning* t) {
    std::memset(t, 0, sizeof(*t));
    t->graphs = g->use_graphs ? 1 : 0;
    t->neighbor_batch = g->nb;
    t->xcd_remap = g->xcd_remap ? 1 : 0;
}

// persistent_loop is refused: a cooperative launch of the whole IISPH solve
// (grid-wide barriers between sweeps) did not finish at 1,792 workgroups on
// the MI355X while it passed at 8, and the cause was not established, so the
// mode is off.
emsph_status tuning_set(Engine* g, emsph_tuning const* t) {
    if (t->persistent_loop) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "persistent_loop is not available");
    }
    if (t->neighbor_batch != 2 && t->neighbor_batch != 4 && t->neighbor_batch != 8) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "neighbor_batch must be 2, 4 or 8");
    }
    g->use_graphs = t->graphs != 0;
    g->nb = t->neighbor_batch;
    g->xcd_remap = t->xcd_remap != 0;
    return EMSPH_OK;
}

// Times each candidate tuning on trial_sub_steps sub-steps of the engine's own
// workload, all from the same state (restored after each trial and at the
// end, so the simulation does not advance), and keeps the fastest. Every
// candidate gives identical results; only the time differs.
emsph_status engine_autotune(Engine* g, int method, int trials, emsph_tuning* chosen,
                             double* times_ms, int* n_candidates) {
    if (trials < 1) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "trial_sub_steps < 1"); }
    // Shapes first (neighbour batch depth x XCD remap, with graphs), then the
    // best shape without graphs.
    std::vector<emsph_tuning> cands;
    for (int nb : {2, 4, 8}) {
        for (int xcd = 1; xcd >= 0; --xcd) {
            emsph_tuning t{};
            t.graphs = 1;
            t.neighbor_batch = nb;
            t.xcd_remap = xcd;
            cands.push_back(t);
        }
    }
    size_t const n_shapes = cands.size();
    Snapshot sn;
    HIP_TRY(snapshot_take(g, sn));
    // One timed trial of candidate k from the snapshot: warm-up (graph
    // capture, first-pass sort), then `trials` sub-steps from the same state.
    auto trial = [&](size_t k, double& ms) -> emsph_status {
        ST_TRY(tuning_set(g, &cands[k]));
        ST_TRY(engine_sub_step(g, method));
        HIP_TRY(snapshot_restore(g, sn));
        ST_TRY(engine_sub_step(g, method));
        HIP_TRY(hipStreamSynchronize(g->stream));
        auto const t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < trials; ++i) { ST_TRY(engine_sub_step(g, method)); }
        HIP_TRY(hipStreamSynchronize(g->stream));
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
               .count();
        HIP_TRY(snapshot_restore(g, sn));
        return EMSPH_OK;
    };
    // Each candidate keeps the best of TUNE_ROUNDS trials, the shapes taken in
    // turn round after round, so that a slow moment of the device does not
    // decide between candidates a few per cent apart.
    constexpr int TUNE_ROUNDS = 3;
    std::vector<double> ms_of(n_shapes + 1, std::numeric_limits<double>::infinity());
    for (int r = 0; r < TUNE_ROUNDS; ++r) {
        for (size_t k = 0; k < n_shapes; ++k) {
            double ms = 0.0;
            ST_TRY(trial(k, ms));
            ms_of[k] = std::min(ms_of[k], ms);
        }
    }
    size_t best_k = 0;
    for (size_t k = 1; k < n_shapes; ++k) {
        if (ms_of[k] < ms_of[best_k]) { best_k = k; }
    }
    {
        emsph_tuning t = cands[best_k];  // the best shape without graphs
        t.graphs = 0;
        cands.push_back(t);
        for (int r = 0; r < TUNE_ROUNDS; ++r) {
            double ms = 0.0;
            ST_TRY(trial(n_shapes, ms));
            ms_of[n_shapes] = std::min(ms_of[n_shapes], ms);
        }
        if (ms_of[n_shapes] < ms_of[best_k]) { best_k = n_shapes; }
    }
    for (size_t k = 0; k < cands.size() && k < 8; ++k) {
        if (times_ms) { times_ms[k] = ms_of[k]; }
    }
    ST_TRY(tuning_set(g, &cands[best_k]));
    if (chosen) { *chosen = cands[best_k]; }
    if (n_candidates) { *n_candidates = static_cast<int>(cands.size()); }
    return EMSPH_OK;
}

int pool_index(emsph_pool const* P, int id) {
    for (size_t k = 0; k < P->live.size(); ++k) {
        if (P->live[k].id == id) { return static_cast<int>(k); }
    }
    return -1;
}

}  // namespace

//==============================================================================
extern "C" {

const char* emsph_last_error(void) { return g_last_error.c_str(); }

int emsph_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) { return 0; }
    return n;
}

const char* emsph_build_info(void) {
    return "emerald_sph gfx950; f32 IEEE (contract=off, correctly rounded div/sqrt, "
           "denormals kept); multi-scene tiles, sorted-slot neighbour lists, LSD radix sort";
}

//------------------------------------------------------------------------------
// Single scene: a batch of one.
static void engine_set_skin(Engine* g, int enable) {
    // Skin lists need the |v|^2 < cut2 form of the acceptance test (every
    // scene's cut2 > 0, i.e. a cut-off above 1e-18).
    bool ok = enable != 0;
    for (auto const& h : g->sh) { ok = ok && step_consts(h.cfg, h.dt).cut2 > 0.0f; }
    g->skin_on = ok;
    g->skin_valid = false;
}

static void engine_get_skin(Engine const* g, int* enabled, int64_t* rebuilds, int64_t* reuses) {
    if (enabled) { *enabled = g->skin_on ? 1 : 0; }
    if (rebuilds) { *rebuilds = g->skin_counts[0]; }
    if (reuses) { *reuses = g->skin_counts[1]; }
}

emsph_status emsph_create(const emsph_config* config, const float* solid_xy, size_t solid_count,
                          size_t fluid_capacity, emsph_context** out) {
    if (!config || !out || (solid_count && !solid_xy)) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument");
    }
    if (fluid_capacity > (1ull << 30) || solid_count > (1ull << 30)) {
        return fail(EMSPH_ERR_CAPACITY, "capacity above 2^30 particles");
    }
    auto* c = new emsph_context();
    emsph_status s = engine_init(&c->e);
    if (s) {
        delete c;
        return s;
    }
    c->cfg = *config;
    SceneHost h;
    h.cfg = *config;
    h.cap = static_cast<uint32_t>(std::max<size_t>(fluid_capacity, 1));
    h.n = 0;
    h.ns = static_cast<uint32_t>(solid_count);
    h.dt = config->seconds_per_sub_step;
    s = engine_configure(&c->e, {h}, {solid_xy});
    if (s) {
        delete c;
        return s;
    }
    engine_set_skin(&c->e, 1);  // skin candidate lists on by default
    *out = c;
    return EMSPH_OK;
}

emsph_status emsph_destroy(emsph_context* ctx) {
    if (!ctx) { return EMSPH_OK; }
    (void)hipStreamSynchronize(ctx->e.stream);
    delete ctx;
    return EMSPH_OK;
}

emsph_status emsph_set_store_geometry(emsph_context* c, int enable) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    c->e.store_geometry = enable != 0;
    if (c->e.store_geometry) { HIP_TRY(ensure_geometry(&c->e)); }
    return EMSPH_OK;
}

emsph_status emsph_set_state(emsph_context* c, const float* pos_xy, const float* vel_xy,
                             size_t count) {
    if (!c || (count && !pos_xy)) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    return engine_set_state(&c->e, 0, pos_xy, vel_xy, count);
}

emsph_status emsph_get_state(emsph_context* c, float* pos_xy, float* vel_xy) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    return engine_get_state(&c->e, 0, pos_xy, vel_xy);
}

size_t emsph_fluid_count(const emsph_context* c) { return c ? c->e.sh[0].n : 0; }
size_t emsph_solid_count(const emsph_context* c) { return c ? c->e.sh[0].ns : 0; }

emsph_status emsph_get_solid_volumes(emsph_context* c, float* out) {
    if (!c || !out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    return engine_solid_volumes(&c->e, 0, out);
}

emsph_status emsph_get_field(emsph_context* c, int field, float* out) {
    if (!c || !out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    return engine_get_field(&c->e, 0, field, out);
}

emsph_status emsph_get_neighborhood(emsph_context* c, int which, uint8_t* counts,
                                    uint32_t* indices, float* distances, float* vectors_xy) {
    if (!c || !counts || !indices) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    return engine_get_neighborhood(&c->e, 0, which, counts, indices, distances, vectors_xy);
}

emsph_status emsph_frame_init(emsph_context* c, int method) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    return engine_frame_init(&c->e, method);
}

emsph_status emsph_sub_step(emsph_context* c, int method, float time_seconds, float dt,
                            emsph_step_stats* stats) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    c->e.step_time = time_seconds;  // passed to a host force callback
    if (stats) { std::memset(stats, 0, sizeof(*stats)); }
    ST_TRY(check_method(method));
    ST_TRY(engine_set_dt(&c->e, 0, dt));
    ST_TRY(engine_sub_step(&c->e, method));
    if (stats) {
        ST_TRY(read_stats(&c->e));
        to_public(c->e.host_stats[0], &c->e, stats);
    }
    return EMSPH_OK;
}

emsph_status emsph_frame_step(emsph_context* c, int method, int64_t global_time_flicks,
                              emsph_frame_stats* fs) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    ST_TRY(emsph_frame_init(c, method));
    Engine* g = &c->e;
    HIP_TRY(hipSetDevice(g->device));
    emsph_frame_stats local{};
    emsph_parameters const& P = c->cfg.params;
    int64_t end_time = global_time_flicks;
    if (method == EMSPH_PCISPH) {
        // pcisph.cpp:166-204: fixed sub-steps of the configured dt, enqueued
        // back to back (no read-back between them).
        int64_t const per = P.time_per_step / P.sub_steps;
        ST_TRY(engine_set_dt(g, 0, c->cfg.seconds_per_sub_step));
        for (int k = 0; k < P.sub_steps; ++k) {
            g->step_time = static_cast<float>(to_seconds(end_time));
            ST_TRY(engine_sub_step(g, method));
            end_time += per;
            local.sub_steps += 1;
            local.total_iterations += g->last_iterations;
        }
    } else if (g->N) {
        // adaptive_time_step.h:35-122 with (60, 6, 0.2) for IISPH
        // (iisph.cpp:175-181) and (60, 4, 0.4) for DFSPH_P, IISPH_AP and
        // IISPH_PSEUDO_AP (dfsph_p.cpp:383-389): each sub-step's dt is picked
        // on the device at its start from the max |v|^2 the previous one left
        // (k_key_range, frame_begin); the host learns the quanta left from the
        // key-state read-back that the solver loop makes anyway, so sub-steps
        // follow each other without a read-back of their own.
        int const min_den = 60;
        int const max_den = (method == EMSPH_IISPH) ? 6 : 4;
        float const cfl = (method == EMSPH_IISPH) ? 0.2f : 0.4f;
        Batch b = fluid_batch(g);
        launch_step_begin(b, g->stream);
        launch_vmax2(b, g->stream);  // max |v|^2 of the state at the frame's start
        HIP_TRY(hipGetLastError());
        SceneDev sd = scene_dev(g->sh[0], 0);
        sd.frame.on = 1;
        sd.frame.per = P.time_per_step / min_den;
        sd.frame.time = global_time_flicks;
        sd.frame.next_time = global_time_flicks;
        sd.frame.remaining = min_den;
        sd.frame.next_remaining = min_den;
        sd.frame.divisor = min_den / max_den;
        sd.frame.cfl = cfl;
        HIP_TRY(hipMemcpyAsync(g->scenes.p, &sd, sizeof(SceneDev), hipMemcpyHostToDevice,
                               g->stream));
        g->frame_on = true;
        emsph_status st = EMSPH_OK;
        while (local.sub_steps < min_den) {
            st = engine_sub_step(g, method);
            if (st != EMSPH_OK) { break; }
            local.sub_steps += 1;
            local.total_iterations += g->last_iterations;
            local.total_divergence_iterations += g->last_div_iterations;
            if (g->keys.frame_remaining <= 0) { break; }
        }
        g->frame_on = false;
        // Back to the host's sub-step dt for later emsph_sub_step calls.
        SceneDev const plain = scene_dev(g->sh[0], 0);
        HIP_TRY(hipMemcpyAsync(g->scenes.p, &plain, sizeof(SceneDev), hipMemcpyHostToDevice,
                               g->stream));
        ST_TRY(st);
        end_time = global_time_flicks + sd.frame.per * min_den;
    } else {
        // No particles: max |v|^2 is 0, every sub-step takes the most quanta.
        int const max_den = (method == EMSPH_IISPH) ? 6 : 4;
        local.sub_steps = max_den;
        end_time = global_time_flicks + (P.time_per_step / 60) * 60;
    }
    emsph_step_stats st{};
    if (g->N) {
        ST_TRY(read_stats(g));
        to_public(g->host_stats[0], g, &st);
    }
    local.last = st;
    local.last_dt = st.dt;
    // adaptive_time_step.h:105-110 / pcisph.cpp:186-192: colors at the frame's
    // end time.
    if (g->N) { ST_TRY(emsph_compute_colors(c, static_cast<float>(to_seconds(end_time)))); }
    if (fs) { *fs = local; }
    return EMSPH_OK;
}

emsph_status emsph_run_sub_steps(emsph_context* c, int method, float dt, int n_sub_steps,
                                 float* elapsed_ms, emsph_step_stats* last) {
    if (!c || n_sub_steps < 0) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument"); }
    ST_TRY(check_method(method));
    ST_TRY(engine_set_dt(&c->e, 0, dt));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, c->e.stream));
    for (int k = 0; k < n_sub_steps; ++k) {
        c->e.step_time = static_cast<float>(k) * dt;  // sub-step k starts at k dt
        emsph_status const s = engine_sub_step(&c->e, method);
        if (s) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return s;
        }
    }
    HIP_TRY(hipEventRecord(e1, c->e.stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) { *elapsed_ms = ms; }
    if (last) {
        ST_TRY(read_stats(&c->e));
        to_public(c->e.host_stats[0], &c->e, last);
    }
    return EMSPH_OK;
}

static void engine_iteration_totals(Engine* g, int reset, int64_t* out) {
    if (out) {
        for (int k = 0; k < 4; ++k) { out[k] = g->totals[k]; }
    }
    if (reset) {
        for (int k = 0; k < 4; ++k) { g->totals[k] = 0; }
    }
}

emsph_status emsph_get_iteration_totals(emsph_context* c, int reset, int64_t* out) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    engine_iteration_totals(&c->e, reset, out);
    return EMSPH_OK;
}

emsph_status emsph_pool_get_iteration_totals(emsph_pool* P, int reset, int64_t* out) {
    if (!P || !out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    engine_iteration_totals(&P->e, reset, out);
    return EMSPH_OK;
}

emsph_status emsph_batch_get_iteration_totals(emsph_batch* b, int reset, int64_t* out) {
    if (!b) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null batch"); }
    engine_iteration_totals(&b->e, reset, out);
    return EMSPH_OK;
}

emsph_status emsph_get_tuning(emsph_context* c, emsph_tuning* out) {
    if (!c || !out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument"); }
    tuning_get(&c->e, out);
    return EMSPH_OK;
}

emsph_status emsph_set_tuning(emsph_context* c, const emsph_tuning* t) {
    if (!c || !t) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument"); }
    return tuning_set(&c->e, t);
}

emsph_status emsph_autotune(emsph_context* c, int method, float dt, int trial_sub_steps,
                            emsph_tuning* chosen, double* times_ms, int* n_candidates) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    ST_TRY(check_method(method));
    ST_TRY(engine_set_dt(&c->e, 0, dt));
    return engine_autotune(&c->e, method, trial_sub_steps, chosen, times_ms, n_candidates);
}

emsph_status emsph_batch_autotune(emsph_batch* b, int method, int trial_sub_steps,
                                  emsph_tuning* chosen, double* times_ms, int* n_candidates) {
    if (!b) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null batch"); }
    ST_TRY(check_method(method));
    return engine_autotune(&b->e, method, trial_sub_steps, chosen, times_ms, n_candidates);
}

emsph_status emsph_batch_set_tuning(emsph_batch* b, const emsph_tuning* t) {
    if (!b || !t) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument"); }
    return tuning_set(&b->e, t);
}

emsph_status emsph_pool_set_tuning(emsph_pool* P, const emsph_tuning* t) {
    if (!P || !t) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument"); }
    return tuning_set(&P->e, t);
}

//------------------------------------------------------------------------------
// Colors (colors.cpp:9-70; computed where adaptive_time_step.h:105-110 and
// pcisph.cpp:186-192 call compute_colors).

namespace {

emsph_status engine_compute_colors(Engine* g, float time_seconds) {
    SceneHost const& h = g->sh[0];
    if (!h.n) {
        g->colors_valid = true;
        return EMSPH_OK;
    }
    if (!g->have_lists) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no sub-step yet"); }
    if (!g->colors.p) { HIP_TRY(g->colors.alloc(g->cap)); }
    if (g->colors_fn) {
        // compute_colors (colors.cpp:22-40): the callback gets the state and the
        // densities in original order on the host.
        size_t const n = h.n;
        std::vector<float> pos(2 * n), vel(2 * n), rho(n), rgba(4 * n, 0.0f);
        ST_TRY(engine_get_state(g, 0, pos.data(), vel.data()));
        ST_TRY(engine_get_field(g, 0, EMSPH_FIELD_DENSITY, rho.data()));
        g->colors_fn(g->colors_data, n, time_seconds, h.cfg.mass_per_particle, h.cfg.draw_radius,
                     rgba.data(), pos.data(), vel.data(), rho.data());
        HIP_TRY(hipMemcpyAsync(g->colors.p, rgba.data(), n * sizeof(float4),
                               hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    } else {
        launch_colors(fluid_batch(g), g->colors.p, g->stream);
        HIP_TRY(hipGetLastError());
    }
    g->colors_valid = true;
    return EMSPH_OK;
}

}  // namespace

emsph_status emsph_set_user_colors(emsph_context* c, emsph_user_colors_fn fn, void* user) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    c->e.colors_fn = fn;
    c->e.colors_data = fn ? user : nullptr;
    return EMSPH_OK;
}

emsph_status emsph_compute_colors(emsph_context* c, float time_seconds) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    HIP_TRY(hipSetDevice(c->e.device));
    return engine_compute_colors(&c->e, time_seconds);
}

emsph_status emsph_get_colors(emsph_context* c, float* rgba) {
    if (!c || !rgba) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    Engine* g = &c->e;
    if (!g->colors_valid) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no colors computed yet"); }
    if (!g->sh[0].n) { return EMSPH_OK; }
    HIP_TRY(hipMemcpyAsync(rgba, g->colors.p, g->sh[0].n * sizeof(float4), hipMemcpyDeviceToHost,
                           g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return EMSPH_OK;
}

emsph_status emsph_get_colors_u8(emsph_context* c, uint8_t* rgba) {
    if (!c || !rgba) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    Engine* g = &c->e;
    if (!g->colors_valid) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no colors computed yet"); }
    uint32_t const n = g->sh[0].n;
    if (!n) { return EMSPH_OK; }
    if (!g->colors8.p) { HIP_TRY(g->colors8.alloc(g->cap)); }
    launch_convert_colors(g->colors.p, g->colors8.p, n, g->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(rgba, g->colors8.p, n * sizeof(uchar4), hipMemcpyDeviceToHost,
                           g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return EMSPH_OK;
}

//------------------------------------------------------------------------------
// External forces.

namespace {

emsph_status check_force_model(emsph_force_model const* m) {
    if (!m) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null force model"); }
    uint32_t const known = EMSPH_FORCE_GRAVITY | EMSPH_FORCE_POLE | EMSPH_FORCE_DRAG |
                           EMSPH_FORCE_ANTI_COUPLING;
    if (m->flags & ~known) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "unknown force flags"); }
    return EMSPH_OK;
}

// After a force model or callback change: whether the per-slot force buffer is
// used, the buffer itself, and the scene table (a new layout generation, so
// no captured graph of the old kernel sequence is replayed).
emsph_status engine_forces_changed(Engine* g) {
    bool use = g->user_fn != nullptr;
    for (auto const& h : g->sh) { use = use || force_model(h).flags != EMSPH_FORCE_GRAVITY; }
    if (use && !g->fext.p && g->cap) { HIP_TRY(g->fext.alloc(g->cap)); }
    g->use_fext = use;
    return upload_layout(g);
}

}  // namespace

emsph_status emsph_default_force_model(const emsph_config* config, emsph_force_model* out) {
    if (!config || !out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    SceneHost h;
    h.cfg = *config;
    *out = force_model(h);
    return EMSPH_OK;
}

emsph_status emsph_set_force_model(emsph_context* c, const emsph_force_model* model) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    ST_TRY(check_force_model(model));
    c->e.sh[0].force = *model;
    c->e.sh[0].force_set = true;
    return engine_forces_changed(&c->e);
}

emsph_status emsph_get_force_model(emsph_context* c, emsph_force_model* out) {
    if (!c || !out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    *out = force_model(c->e.sh[0]);
    return EMSPH_OK;
}

emsph_status emsph_set_user_forces(emsph_context* c, emsph_user_forces_fn fn, void* user) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    c->e.user_fn = fn;
    c->e.user_data = fn ? user : nullptr;
    return engine_forces_changed(&c->e);
}

emsph_status emsph_batch_set_force_model(emsph_batch* b, int scene,
                                         const emsph_force_model* model) {
    if (!b || scene < 0 || static_cast<uint32_t>(scene) >= b->e.M) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad batch or scene");
    }
    ST_TRY(check_force_model(model));
    b->e.sh[scene].force = *model;
    b->e.sh[scene].force_set = true;
    return engine_forces_changed(&b->e);
}

emsph_status emsph_set_graphs(emsph_context* c, int enable) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    c->e.use_graphs = enable != 0;
    return EMSPH_OK;
}

emsph_status emsph_set_skin_lists(emsph_context* c, int enable) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    engine_set_skin(&c->e, enable);
    return EMSPH_OK;
}

emsph_status emsph_get_skin_lists(emsph_context* c, int* enabled, int64_t* rebuilds,
                                  int64_t* reuses) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    engine_get_skin(&c->e, enabled, rebuilds, reuses);
    return EMSPH_OK;
}

emsph_status emsph_batch_set_skin_lists(emsph_batch* b, int enable) {
    if (!b) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null batch"); }
    engine_set_skin(&b->e, enable);
    return EMSPH_OK;
}

emsph_status emsph_batch_get_skin_lists(emsph_batch* b, int* enabled, int64_t* rebuilds,
                                        int64_t* reuses) {
    if (!b) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null batch"); }
    engine_get_skin(&b->e, enabled, rebuilds, reuses);
    return EMSPH_OK;
}

emsph_status emsph_batch_set_graphs(emsph_batch* b, int enable) {
    if (!b) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null batch"); }
    b->e.use_graphs = enable != 0;
    return EMSPH_OK;
}

emsph_status emsph_set_cfl(emsph_context* c, float factor) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    return engine_set_cfl(&c->e, factor);
}

static emsph_status engine_set_kernel_timing(Engine* g, int enable) {
    HIP_TRY(timer_flush(g));
    KernelTimer& tm = g->timer;
    if (enable) {
        // The whole event pool up front: creating events between launches
        // delays the host, the device drains, and the wait lands inside the
        // next stage's interval.
        while (tm.ev.size() < 2 * KernelTimer::POOL) {
            hipEvent_t e = nullptr;
            HIP_TRY(hipEventCreate(&e));
            tm.ev.push_back(e);
        }
    }
    tm.enabled = enable != 0;
    tm.acc.clear();
    return EMSPH_OK;
}

static int engine_get_kernel_timing(Engine* g, const char** names, double* total_ms,
                                    int64_t* calls, int max) {
    if (timer_flush(g)) { return -1; }
    int k = 0;
    for (auto const& kv : g->timer.acc) {
        if (k >= max) { break; }
        if (names) { names[k] = kv.first.c_str(); }
        if (total_ms) { total_ms[k] = kv.second.first; }
        if (calls) { calls[k] = kv.second.second; }
        ++k;
    }
    return k;
}

emsph_status emsph_set_kernel_timing(emsph_context* c, int enable) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    return engine_set_kernel_timing(&c->e, enable);
}

int emsph_get_kernel_timing(emsph_context* c, const char** names, double* total_ms,
                            int64_t* calls, int max) {
    if (!c) { return 0; }
    return engine_get_kernel_timing(&c->e, names, total_ms, calls, max);
}

emsph_status emsph_batch_set_kernel_timing(emsph_batch* b, int enable) {
    if (!b) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null batch"); }
    return engine_set_kernel_timing(&b->e, enable);
}

int emsph_batch_get_kernel_timing(emsph_batch* b, const char** names, double* total_ms,
                                  int64_t* calls, int max) {
    if (!b) { return 0; }
    return engine_get_kernel_timing(&b->e, names, total_ms, calls, max);
}

//------------------------------------------------------------------------------
// Batches of scenes stepped together.
emsph_status emsph_batch_create(int n_scenes, const emsph_config* configs,
                                const float* const* solid_xy, const size_t* solid_counts,
                                const size_t* fluid_counts, emsph_batch** out) {
    if (n_scenes <= 0 || !configs || !fluid_counts || !out) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument");
    }
    size_t total = 0;
    for (int s = 0; s < n_scenes; ++s) { total += fluid_counts[s]; }
    if (total > (1ull << 30)) { return fail(EMSPH_ERR_CAPACITY, "batch above 2^30 particles"); }
    auto* b = new emsph_batch();
    emsph_status st = engine_init(&b->e);
    if (st) {
        delete b;
        return st;
    }
    std::vector<SceneHost> hs(n_scenes);
    std::vector<float const*> sx(n_scenes, nullptr);
    for (int s = 0; s < n_scenes; ++s) {
        hs[s].cfg = configs[s];
        hs[s].cap = static_cast<uint32_t>(fluid_counts[s]);
        hs[s].n = static_cast<uint32_t>(fluid_counts[s]);
        hs[s].ns = solid_counts ? static_cast<uint32_t>(solid_counts[s]) : 0u;
        hs[s].dt = configs[s].seconds_per_sub_step;
        sx[s] = solid_xy ? solid_xy[s] : nullptr;
        if (hs[s].ns && !sx[s]) {
            delete b;
            return fail(EMSPH_ERR_INVALID_ARGUMENT, "solid positions missing");
        }
    }
    st = engine_configure(&b->e, std::move(hs), sx);
    if (st) {
        delete b;
        return st;
    }
    engine_set_skin(&b->e, 1);  // skin candidate lists on by default
    *out = b;
    return EMSPH_OK;
}

emsph_status emsph_batch_destroy(emsph_batch* b) {
    if (b) {
        (void)hipStreamSynchronize(b->e.stream);
        delete b;
    }
    return EMSPH_OK;
}

int emsph_batch_scene_count(const emsph_batch* b) { return b ? static_cast<int>(b->e.M) : 0; }

emsph_status emsph_batch_set_state(emsph_batch* b, int scene, const float* pos_xy,
                                   const float* vel_xy) {
    if (!b || scene < 0 || scene >= static_cast<int>(b->e.M) || !pos_xy) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument");
    }
    return engine_set_state(&b->e, scene, pos_xy, vel_xy, b->e.sh[scene].cap);
}

emsph_status emsph_batch_get_state(emsph_batch* b, int scene, float* pos_xy, float* vel_xy) {
    if (!b || scene < 0 || scene >= static_cast<int>(b->e.M)) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument");
    }
    return engine_get_state(&b->e, scene, pos_xy, vel_xy);
}

emsph_status emsph_batch_get_field(emsph_batch* b, int scene, int field, float* out) {
    if (!b || scene < 0 || scene >= static_cast<int>(b->e.M) || !out) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument");
    }
    return engine_get_field(&b->e, scene, field, out);
}

emsph_status emsph_batch_set_cfl(emsph_batch* b, float factor) {
    if (!b) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null batch"); }
    return engine_set_cfl(&b->e, factor);
}

emsph_status emsph_batch_frame_init(emsph_batch* b, int method) {
    if (!b) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null batch"); }
    return engine_frame_init(&b->e, method);
}

emsph_status emsph_batch_sub_step(emsph_batch* b, int method, emsph_step_stats* stats) {
    if (!b) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null batch"); }
    ST_TRY(engine_sub_step(&b->e, method));
    if (stats) {
        ST_TRY(read_stats(&b->e));
        for (uint32_t s = 0; s < b->e.M; ++s) { to_public(b->e.host_stats[s], &b->e, &stats[s]); }
    }
    return EMSPH_OK;
}

emsph_status emsph_batch_run_sub_steps(emsph_batch* b, int method, int n_sub_steps,
                                       float* elapsed_ms) {
    if (!b || n_sub_steps < 0) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument"); }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, b->e.stream));
    for (int k = 0; k < n_sub_steps; ++k) {
        emsph_status const s = engine_sub_step(&b->e, method);
        if (s) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return s;
        }
    }
    HIP_TRY(hipEventRecord(e1, b->e.stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) { *elapsed_ms = ms; }
    return EMSPH_OK;
}

//------------------------------------------------------------------------------
// Stand-alone search: one scene, no solids, cut-off = cell size.
emsph_status emsph_search_create(size_t capacity, emsph_search** out) {
    if (!out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    if (capacity > (1ull << 30)) { return fail(EMSPH_ERR_CAPACITY, "capacity above 2^30"); }
    auto* s = new emsph_search();
    emsph_status st = engine_init(&s->e);
    if (st) {
        delete s;
        return st;
    }
    SceneHost h;
    emsph_default_parameters(&h.cfg.params);
    h.cfg.mass_per_particle = 1.0f;
    h.cap = static_cast<uint32_t>(std::max<size_t>(capacity, 1));
    h.n = 0;
    h.ns = 0;
    st = engine_configure(&s->e, {h}, {nullptr});
    if (st) {
        delete s;
        return st;
    }
    *out = s;
    return EMSPH_OK;
}

emsph_status emsph_search_destroy(emsph_search* s) {
    if (s) {
        (void)hipStreamSynchronize(s->e.stream);
        delete s;
    }
    return EMSPH_OK;
}

emsph_status emsph_search_set_points(emsph_search* s, const float* xy, size_t count) {
    if (!s || (count && !xy)) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    s->ran = false;
    return engine_set_state(&s->e, 0, xy, nullptr, count);
}

emsph_status emsph_search_run(emsph_search* s, float cell_size, int store_geometry,
                              float* elapsed_ms) {
    if (!s || !(cell_size > 0.0f)) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument"); }
    Engine* g = &s->e;
    g->store_geometry = store_geometry != 0;
    if (g->store_geometry) { HIP_TRY(ensure_geometry(g)); }
    if (s->cell != cell_size) {
        // The scene's cut-off (2h) is the requested cell size.
        g->sh[0].cfg.params.support = cell_size * 0.5f;
        ST_TRY(upload_layout(g));
        g->N = g->sh[0].n;
        s->cell = cell_size;
    }
    // After a first run, the next one sorts with that run's key range (the
    // dense cell index when its table fits) and is redone exactly, inside the
    // timed region, when the range no longer fits.
    int passes = -1;
    int dense = 0;
    if (s->ran && g->passes_hint >= 0) {
        passes = g->passes_hint;
        dense_prepare(g);
        dense = dense_for(g);
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, g->stream));
    for (int attempt = 0; attempt < 2 && g->N; ++attempt) {
        IndexArgs ga{};
        ga.pos_out = g->spos.p;
        HIP_TRY(index_points(g, g->fidx, g->pos.p, g->N, g->scene_of.p, 1,
                             attempt ? -1 : passes, ga, false, nullptr, nullptr,
                             attempt ? 0 : dense));
        Batch b = fluid_batch(g);
        b.fl.gw = nullptr;
        launch_build(BUILD_LISTS, g->build_cap, b, g->stream);
        HIP_TRY(hipGetLastError());
        g->have_lists = true;
        // The key state of this run: list capacity, pass count and dense
        // table of the next run, and whether this one was short.
        g->keys_checked = false;
        HIP_TRY(read_keystate(g));
        if (!g->sort_short) { break; }
    }
    HIP_TRY(hipEventRecord(e1, g->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) { *elapsed_ms = ms; }
    s->ran = true;
    return EMSPH_OK;
}

emsph_status emsph_search_get(emsph_search* s, int32_t* grid_xy, uint64_t* keys,
                              uint64_t* sorted_keys, uint32_t* sorted_idx,
                              uint32_t* block_indices, uint32_t* block_begin,
                              uint32_t* block_end, size_t* block_count, uint8_t* counts,
                              uint32_t* indices, float* distances, float* vectors_xy) {
    if (!s || !s->ran) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "search not run"); }
    Engine* g = &s->e;
    uint32_t const n = g->N;
    if (n == 0) {
        if (block_count) { *block_count = 0; }
        return EMSPH_OK;
    }
    hipStream_t const st = g->stream;
    // Block count = last block index + 1 (neighborhoods_from_state.cpp:33).
    uint32_t last = 0;
    HIP_TRY(hipMemcpyAsync(&last, g->fidx.block_index.p + (n - 1), 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    uint32_t const nb = last + 1;
    if (block_count) { *block_count = nb; }
    if (grid_xy || keys) {
        // Recompute the keys in original order (the sort's buffers were reused).
        IndexArgs ka{};
        ka.pos = g->pos.p;
        ka.n = n;
        ka.scenes = g->scenes.p;
        ka.scene_of = g->scene_of.p;
        DBuf<uint64_t> k2;
        DBuf<uint32_t> v2;
        DBuf<int2> g2;
        HIP_TRY(k2.alloc(n));
        HIP_TRY(v2.alloc(n));
        HIP_TRY(g2.alloc(n));
        ka.keys = k2.p;
        ka.vals = v2.p;
        ka.grid = g2.p;
        ka.keystate = g->keystate.p;
        ka.key_partials = g->key_partials.p;
        launch_keys(ka, st);
        HIP_TRY(hipGetLastError());
        if (keys) { HIP_TRY(hipMemcpyAsync(keys, k2.p, n * 8, hipMemcpyDeviceToHost, st)); }
        if (grid_xy) { HIP_TRY(hipMemcpyAsync(grid_xy, g2.p, n * 8, hipMemcpyDeviceToHost, st)); }
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (sorted_keys) {
        HIP_TRY(hipMemcpyAsync(sorted_keys, g->fidx.sorted_keys, n * 8, hipMemcpyDeviceToHost, st));
    }
    if (sorted_idx) {
        HIP_TRY(hipMemcpyAsync(sorted_idx, g->fidx.orig.p, n * 4, hipMemcpyDeviceToHost, st));
    }
    if (block_indices) {
        HIP_TRY(hipMemcpyAsync(block_indices, g->fidx.block_index.p, n * 4, hipMemcpyDeviceToHost,
                               st));
    }
    if (block_begin) {
        HIP_TRY(hipMemcpyAsync(block_begin, g->fidx.cell_begin.p, nb * 4, hipMemcpyDeviceToHost,
                               st));
    }
    if (block_end) {
        HIP_TRY(hipMemcpyAsync(block_end, g->fidx.cell_end.p, nb * 4, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (counts && indices) {
        if ((distances || vectors_xy) && !g->store_geometry) {
            return fail(EMSPH_ERR_INVALID_ARGUMENT, "geometry not stored by emsph_search_run");
        }
        ST_TRY(engine_get_neighborhood(g, 0, 0, counts, indices, distances, vectors_xy));
    }
    return EMSPH_OK;
}

//------------------------------------------------------------------------------
// Pools (the served form).
int64_t emsph_device_bytes(void) { return g_device_bytes.load(); }

emsph_status emsph_pool_create(size_t fluid_capacity, size_t solid_capacity, int max_scenes,
                               int method, emsph_pool** out) {
    if (!out || fluid_capacity == 0 || max_scenes <= 0 || fluid_capacity > (1ull << 30) ||
        solid_capacity > (1ull << 30)) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument");
    }
    ST_TRY(check_method(method));
    auto* P = new emsph_pool();
    P->method = method;
    P->fluid_cap = static_cast<uint32_t>(fluid_capacity);
    P->solid_cap = static_cast<uint32_t>(solid_capacity);
    P->max_scenes = static_cast<uint32_t>(max_scenes);
    Engine* g = &P->e;
    emsph_status st = engine_init(g);
    if (st) {
        delete P;
        return st;
    }
    hipError_t e = alloc_fluid(g, P->fluid_cap);
    if (!e && solid_capacity) {
        uint32_t const NS = P->solid_cap;
        if (!e) { e = g->sidx.ensure(NS); }
        if (!e) { e = g->s_pos.ensure(NS); }
        if (!e) { e = g->s_vel.ensure(NS); }
        if (!e) { e = g->s_pos_orig.ensure(NS); }
        if (!e) { e = g->s_vol.ensure(NS); }
        if (!e) { e = g->s_scene_of.ensure(NS); }
    }
    if (e) {
        delete P;
        return fail(e == hipErrorOutOfMemory ? EMSPH_ERR_OUT_OF_MEMORY : EMSPH_ERR_HIP,
                    std::string("pool allocation: ") + hipGetErrorString(e));
    }
    g->M = 0;
    g->N = 0;
    g->keep_solid_scratch = true;
    st = upload_layout(g);
    if (st) {
        delete P;
        return st;
    }
    *out = P;
    return EMSPH_OK;
}

emsph_status emsph_pool_destroy(emsph_pool* P) {
    if (P) {
        (void)hipStreamSynchronize(P->e.stream);
        delete P;
    }
    return EMSPH_OK;
}

emsph_status emsph_pool_add(emsph_pool* P, const emsph_config* config, const float* solid_xy,
                            size_t solid_count, const float* pos_xy, const float* vel_xy,
                            size_t count, int* id) {
    if (!P || !config || !pos_xy || count == 0 || (solid_count && !solid_xy)) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument");
    }
    uint64_t n, ns, m;
    pool_projection(P, n, ns, m);
    if (n + count > P->fluid_cap || ns + solid_count > P->solid_cap || m + 1 > P->max_scenes) {
        return fail(EMSPH_ERR_CAPACITY, "pool capacity");
    }
    PoolScene sc;
    sc.id = P->next_id++;
    sc.cfg = *config;
    sc.n = static_cast<uint32_t>(count);
    sc.ns = static_cast<uint32_t>(solid_count);
    sc.solids.assign(solid_xy, solid_xy + 2 * solid_count);
    sc.pos.assign(pos_xy, pos_xy + 2 * count);
    if (vel_xy) { sc.vel.assign(vel_xy, vel_xy + 2 * count); }
    if (id) { *id = sc.id; }
    P->adds.push_back(std::move(sc));
    return EMSPH_OK;
}

emsph_status emsph_pool_remove(emsph_pool* P, int id) {
    if (!P) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null pool"); }
    for (size_t k = 0; k < P->adds.size(); ++k) {
        if (P->adds[k].id == id) {
            P->adds.erase(P->adds.begin() + static_cast<std::ptrdiff_t>(k));
            return EMSPH_OK;
        }
    }
    int const k = pool_index(P, id);
    if (k < 0) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no such scene"); }
    if (std::find(P->removes.begin(), P->removes.end(), id) == P->removes.end()) {
        P->removes.push_back(id);
        // Until the commit the scene still rides along in the sort and the
        // neighbour build, but no longer in the solver iterations.
        Engine* g = &P->e;
        SceneHost& h = g->sh[static_cast<size_t>(k)];
        h.retired = true;
        uint32_t first_tile = 0;
        for (int j = 0; j < k; ++j) { first_tile += (g->sh[static_cast<size_t>(j)].n + TILE - 1) / TILE; }
        SceneDev const sd = scene_dev(h, first_tile);
        HIP_TRY(hipSetDevice(g->device));
        HIP_TRY(hipMemcpyAsync(g->scenes.p + k, &sd, sizeof(SceneDev), hipMemcpyHostToDevice,
                               g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));  // sd lives on this stack frame
    }
    return EMSPH_OK;
}

emsph_status emsph_pool_commit(emsph_pool* P) {
    if (!P) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null pool"); }
    return pool_commit(P);
}

int emsph_pool_live(const emsph_pool* P, int* ids, int max) {
    if (!P) { return 0; }
    int const k = static_cast<int>(P->live.size());
    for (int i = 0; ids && i < k && i < max; ++i) { ids[i] = P->live[i].id; }
    return k;
}

emsph_status emsph_pool_set_cfl(emsph_pool* P, float factor) {
    if (!P) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null pool"); }
    if (!(factor >= 0.0f)) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "CFL factor < 0"); }
    P->cfl = factor;
    return engine_set_cfl(&P->e, factor);
}

emsph_status emsph_pool_sub_step(emsph_pool* P, emsph_step_stats* stats) {
    if (!P) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null pool"); }
    Engine* g = &P->e;
    if (!g->M) { return EMSPH_OK; }
    ST_TRY(engine_sub_step(g, P->method));
    if (stats) {
        ST_TRY(read_stats(g));
        for (uint32_t s = 0; s < g->M; ++s) { to_public(g->host_stats[s], g, &stats[s]); }
    }
    return EMSPH_OK;
}

emsph_status emsph_pool_get_state(emsph_pool* P, int id, float* pos_xy, float* vel_xy) {
    if (!P) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null pool"); }
    int const k = pool_index(P, id);
    if (k < 0) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no such live scene"); }
    return engine_get_state(&P->e, static_cast<uint32_t>(k), pos_xy, vel_xy);
}

//------------------------------------------------------------------------------
// Snapshots of device state (format and encoding in snapshot.cpp).

namespace {

emsph_status save_state(Engine* g, uint32_t s, const char* path, int64_t time_flicks,
                        int32_t method, bool with_colors) {
    if (!path) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null path"); }
    size_t const n = g->sh[s].n;
    std::vector<float> pos(2 * n), vel(2 * n), rgba;
    ST_TRY(engine_get_state(g, s, pos.data(), vel.data()));
    if (with_colors && n) {
        rgba.resize(4 * n);
        HIP_TRY(hipMemcpyAsync(rgba.data(), g->colors.p, n * sizeof(float4),
                               hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    std::vector<unsigned char> buf(emsph_snapshot_bytes(n, with_colors ? 3u : 1u));
    size_t written = 0;
    ST_TRY(emsph_snapshot_encode(pos.data(), vel.data(), with_colors ? rgba.data() : nullptr, n,
                                 time_flicks, method, buf.data(), buf.size(), &written));
    return emsph_snapshot_write_file(path, buf.data(), written);
}

}  // namespace

emsph_status emsph_save_snapshot(emsph_context* c, const char* path, int64_t time_flicks,
                                 int32_t method) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    return save_state(&c->e, 0, path, time_flicks, method, c->e.colors_valid);
}

emsph_status emsph_load_snapshot(emsph_context* c, const char* path, int64_t* time_flicks,
                                 int32_t* method) {
    if (!c || !path) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    std::vector<unsigned char> buf;
    if (!emsph_snapshot_read_file(path, buf)) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, std::string("cannot read ") + path);
    }
    emsph_snapshot_header h{};
    ST_TRY(emsph_snapshot_info(buf.data(), buf.size(), &h));
    size_t const n = static_cast<size_t>(h.count);
    std::vector<float> pos(2 * n), vel(2 * n);
    ST_TRY(emsph_snapshot_decode(buf.data(), buf.size(), pos.data(), vel.data(), nullptr));
    ST_TRY(engine_set_state(&c->e, 0, pos.data(), vel.data(), n));
    if (time_flicks) { *time_flicks = h.time_flicks; }
    if (method) { *method = h.method; }
    return EMSPH_OK;
}

emsph_status emsph_pool_save_snapshot(emsph_pool* P, int id, const char* path,
                                      int64_t time_flicks) {
    if (!P) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null pool"); }
    int const k = pool_index(P, id);
    if (k < 0) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no such live scene"); }
    return save_state(&P->e, static_cast<uint32_t>(k), path, time_flicks, P->method, false);
}

}  // extern "C"

//------------------------------------------------------------------------------
// Neighbour lists without allocation: the stand-alone search
// (emsph_search_run + emsph_search_get's lists) for callers that cannot
// allocate device memory, on caller-owned device buffers. Every intermediate
// buffer of the search (keys, sort passes, cell tables, cell windows, sorted
// positions, slot-major lists, scene and tile tables, key state) lives in one
// file-scope device array, carved at fixed offsets for up to
// NOALLOC_MAX_POINTS points. Each GPU holds its own copy of the array until the
// process ends; its address on a device is taken once, with that device
// current, by hipGetSymbolAddress, and only that address is passed to kernels
// (the array's name is not a device address in host code). Calls on one GPU
// share the array: calls that overlap on a device race.
namespace {

constexpr uint32_t NOALLOC_MAX_POINTS = 65536;
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t{32} << 20;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Scratch address of the current device (taken on the first call there).
emsph_status noalloc_scratch(int device, unsigned char** out) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "device index above the supported range");
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            return fail(EMSPH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
        }
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
    return EMSPH_OK;
}

// Fixed carving of the scratch for n points: each buffer 256-byte aligned and
// followed by DBuf's tail padding (the build reads up to 3 positions past a
// cell's range).
struct NoallocLayout {
    size_t at = 0;
    unsigned char* base = nullptr;
    template <class T>
    T* take(size_t count) {
        at = (at + 255) & ~size_t{255};
        T* p = reinterpret_cast<T*>(base ? base + at : nullptr);
        at += (count + DBuf<T>::TAIL_PAD) * sizeof(T);
        return p;
    }
};

struct NoallocBuffers {
    SceneDev* scenes;
    Tile* tiles;
    uint32_t* scene_of;
    KeyState* keystate;
    unsigned long long* key_partials;
    uint64_t *keys0, *keys1, *cell_key;
    uint32_t *vals0, *vals1, *counts, *scan, *heads, *head_scan, *block_index, *cell_begin,
      *cell_end, *cell_tag, *ncells, *orig, *rank, *fl_idx;
    uint2* win;
    uint2* overflow;
    V2* spos;
    uint8_t* fcount;
};

size_t noalloc_carve(unsigned char* base, uint32_t n, uint32_t stride, NoallocBuffers* b) {
    NoallocLayout L;
    L.base = base;
    uint32_t const ntiles = (n + TILE - 1) / TILE;
    b->scenes = L.take<SceneDev>(1);
    b->tiles = L.take<Tile>(ntiles ? ntiles : 1);
    b->scene_of = L.take<uint32_t>(n);
    b->keystate = L.take<KeyState>(1);
    b->key_partials = L.take<unsigned long long>(2 * KEY_PARTS);
    b->keys0 = L.take<uint64_t>(n);
    b->keys1 = L.take<uint64_t>(n);
    b->vals0 = L.take<uint32_t>(n);
    b->vals1 = L.take<uint32_t>(n);
    b->counts = L.take<uint32_t>(radix_counts_words(n) + 256);
    b->scan = L.take<uint32_t>(scan_scratch_words(radix_counts_words(n) + n) + 64);
    b->heads = L.take<uint32_t>(n);
    b->head_scan = L.take<uint32_t>(n);
    b->block_index = L.take<uint32_t>(n);
    b->cell_begin = L.take<uint32_t>(n);
    b->cell_end = L.take<uint32_t>(n);
    b->cell_tag = L.take<uint32_t>(n);
    b->cell_key = L.take<uint64_t>(n);
    b->ncells = L.take<uint32_t>(1);
    b->orig = L.take<uint32_t>(n);
    b->rank = L.take<uint32_t>(n);
    b->win = L.take<uint2>(static_cast<size_t>(n) * 9);
    b->overflow = L.take<uint2>(n);
    b->spos = L.take<V2>(n);
    b->fcount = L.take<uint8_t>(stride);
    b->fl_idx = L.take<uint32_t>(static_cast<size_t>(MAXN) * stride);
    return L.at;
}

// The scene entry and the tile table of one scene of n points, written on the
// device from kernel arguments (no host copy).
__global__ __launch_bounds__(256) void k_noalloc_setup(SceneDev sd, SceneDev* scenes, Tile* tiles,
                                                     uint32_t n, uint32_t ntiles) {
    uint32_t const k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) { scenes[0] = sd; }
    if (k < ntiles) {
        uint32_t const first = k * TILE;
        tiles[k] = Tile{0u, first, std::min<uint32_t>(TILE, n - first), 0u};
    }
}

}  // namespace

extern "C" {

emsph_status emsph_search_lists_noalloc(const float* xy_dev, size_t count, float cell_size,
                                        uint8_t* counts_dev, uint32_t* indices_dev,
                                        void* stream) {
    if (count > NOALLOC_MAX_POINTS) {
        return fail(EMSPH_ERR_CAPACITY, "emsph_search_lists_noalloc: " + std::to_string(count) +
                                          " points above the maximum of " +
                                          std::to_string(NOALLOC_MAX_POINTS));
    }
    if (!(cell_size > 0.0f)) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "cell size <= 0"); }
    if (count == 0) { return EMSPH_OK; }
    if (!xy_dev || !counts_dev || !indices_dev) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "null buffer");
    }
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    unsigned char* base = nullptr;
    ST_TRY(noalloc_scratch(device, &base));
    auto const n = static_cast<uint32_t>(count);
    uint32_t const stride = round_up(n, 256);
    uint32_t const ntiles = (n + TILE - 1) / TILE;
    NoallocBuffers B{};
    if (noalloc_carve(nullptr, NOALLOC_MAX_POINTS, round_up(NOALLOC_MAX_POINTS, 256), &B) >
        NOALLOC_SCRATCH_BYTES) {
        return fail(EMSPH_ERR_CAPACITY, "scratch layout exceeds the scratch array");
    }
    noalloc_carve(base, n, stride, &B);
    auto const s = static_cast<hipStream_t>(stream);

    // The scene: one point set, cut-off and cell = cell_size (2h), as
    // emsph_search_run configures it.
    SceneHost h;
    emsph_default_parameters(&h.cfg.params);
    h.cfg.mass_per_particle = 1.0f;
    h.cfg.params.support = cell_size * 0.5f;
    h.cap = n;
    h.n = n;
    SceneDev const sd = scene_dev(h, 0);
    hipLaunchKernelGGL(k_noalloc_setup, dim3((ntiles + 255) / 256), dim3(256), 0, s, sd, B.scenes,
                       B.tiles, n, ntiles);
    launch_scene_of(B.tiles, ntiles, B.scene_of, s);

    // Keys, all eight 8-bit digits of the sort (no key range read back),
    // cells and cell windows (neighborhoods_from_state.cpp:14-40).
    IndexArgs a{};
    a.pos = reinterpret_cast<V2 const*>(xy_dev);
    a.n = n;
    a.scenes = B.scenes;
    a.scene_of = B.scene_of;
    a.keys = B.keys0;
    a.vals = B.vals0;
    a.keystate = B.keystate;
    a.key_partials = B.key_partials;
    a.pos_out = B.spos;
    launch_keys(a, s);
    int const which = launch_radix_sort(B.keys0, B.vals0, B.keys1, B.vals1, n, 8, B.scene_of, 0,
                                        B.counts, B.scan, s);
    a.key_passes = 8;
    a.check_keys = 0;
    a.sorted_keys = which ? B.keys1 : B.keys0;
    a.sorted_vals = which ? B.vals1 : B.vals0;
    a.heads = B.heads;
    a.head_scan = B.head_scan;
    a.block_index = B.block_index;
    a.cell_begin = B.cell_begin;
    a.cell_end = B.cell_end;
    a.cell_key = B.cell_key;
    a.cell_tag = B.cell_tag;
    a.orig = B.orig;
    a.rank = B.rank;
    launch_cell_heads(a, s);
    launch_exclusive_scan(B.heads, B.head_scan, n, B.scan, B.ncells, s);
    launch_cells_finalize(a, s);
    WindowArgs w{};
    w.self = CellList{B.cell_key, B.cell_tag, B.cell_begin, B.cell_end, B.ncells};
    w.max_cells = n;
    w.win = B.win;
    w.report = &B.keystate->cells;
    launch_cell_windows(w, s);

    // 64-capped lists in reference order (neighborhood.h:198-282), then in
    // original order into the caller's buffers.
    Batch b{};
    b.scenes = B.scenes;
    b.tiles = B.tiles;
    b.ntiles = ntiles;
    b.nscenes = 1;
    b.stride = stride;
    b.keystate = B.keystate;
    b.pos = B.spos;
    b.orig = B.orig;
    b.rank = B.rank;
    b.cell_of = B.block_index;
    b.win = B.win;
    b.overflow = B.overflow;
    b.fl = Lists{B.fcount, B.fl_idx, nullptr, nullptr, nullptr};
    b.nb = 4;
    b.xcd_remap = 1;
    launch_build(BUILD_LISTS, MAXN, b, s);
    launch_export_lists(b.fl, n, stride, B.orig, B.orig, B.scene_of, B.scenes, 0, counts_dev,
                        indices_dev, nullptr, nullptr, s);
    HIP_TRY(hipGetLastError());
    return EMSPH_OK;
}

}  // extern "C"
