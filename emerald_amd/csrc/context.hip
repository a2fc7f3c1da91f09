// Host orchestration of the 2D SPH sub-step on one gfx950 GPU and the C ABI
// of include/emerald_sph.h.
//
// One sub-step (the reference's <method>_sub_step, e.g. iisph.cpp:23-151):
//   1. keys      grid coords + 64-bit Morton keys of the fluid, in original order
//   2. sort      stable LSD radix sort of (key, index) over the significant digits
//   3. cells     run heads, exclusive scan -> block indices, cell ranges, cell
//                table; gather of the per-particle state into sorted order
//   4. build     64-capped neighbour lists in reference order (fluid and solid),
//                with the first per-particle sums of the method fused in
//   5. solver    method sweeps; each iteration's residual reduction is read back
//                for the reference's host-side convergence branch
//   6. finalize  x += dt v, state scattered back to original order
// Solids are static: their cell table, lists and volumes are built once at
// context creation (solids.cpp:76-137) with the same kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
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

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t const e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                          \
            return fail(EMSPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                \
    } while (0)

constexpr int64_t kFlicksPerSecond = 705600000;
constexpr float kPi = 3.14159274101257324218750f;

double to_seconds(int64_t f) {
    return static_cast<double>(f) / static_cast<double>(kFlicksPerSecond);
}
int64_t to_flicks(double s) {
    return static_cast<int64_t>(s * static_cast<double>(kFlicksPerSecond));
}

KernelConsts make_kernel_consts(float h) {
    KernelConsts c;
    c.h = h;
    c.hh = h * h;
    c.hhh = h * h * h;
    c.k = h * h * h * kPi;
    c.four_h = 4.0f * h;
    // W(0, h) = _wa(0) / k = 1 / k (kernels.h:25-35)
    float const q = 0.0f / h;
    c.w0 = (1.0f + q * q * (q * 0.75f - 1.5f)) / c.k;
    return c;
}

uint32_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) { p <<= 1; }
    return static_cast<uint32_t>(p);
}

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
        if (p) { (void)hipFree(p); }
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        if (count == 0) { count = 1; }
        hipError_t const e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e == hipSuccess) { n = count; }
        return e;
    }
};

// Keys, sort and cell table of one point set.
struct SpatialIndex {
    uint32_t cap = 0;        // point capacity
    uint32_t table_cap = 0;  // power of two
    DBuf<uint64_t> keys0, keys1;
    DBuf<uint32_t> vals0, vals1;
    DBuf<uint32_t> counts, scan_scratch;
    DBuf<uint32_t> heads, head_scan, block_index, cell_begin, cell_end;
    DBuf<uint64_t> table_keys;
    DBuf<uint32_t> table_vals;
    DBuf<uint32_t> orig, rank;
    DBuf<int2> grid;
    uint64_t* sorted_keys = nullptr;  // points into keys0/keys1
    uint32_t* sorted_vals = nullptr;
    int passes = 0;

    hipError_t alloc(uint32_t c, bool with_grid) {
        cap = c;
        table_cap = next_pow2(std::max<uint64_t>(2ull * c, 1024));
        hipError_t e;
        if ((e = keys0.alloc(c)) || (e = keys1.alloc(c)) || (e = vals0.alloc(c)) ||
            (e = vals1.alloc(c)) || (e = counts.alloc(radix_counts_words(c) + 256)) ||
            (e = scan_scratch.alloc(scan_scratch_words(radix_counts_words(c) + c) + 64)) ||
            (e = heads.alloc(c)) || (e = head_scan.alloc(c)) || (e = block_index.alloc(c)) ||
            (e = cell_begin.alloc(c)) || (e = cell_end.alloc(c)) ||
            (e = table_keys.alloc(table_cap)) || (e = table_vals.alloc(table_cap)) ||
            (e = orig.alloc(c)) || (e = rank.alloc(c))) {
            return e;
        }
        if (with_grid) { return grid.alloc(c); }
        return hipSuccess;
    }

    CellTable table() const {
        return CellTable{table_keys.p, table_vals.p, cell_begin.p, cell_end.p, table_cap - 1};
    }
};

struct KernelTimer {
    bool enabled = false;
    std::map<std::string, std::pair<double, int64_t>> acc;
    hipEvent_t a = nullptr, b = nullptr;
};

// Small pinned read-back block.
struct Pinned {
    KeyState keys_reset;   // source of the per-sort KeyState reset
    KeyState keys;         // read back
    Reduction red[MAX_RED_SLOTS];
    unsigned int vmax2;
    unsigned int cells;
};

}  // namespace

struct emsph_context {
    emsph_config cfg{};
    hipStream_t stream = nullptr;
    int device = 0;
    uint32_t cap = 0;
    uint32_t stride = 0;
    uint32_t n = 0;
    bool store_geometry = false;
    bool have_lists = false;
    int last_method = -1;

    // Fluid state in original order.
    DBuf<V2> pos, vel;
    DBuf<float> kv, kd;  // DFSPH warm-start stiffnesses (original order)

    SpatialIndex fidx;

    // Sorted-space fields.
    DBuf<V2> spos, svel, dii, sdp, vext, xstar, vstar0, vstar1, force;
    DBuf<float> rho, rho_star, aii, p0, p1, alpha, kvs, kds;
    float* cur_p = nullptr;  // pressure buffer holding the last sub-step's result

    // Neighbour lists (slot-major).
    DBuf<uint8_t> fcount, scount;
    DBuf<uint32_t> fl_idx, sl_idx;
    DBuf<V2> fl_gw, sl_gw, fl_vec, sl_vec;
    DBuf<float> fl_dist, sl_dist;

    // Solids.
    uint32_t ns = 0;
    SpatialIndex sidx;
    DBuf<V2> s_pos, s_vel, s_pos_orig;
    DBuf<float> s_vol;

    // Reductions.
    DBuf<Reduction> red;
    DBuf<KeyState> keystate;
    DBuf<unsigned int> vmax2;
    DBuf<unsigned int> cells_total;
    Pinned* host = nullptr;
    // Radix passes of the previous sort (-1: read the key range before sorting).
    int passes_hint = -1;
    bool keys_checked = false;  // this sub-step's KeyState has been read back
    bool sort_short = false;    // ... and showed the hint was too small

    KernelTimer timer;

    ~emsph_context() {
        if (host) { (void)hipHostFree(host); }
        if (timer.a) { (void)hipEventDestroy(timer.a); }
        if (timer.b) { (void)hipEventDestroy(timer.b); }
        if (stream) { (void)hipStreamDestroy(stream); }
    }
};

struct emsph_search {
    hipStream_t stream = nullptr;
    uint32_t cap = 0;
    uint32_t stride = 0;
    uint32_t n = 0;
    float cell = 0.0f;
    bool ran = false;
    bool geometry = false;
    DBuf<V2> pos, spos;
    SpatialIndex idx;
    DBuf<uint8_t> count;
    DBuf<uint32_t> lidx;
    DBuf<float> dist;
    DBuf<V2> vec;
    DBuf<KeyState> keystate;
    DBuf<unsigned int> cells_total;
    Pinned* host = nullptr;
    ~emsph_search() {
        if (host) { (void)hipHostFree(host); }
        if (stream) { (void)hipStreamDestroy(stream); }
    }
};

namespace {

// Time a launch when kernel timing is enabled (measurement mode only).
template <class F>
hipError_t timed(emsph_context* c, char const* name, F&& launch) {
    if (!c->timer.enabled) {
        launch();
        return hipGetLastError();
    }
    (void)hipEventRecord(c->timer.a, c->stream);
    launch();
    (void)hipEventRecord(c->timer.b, c->stream);
    hipError_t const e = hipEventSynchronize(c->timer.b);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->timer.a, c->timer.b);
    auto& slot = c->timer.acc[name];
    slot.first += ms;
    slot.second += 1;
    return e;
}

// Keys -> sort -> cells of n points at pos (original order), plus gathers.
//
// passes_hint < 0 reads the key range back before sorting (one host sync) and
// sorts exactly the significant digits. passes_hint >= 0 sorts that many
// digits without waiting; k_cell_heads then flags (KeyState::sort_short) a key
// range the digits do not cover, and the caller must redo the work with the
// exact count before anything it produced is used.
hipError_t index_points(SpatialIndex& si, V2 const* pos, uint32_t n, float cell,
                        hipStream_t s, Pinned* host, KeyState* d_keys, int passes_hint,
                        unsigned int* d_cells, Reduction* red_reset, unsigned int* vmax2_reset,
                        CellsArgs gathers, emsph_context* timer_ctx) {
    auto run = [&](char const* name, auto&& f) -> hipError_t {
        if (timer_ctx) { return timed(timer_ctx, name, f); }
        f();
        return hipGetLastError();
    };
    hipError_t e;
    KeysArgs ka{};
    ka.pos = pos;
    ka.n = n;
    ka.cell = cell;
    ka.keys = si.keys0.p;
    ka.vals = si.vals0.p;
    ka.grid = si.grid.p;
    ka.keystate = d_keys;
    ka.red_reset = red_reset;
    ka.vmax2_reset = vmax2_reset;
    host->keys_reset.kmin = ~0ull;
    host->keys_reset.kmax = 0ull;
    host->keys_reset.sort_short = 0u;
    host->keys_reset.pad = 0u;
    if ((e = hipMemcpyAsync(d_keys, &host->keys_reset, sizeof(KeyState), hipMemcpyHostToDevice,
                            s))) {
        return e;
    }
    if ((e = run("keys", [&] { launch_keys(ka, s); }))) { return e; }
    if (passes_hint < 0) {
        if ((e = hipMemcpyAsync(&host->keys, d_keys, sizeof(KeyState), hipMemcpyDeviceToHost,
                                s))) {
            return e;
        }
        if ((e = hipStreamSynchronize(s))) { return e; }
        si.passes = radix_pass_count(host->keys.kmin, host->keys.kmax);
    } else {
        si.passes = passes_hint;
    }
    int which = 0;
    if ((e = run("radix_sort", [&] {
             which = launch_radix_sort(si.keys0.p, si.vals0.p, si.keys1.p, si.vals1.p, n,
                                       si.passes, si.counts.p, si.scan_scratch.p, s);
         }))) {
        return e;
    }
    si.sorted_keys = which ? si.keys1.p : si.keys0.p;
    si.sorted_vals = which ? si.vals1.p : si.vals0.p;

    CellsArgs ca = gathers;
    ca.sorted_keys = si.sorted_keys;
    ca.sorted_vals = si.sorted_vals;
    ca.n = n;
    ca.heads = si.heads.p;
    ca.head_scan = si.head_scan.p;
    ca.block_index = si.block_index.p;
    ca.cell_begin = si.cell_begin.p;
    ca.cell_end = si.cell_end.p;
    ca.table_keys = si.table_keys.p;
    ca.table_vals = si.table_vals.p;
    ca.mask = si.table_cap - 1;
    ca.orig = si.orig.p;
    ca.rank = si.rank.p;
    ca.keystate = (passes_hint >= 0) ? d_keys : nullptr;
    ca.passes = si.passes;
    if ((e = run("cells", [&] {
             launch_cell_heads(ca, s);
             launch_exclusive_scan(si.heads.p, si.head_scan.p, n, si.scan_scratch.p, d_cells, s);
             launch_fill_u32(si.table_vals.p, EMPTY_CELL, si.table_cap, s);
             launch_cells_finalize(ca, s);
         }))) {
        return e;
    }
    return hipSuccess;
}

StepConsts step_consts(emsph_context const* c, float dt) {
    StepConsts k{};
    emsph_parameters const& p = c->cfg.params;
    k.kc = make_kernel_consts(p.support);
    k.cell = p.support * 2.0f;
    k.rho0 = p.target_density;
    k.fluid_volume = c->cfg.mass_per_particle / p.target_density;
    k.mass = c->cfg.mass_per_particle;
    k.gravity = p.gravity;
    k.dt = dt;
    k.omega = p.iisph_omega;
    k.pci_denom = c->cfg.pressure_correction_denom;
    k.viscosity = p.viscosity;
    return k;
}

SolidView solid_view(emsph_context const* c) {
    SolidView v{};
    v.pos = c->s_pos.p;
    v.vel = c->s_vel.p;
    v.vol = c->s_vol.p;
    v.rank = c->sidx.rank.p;
    v.orig = c->sidx.orig.p;
    v.table = c->sidx.table();
    v.n = c->ns;
    return v;
}

Lists fluid_lists(emsph_context* c) {
    return Lists{c->fcount.p, c->fl_idx.p, c->fl_gw.p, c->fl_dist.p, c->fl_vec.p};
}
Lists solid_lists(emsph_context* c) {
    return Lists{c->scount.p, c->sl_idx.p, c->sl_gw.p, c->sl_dist.p, c->sl_vec.p};
}

hipError_t ensure_geometry(emsph_context* c) {
    hipError_t e = hipSuccess;
    size_t const m = static_cast<size_t>(MAXN) * c->stride;
    if (!c->fl_dist.p && (e = c->fl_dist.alloc(m))) { return e; }
    if (!c->fl_vec.p && (e = c->fl_vec.alloc(m))) { return e; }
    if (!c->sl_dist.p && (e = c->sl_dist.alloc(m))) { return e; }
    if (!c->sl_vec.p && (e = c->sl_vec.alloc(m))) { return e; }
    return e;
}

// Steps 1-4 of the sub-step.
emsph_status prepare(emsph_context* c, int build_mode, bool dfsph, float dt,
                     emsph_step_stats* st) {
    uint32_t const n = c->n;
    CellsArgs g{};
    g.pos_in = c->pos.p;
    g.vel_in = c->vel.p;
    g.pos_out = c->spos.p;
    g.vel_out = c->svel.p;
    if (dfsph) {
        g.kv_in = c->kv.p;
        g.kd_in = c->kd.p;
        g.kv_out = c->kvs.p;
        g.kd_out = c->kds.p;
    }
    c->keys_checked = c->passes_hint < 0;  // an exact sort needs no check
    c->sort_short = false;
    HIP_TRY(index_points(c->fidx, c->pos.p, n, c->cfg.params.support * 2.0f, c->stream, c->host,
                         c->keystate.p, c->passes_hint, c->cells_total.p, c->red.p, c->vmax2.p,
                         g, c->timer.enabled ? c : nullptr));
    if (c->passes_hint < 0) { c->passes_hint = c->fidx.passes; }
    BuildArgs b{};
    b.c = step_consts(c, dt);
    b.n = n;
    b.stride = c->stride;
    b.pos = c->spos.p;
    b.vel = c->svel.p;
    b.orig = c->fidx.orig.p;
    b.rank = c->fidx.rank.p;
    b.table = c->fidx.table();
    b.solids = solid_view(c);
    b.fl = fluid_lists(c);
    b.sl = solid_lists(c);
    b.rho = c->rho.p;
    b.dii = c->dii.p;
    b.alpha = c->alpha.p;
    b.store_geometry = c->store_geometry ? 1 : 0;
    HIP_TRY(timed(c, "build", [&] { launch_build(build_mode, b, c->stream); }));
    c->have_lists = true;
    if (st) { st->sort_passes = static_cast<uint32_t>(c->fidx.passes); }
    return EMSPH_OK;
}

SweepArgs sweep_args(emsph_context* c, float dt) {
    SweepArgs a{};
    a.c = step_consts(c, dt);
    a.n = c->n;
    a.stride = c->stride;
    a.pos = c->spos.p;
    a.vel = c->svel.p;
    a.solids = solid_view(c);
    a.fl = fluid_lists(c);
    a.sl = solid_lists(c);
    a.rho = c->rho.p;
    a.rho_star = c->rho_star.p;
    a.aii = c->aii.p;
    a.dii = c->dii.p;
    a.sdp = c->sdp.p;
    a.alpha = c->alpha.p;
    a.red = c->red.p;
    a.vmax2_bits = c->vmax2.p;
    a.orig = c->fidx.orig.p;
    a.pos_orig = c->pos.p;
    a.vel_orig = c->vel.p;
    a.kappa_v_sorted = c->kvs.p;
    a.kappa_d_sorted = c->kds.p;
    a.kappa_v_orig = c->kv.p;
    a.kappa_d_orig = c->kd.p;
    a.vext = c->vext.p;
    a.xstar = c->xstar.p;
    a.force = c->force.p;
    return a;
}

// Read one residual slot: {average, max} as the reference computes them
// (iisph_ops.cpp:422-424, common.h:228-231).
// The first read-back of a sub-step also fetches its KeyState: when the sort's
// pass hint was too small, c->sort_short is set and the caller abandons the
// sub-step (nothing in original order has been written yet) to redo it.
hipError_t fetch_keystate_async(emsph_context* c) {
    if (c->keys_checked) { return hipSuccess; }
    return hipMemcpyAsync(&c->host->keys, c->keystate.p, sizeof(KeyState),
                          hipMemcpyDeviceToHost, c->stream);
}
void note_keystate(emsph_context* c) {
    if (c->keys_checked) { return; }
    c->keys_checked = true;
    int const exact = radix_pass_count(c->host->keys.kmin, c->host->keys.kmax);
    c->sort_short = c->host->keys.sort_short != 0u;
    // Next sub-step: the exact count of this one (keys move little per step).
    c->passes_hint = exact;
}
hipError_t check_keys(emsph_context* c) {
    if (c->keys_checked) { return hipSuccess; }
    hipError_t e = fetch_keystate_async(c);
    if (e) { return e; }
    if ((e = hipStreamSynchronize(c->stream))) { return e; }
    note_keystate(c);
    return hipSuccess;
}

hipError_t read_residual(emsph_context* c, int slot, float* avg, float* mx) {
    hipError_t e = hipMemcpyAsync(&c->host->red[slot], &c->red.p[slot], sizeof(Reduction),
                                  hipMemcpyDeviceToHost, c->stream);
    if (e) { return e; }
    if ((e = fetch_keystate_async(c))) { return e; }
    if ((e = hipStreamSynchronize(c->stream))) { return e; }
    note_keystate(c);
    double const num = static_cast<double>(c->host->red[slot].isum);
    double const den = static_cast<double>(1000ull * c->n);
    *avg = static_cast<float>(num / den);
    unsigned int const bits = c->host->red[slot].max_bits;
    std::memcpy(mx, &bits, 4);
    return hipSuccess;
}

hipError_t read_vmax2(emsph_context* c, float* out) {
    hipError_t e = hipMemcpyAsync(&c->host->vmax2, c->vmax2.p, 4, hipMemcpyDeviceToHost,
                                  c->stream);
    if (e) { return e; }
    if ((e = hipStreamSynchronize(c->stream))) { return e; }
    std::memcpy(out, &c->host->vmax2, 4);
    return hipSuccess;
}

// iisph.cpp:23-151
emsph_status iisph_sub_step(emsph_context* c, float dt, emsph_step_stats* st) {
    emsph_status s = prepare(c, BUILD_IISPH, false, dt, st);
    if (s) { return s; }
    SweepArgs a = sweep_args(c, dt);
    float* pa = c->p0.p;
    float* pb = c->p1.p;
    a.p_out = pa;
    HIP_TRY(timed(c, "iisph_aii", [&] { launch_iisph_aii(a, c->stream); }));
    emsph_parameters const& P = c->cfg.params;
    int const max_it = std::min(P.iisph_max_pressure_iterations, MAX_RED_SLOTS);
    int iter = 0;
    float avg = 0.0f;
    float mx = 0.0f;
    for (; iter < max_it; ++iter) {
        a.p_in = pa;
        HIP_TRY(timed(c, "iisph_sdp", [&] { launch_iisph_sdp(a, c->stream); }));
        a.p_out = pb;
        HIP_TRY(timed(c, "iisph_beta_p", [&] { launch_iisph_beta_p(a, iter, c->stream); }));
        std::swap(pa, pb);
        // iisph.cpp:126: the exit test only runs from the third iteration, so
        // the first two never wait for their residuals.
        if (iter > 1 || iter + 1 == max_it) {
            HIP_TRY(read_residual(c, iter, &avg, &mx));
            if (c->sort_short) { return EMSPH_OK; }
        }
        if (iter > 1 && avg <= P.iisph_error_average_threshold &&
            mx <= P.iisph_error_max_threshold) {
            break;
        }
    }
    HIP_TRY(check_keys(c));
    if (c->sort_short) { return EMSPH_OK; }
    a.p_in = pa;
    HIP_TRY(timed(c, "apply_integrate", [&] { launch_apply(a, 0, 1, c->stream); }));
    c->cur_p = pa;
    if (st) {
        st->iterations = (iter < max_it) ? iter + 1 : iter;
        st->error_average = avg;
        st->error_max = mx;
    }
    return EMSPH_OK;
}

// dfsph_p.cpp:265-374 (with the correction loops of :61-212)
emsph_status dfsph_sub_step(emsph_context* c, float dt, emsph_step_stats* st) {
    emsph_status s = prepare(c, BUILD_DFSPH, true, dt, st);
    if (s) { return s; }
    SweepArgs a = sweep_args(c, dt);
    emsph_parameters const& P = c->cfg.params;
    int const max_it = std::min(P.dfsph_max_correction_iterations, MAX_RED_SLOTS / 2);
    // Divergence solve (dfsph_p.cpp:136-212).
    a.p_in = c->kvs.p;
    a.kappa = c->kvs.p;
    HIP_TRY(timed(c, "apply", [&] { launch_apply(a, 0, 0, c->stream); }));
    int it = 0;
    float avg = 0.0f;
    float mx = 0.0f;
    for (; it < max_it; ++it) {
        HIP_TRY(timed(c, "dfsph_div", [&] { launch_dfsph_div(a, it, c->stream); }));
        HIP_TRY(timed(c, "apply", [&] { launch_apply(a, 0, 0, c->stream); }));
        if (it > 0) {
            HIP_TRY(read_residual(c, it, &avg, &mx));
            if (c->sort_short) { return EMSPH_OK; }
            if (avg <= P.dfsph_divergence_error_average_threshold &&
                mx <= P.dfsph_divergence_error_max_threshold) {
                break;
            }
        }
    }
    if (st) {
        st->divergence_iterations = (it < max_it) ? it + 1 : it;
        st->divergence_error_average = avg;
        st->divergence_error_max = mx;
    }
    // External forces, then the density solve (dfsph_p.cpp:350-366, :61-134).
    a.p_in = c->kds.p;
    a.kappa = c->kds.p;
    HIP_TRY(timed(c, "apply", [&] { launch_apply(a, 1, 0, c->stream); }));
    int const base = MAX_RED_SLOTS / 2;
    avg = 0.0f;
    mx = 0.0f;
    it = 0;
    for (; it < max_it; ++it) {
        HIP_TRY(timed(c, "dfsph_dens", [&] { launch_dfsph_dens(a, base + it, c->stream); }));
        HIP_TRY(timed(c, "apply", [&] { launch_apply(a, 0, 0, c->stream); }));
        if (it > 1) {
            HIP_TRY(read_residual(c, base + it, &avg, &mx));
            if (c->sort_short) { return EMSPH_OK; }
            if (avg <= P.dfsph_density_error_average_threshold &&
                mx <= P.dfsph_density_error_max_threshold) {
                break;
            }
        }
    }
    if (st) {
        st->iterations = (it < max_it) ? it + 1 : it;
        st->error_average = avg;
        st->error_max = mx;
    }
    HIP_TRY(check_keys(c));
    if (c->sort_short) { return EMSPH_OK; }
    HIP_TRY(timed(c, "finalize", [&] { launch_finalize(a, 1, c->stream); }));
    c->cur_p = c->kds.p;
    return EMSPH_OK;
}

// pcisph.cpp:36-164
emsph_status pcisph_sub_step(emsph_context* c, float dt, emsph_step_stats* st) {
    HIP_TRY(ensure_geometry(c));
    bool const keep = c->store_geometry;
    emsph_status s = prepare(c, BUILD_LISTS, false, dt, st);
    if (s) { return s; }
    (void)keep;
    SweepArgs a = sweep_args(c, dt);
    a.p_out = c->p0.p;
    V2* vin = c->vstar0.p;
    V2* vout = c->vstar1.p;
    a.vstar_out = vin;
    HIP_TRY(timed(c, "pci_init", [&] { launch_pci_init(a, c->stream); }));
    float err = 0.0f;
    int it = 0;
    for (; it < 6; ++it) {
        HIP_TRY(timed(c, "pci_density", [&] { launch_pci_density(a, it, c->stream); }));
        a.vstar_in = vin;
        a.vstar_out = vout;
        HIP_TRY(timed(c, "pci_force", [&] { launch_pci_force(a, c->stream); }));
        std::swap(vin, vout);
        if (it > 1) {
            float avg_unused = 0.0f;
            float me = 0.0f;
            HIP_TRY(read_residual(c, it, &avg_unused, &me));
            if (c->sort_short) { return EMSPH_OK; }
            err = me / c->cfg.params.target_density;
            if (err <= 0.01f) { break; }
        }
    }
    HIP_TRY(check_keys(c));
    if (c->sort_short) { return EMSPH_OK; }
    a.vstar_in = vin;
    HIP_TRY(timed(c, "pci_finalize", [&] { launch_pci_finalize(a, c->stream); }));
    c->cur_p = c->p0.p;
    if (st) {
        st->iterations = (it < 6) ? it + 1 : it;
        st->error_average = 0.0f;
        st->error_max = err;
    }
    return EMSPH_OK;
}

emsph_status check_method(int method) {
    if (method == EMSPH_IISPH || method == EMSPH_PCISPH || method == EMSPH_DFSPH_P) {
        return EMSPH_OK;
    }
    return fail(EMSPH_ERR_UNKNOWN_METHOD,
                "method " + std::to_string(method) + " is not implemented on the GPU path");
}

emsph_status sub_step_impl(emsph_context* c, int method, float dt, emsph_step_stats* st) {
    if (st) { std::memset(st, 0, sizeof(*st)); }
    emsph_status s = check_method(method);
    if (s) { return s; }
    if (c->n == 0) { return EMSPH_OK; }  // the reference reads block_indices.back() on N=0
    HIP_TRY(hipSetDevice(c->device));
    // A sort run with the previous sub-step's pass count is checked at the
    // first read-back; when it was short the sub-step is redone with the exact
    // count (the second attempt sorts exactly and cannot be short).
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (attempt == 1) { c->passes_hint = -1; }
        switch (method) {
        case EMSPH_IISPH: s = iisph_sub_step(c, dt, st); break;
        case EMSPH_DFSPH_P: s = dfsph_sub_step(c, dt, st); break;
        default: s = pcisph_sub_step(c, dt, st); break;
        }
        if (s) { return s; }
        if (!c->sort_short) { break; }
    }
    c->last_method = method;
    if (st) {
        HIP_TRY(hipMemcpyAsync(&c->host->cells, c->cells_total.p, 4, hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(read_vmax2(c, &st->max_velocity_sq));
        st->cell_count = c->host->cells;
    }
    return EMSPH_OK;
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
           "denormals kept); sorted-slot neighbour lists, LSD radix sort";
}

emsph_status emsph_create(const emsph_config* config, const float* solid_xy, size_t solid_count,
                          size_t fluid_capacity, emsph_context** out) {
    if (!config || !out || (solid_count && !solid_xy)) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument");
    }
    if (fluid_capacity > (1ull << 31) || solid_count > (1ull << 31)) {
        return fail(EMSPH_ERR_CAPACITY, "capacity above 2^31 particles");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        return fail(EMSPH_ERR_NO_DEVICE, "no HIP device");
    }
    auto* c = new emsph_context();
    c->cfg = *config;
    HIP_TRY(hipGetDevice(&c->device));
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&c->host), sizeof(Pinned)));
    HIP_TRY(hipEventCreate(&c->timer.a));
    HIP_TRY(hipEventCreate(&c->timer.b));
    uint32_t const cap = static_cast<uint32_t>(std::max<size_t>(fluid_capacity, 1));
    c->cap = cap;
    c->stride = (cap + 255u) & ~255u;
    size_t const L = static_cast<size_t>(MAXN) * c->stride;
    hipError_t e;
    if ((e = c->pos.alloc(cap)) || (e = c->vel.alloc(cap)) || (e = c->kv.alloc(cap)) ||
        (e = c->kd.alloc(cap)) || (e = c->fidx.alloc(cap, false)) || (e = c->spos.alloc(cap)) ||
        (e = c->svel.alloc(cap)) || (e = c->dii.alloc(cap)) || (e = c->sdp.alloc(cap)) ||
        (e = c->vext.alloc(cap)) || (e = c->xstar.alloc(cap)) || (e = c->vstar0.alloc(cap)) ||
        (e = c->vstar1.alloc(cap)) || (e = c->force.alloc(cap)) || (e = c->rho.alloc(cap)) ||
        (e = c->rho_star.alloc(cap)) || (e = c->aii.alloc(cap)) || (e = c->p0.alloc(cap)) ||
        (e = c->p1.alloc(cap)) || (e = c->alpha.alloc(cap)) || (e = c->kvs.alloc(cap)) ||
        (e = c->kds.alloc(cap)) || (e = c->fcount.alloc(c->stride)) ||
        (e = c->scount.alloc(c->stride)) || (e = c->fl_idx.alloc(L)) ||
        (e = c->sl_idx.alloc(L)) || (e = c->fl_gw.alloc(L)) || (e = c->sl_gw.alloc(L)) ||
        (e = c->red.alloc(MAX_RED_SLOTS)) || (e = c->keystate.alloc(1)) ||
        (e = c->vmax2.alloc(1)) || (e = c->cells_total.alloc(1))) {
        delete c;
        return fail(e == hipErrorOutOfMemory ? EMSPH_ERR_OUT_OF_MEMORY : EMSPH_ERR_HIP,
                    std::string("device allocation: ") + hipGetErrorString(e));
    }
    HIP_TRY(hipMemsetAsync(c->kv.p, 0, cap * sizeof(float), c->stream));
    HIP_TRY(hipMemsetAsync(c->kd.p, 0, cap * sizeof(float), c->stream));
    HIP_TRY(hipMemsetAsync(c->red.p, 0, MAX_RED_SLOTS * sizeof(Reduction), c->stream));

    // Solids: index, solid-solid lists with geometry, volumes (solids.cpp:76-137).
    c->ns = static_cast<uint32_t>(solid_count);
    if (c->ns) {
        uint32_t const ns = c->ns;
        uint32_t const sstride = (ns + 255u) & ~255u;
        size_t const SL = static_cast<size_t>(MAXN) * sstride;
        DBuf<uint8_t> scnt;
        DBuf<uint32_t> sli;
        DBuf<float> sld;
        DBuf<V2> slv, slg;
        DBuf<float> vol_unused;
        if ((e = c->sidx.alloc(ns, false)) || (e = c->s_pos.alloc(ns)) ||
            (e = c->s_vel.alloc(ns)) || (e = c->s_pos_orig.alloc(ns)) ||
            (e = c->s_vol.alloc(ns)) || (e = scnt.alloc(sstride)) || (e = sli.alloc(SL)) ||
            (e = sld.alloc(SL)) || (e = slv.alloc(SL)) || (e = slg.alloc(SL))) {
            delete c;
            return fail(EMSPH_ERR_OUT_OF_MEMORY, "solid allocation");
        }
        HIP_TRY(hipMemcpyAsync(c->s_pos_orig.p, solid_xy, ns * sizeof(V2), hipMemcpyHostToDevice,
                               c->stream));
        HIP_TRY(hipMemsetAsync(c->s_vel.p, 0, ns * sizeof(V2), c->stream));
        CellsArgs g{};
        g.pos_in = c->s_pos_orig.p;
        g.pos_out = c->s_pos.p;
        HIP_TRY(index_points(c->sidx, c->s_pos_orig.p, ns, config->params.support * 2.0f,
                             c->stream, c->host, c->keystate.p, -1, c->cells_total.p, nullptr, nullptr,
                             g, nullptr));
        BuildArgs b{};
        b.c = step_consts(c, 0.0f);
        b.n = ns;
        b.stride = sstride;
        b.pos = c->s_pos.p;
        b.orig = c->sidx.orig.p;
        b.rank = c->sidx.rank.p;
        b.table = c->sidx.table();
        b.solids = SolidView{};  // no solid-vs-solid second list
        b.fl = Lists{scnt.p, sli.p, slg.p, sld.p, slv.p};
        b.sl = Lists{scnt.p, sli.p, slg.p, sld.p, slv.p};
        b.store_geometry = 1;
        launch_build(BUILD_LISTS, b, c->stream);
        launch_solid_volumes(b.c.kc, ns, sstride, b.fl, c->s_vol.p, c->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    *out = c;
    return EMSPH_OK;
}

emsph_status emsph_destroy(emsph_context* ctx) {
    if (!ctx) { return EMSPH_OK; }
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return EMSPH_OK;
}

emsph_status emsph_set_store_geometry(emsph_context* c, int enable) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    c->store_geometry = enable != 0;
    if (c->store_geometry) { HIP_TRY(ensure_geometry(c)); }
    return EMSPH_OK;
}

emsph_status emsph_set_state(emsph_context* c, const float* pos_xy, const float* vel_xy,
                             size_t count) {
    if (!c || (count && !pos_xy)) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    if (count > c->cap) { return fail(EMSPH_ERR_CAPACITY, "fluid count above capacity"); }
    c->n = static_cast<uint32_t>(count);
    c->have_lists = false;
    c->passes_hint = -1;
    if (count) {
        HIP_TRY(hipMemcpyAsync(c->pos.p, pos_xy, count * sizeof(V2), hipMemcpyHostToDevice,
                               c->stream));
        if (vel_xy) {
            HIP_TRY(hipMemcpyAsync(c->vel.p, vel_xy, count * sizeof(V2), hipMemcpyHostToDevice,
                                   c->stream));
        } else {
            HIP_TRY(hipMemsetAsync(c->vel.p, 0, count * sizeof(V2), c->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return EMSPH_OK;
}

emsph_status emsph_get_state(emsph_context* c, float* pos_xy, float* vel_xy) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    if (c->n) {
        if (pos_xy) {
            HIP_TRY(hipMemcpyAsync(pos_xy, c->pos.p, c->n * sizeof(V2), hipMemcpyDeviceToHost,
                                   c->stream));
        }
        if (vel_xy) {
            HIP_TRY(hipMemcpyAsync(vel_xy, c->vel.p, c->n * sizeof(V2), hipMemcpyDeviceToHost,
                                   c->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return EMSPH_OK;
}

size_t emsph_fluid_count(const emsph_context* c) { return c ? c->n : 0; }
size_t emsph_solid_count(const emsph_context* c) { return c ? c->ns : 0; }

emsph_status emsph_get_solid_volumes(emsph_context* c, float* out) {
    if (!c || !out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    if (!c->ns) { return EMSPH_OK; }
    DBuf<float> tmp;
    HIP_TRY(tmp.alloc(c->ns));
    launch_scatter_f32(c->s_vol.p, c->sidx.orig.p, tmp.p, c->ns, c->stream);
    HIP_TRY(hipMemcpyAsync(out, tmp.p, c->ns * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return EMSPH_OK;
}

emsph_status emsph_get_field(emsph_context* c, int field, float* out) {
    if (!c || !out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    if (!c->have_lists || !c->n) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no sub-step yet"); }
    float const* f1 = nullptr;
    V2 const* f2 = nullptr;
    switch (field) {
    case EMSPH_FIELD_DENSITY:
        // PCISPH keeps its predicted densities in the rho_star buffer.
        f1 = (c->last_method == EMSPH_PCISPH) ? c->rho_star.p : c->rho.p;
        break;
    case EMSPH_FIELD_DENSITY_STAR: f1 = c->rho_star.p; break;
    case EMSPH_FIELD_PRESSURE: f1 = c->cur_p ? c->cur_p : c->p0.p; break;
    case EMSPH_FIELD_AII: f1 = c->aii.p; break;
    case EMSPH_FIELD_DII: f2 = c->dii.p; break;
    case EMSPH_FIELD_SUM_DIJ_PJ: f2 = c->sdp.p; break;
    case EMSPH_FIELD_ALPHA: f1 = c->alpha.p; break;
    case EMSPH_FIELD_KAPPA_DENSITY: f1 = c->kds.p; break;
    case EMSPH_FIELD_KAPPA_DIVERGENCE: f1 = c->kvs.p; break;
    case EMSPH_FIELD_PRESSURE_FORCE: f2 = c->force.p; break;
    default: return fail(EMSPH_ERR_INVALID_ARGUMENT, "unknown field");
    }
    if (f1) {
        DBuf<float> tmp;
        HIP_TRY(tmp.alloc(c->n));
        launch_scatter_f32(f1, c->fidx.orig.p, tmp.p, c->n, c->stream);
        HIP_TRY(hipMemcpyAsync(out, tmp.p, c->n * sizeof(float), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    } else {
        DBuf<V2> tmp;
        HIP_TRY(tmp.alloc(c->n));
        launch_scatter_v2(f2, c->fidx.orig.p, tmp.p, c->n, c->stream);
        HIP_TRY(hipMemcpyAsync(out, tmp.p, c->n * sizeof(V2), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return EMSPH_OK;
}

emsph_status emsph_get_neighborhood(emsph_context* c, int which, uint8_t* counts,
                                    uint32_t* indices, float* distances, float* vectors_xy) {
    if (!c || !counts || !indices) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    if (!c->have_lists || !c->n) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "no sub-step yet"); }
    bool const geo = (distances || vectors_xy);
    if (geo && !c->fl_dist.p) {
        return fail(EMSPH_ERR_INVALID_ARGUMENT, "geometry not stored (emsph_set_store_geometry)");
    }
    Lists l = which ? solid_lists(c) : fluid_lists(c);
    uint32_t const* other = which ? c->sidx.orig.p : c->fidx.orig.p;
    size_t const m = static_cast<size_t>(c->n) * MAXN;
    DBuf<uint8_t> dc;
    DBuf<uint32_t> di;
    DBuf<float> dd;
    DBuf<V2> dv;
    HIP_TRY(dc.alloc(c->n));
    HIP_TRY(di.alloc(m));
    if (distances) { HIP_TRY(dd.alloc(m)); }
    if (vectors_xy) { HIP_TRY(dv.alloc(m)); }
    launch_export_lists(l, c->n, c->stride, c->fidx.orig.p, other, dc.p, di.p,
                        distances ? dd.p : nullptr, vectors_xy ? dv.p : nullptr, c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(counts, dc.p, c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(indices, di.p, m * 4, hipMemcpyDeviceToHost, c->stream));
    if (distances) {
        HIP_TRY(hipMemcpyAsync(distances, dd.p, m * 4, hipMemcpyDeviceToHost, c->stream));
    }
    if (vectors_xy) {
        HIP_TRY(hipMemcpyAsync(vectors_xy, dv.p, m * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return EMSPH_OK;
}

emsph_status emsph_frame_init(emsph_context* c, int method) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    emsph_status s = check_method(method);
    if (s) { return s; }
    if (method == EMSPH_DFSPH_P && c->n) {
        HIP_TRY(hipMemsetAsync(c->kv.p, 0, c->n * sizeof(float), c->stream));
        HIP_TRY(hipMemsetAsync(c->kd.p, 0, c->n * sizeof(float), c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return EMSPH_OK;
}

emsph_status emsph_sub_step(emsph_context* c, int method, float time_seconds, float dt,
                            emsph_step_stats* stats) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    (void)time_seconds;  // the built-in gravity force model does not depend on time
    return sub_step_impl(c, method, dt, stats);
}

emsph_status emsph_frame_step(emsph_context* c, int method, int64_t global_time_flicks,
                              emsph_frame_stats* fs) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    emsph_status s = emsph_frame_init(c, method);
    if (s) { return s; }
    emsph_frame_stats local{};
    emsph_step_stats st{};
    emsph_parameters const& P = c->cfg.params;
    if (method == EMSPH_PCISPH) {
        // pcisph.cpp:166-204: fixed sub-steps of the configured dt.
        int64_t const per = P.time_per_step / P.sub_steps;
        int64_t time = global_time_flicks;
        for (int k = 0; k < P.sub_steps; ++k) {
            s = emsph_sub_step(c, method, static_cast<float>(to_seconds(time)),
                               c->cfg.seconds_per_sub_step, &st);
            if (s) { return s; }
            time += per;
            local.sub_steps += 1;
            local.total_iterations += st.iterations;
            local.last_dt = c->cfg.seconds_per_sub_step;
        }
    } else {
        // adaptive_time_step.h:35-122 with (60, 6, 0.2) for IISPH
        // (iisph.cpp:175-181) and (60, 4, 0.4) for DFSPH (dfsph_p.cpp:383-389).
        int const min_den = 60;
        int const max_den = (method == EMSPH_IISPH) ? 6 : 4;
        float const cfl = (method == EMSPH_IISPH) ? 0.2f : 0.4f;
        int64_t const per = P.time_per_step / min_den;
        int const divisor = min_den / max_den;
        int remaining = min_den;
        int64_t time = global_time_flicks;
        HIP_TRY(hipMemsetAsync(c->vmax2.p, 0, 4, c->stream));
        launch_vmax2(c->vel.p, c->n, c->vmax2.p, c->stream);
        float vmax2 = 0.0f;
        HIP_TRY(read_vmax2(c, &vmax2));
        while (remaining > 0) {
            // cfl.cpp:12-21
            double const numer = cfl * P.support;
            double const denom = std::sqrt(static_cast<double>(vmax2));
            bool safe = false;
            if (denom != 0.0) {
                safe = std::fabs(denom) >= 1.0 ||
                       std::fabs(denom) / std::numeric_limits<double>::min() > std::fabs(numer);
            }
            int64_t const cfl_step = to_flicks(safe ? numer / denom : 1.0);
            int ns = static_cast<int>(std::ceil(to_seconds(cfl_step) / to_seconds(per)));
            ns = std::clamp(ns, 1, divisor);
            ns = std::min(ns, remaining);
            int64_t const sub = ns * per;
            float const dt = static_cast<float>(to_seconds(sub));
            s = emsph_sub_step(c, method, static_cast<float>(to_seconds(time)), dt, &st);
            if (s) { return s; }
            vmax2 = st.max_velocity_sq;
            remaining -= ns;
            local.sub_steps += 1;
            local.total_iterations += st.iterations;
            local.total_divergence_iterations += st.divergence_iterations;
            local.last_dt = dt;
            time += sub;
        }
    }
    local.last = st;
    if (fs) { *fs = local; }
    return EMSPH_OK;
}

emsph_status emsph_run_sub_steps(emsph_context* c, int method, float dt, int n_sub_steps,
                                 float* elapsed_ms, emsph_step_stats* last) {
    if (!c || n_sub_steps < 0) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument"); }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, c->stream));
    emsph_step_stats st{};
    for (int k = 0; k < n_sub_steps; ++k) {
        emsph_status const s = sub_step_impl(c, method, dt, (k + 1 == n_sub_steps) ? &st : nullptr);
        if (s) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return s;
        }
    }
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) { *elapsed_ms = ms; }
    if (last) { *last = st; }
    return EMSPH_OK;
}

emsph_status emsph_set_kernel_timing(emsph_context* c, int enable) {
    if (!c) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null context"); }
    c->timer.enabled = enable != 0;
    c->timer.acc.clear();
    return EMSPH_OK;
}

int emsph_get_kernel_timing(emsph_context* c, const char** names, double* total_ms,
                            int64_t* calls, int max) {
    if (!c) { return 0; }
    int k = 0;
    for (auto const& kv : c->timer.acc) {
        if (k >= max) { break; }
        if (names) { names[k] = kv.first.c_str(); }
        if (total_ms) { total_ms[k] = kv.second.first; }
        if (calls) { calls[k] = kv.second.second; }
        ++k;
    }
    return k;
}

//------------------------------------------------------------------------------
emsph_status emsph_search_create(size_t capacity, emsph_search** out) {
    if (!out) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    if (capacity > (1ull << 31)) { return fail(EMSPH_ERR_CAPACITY, "capacity above 2^31"); }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        return fail(EMSPH_ERR_NO_DEVICE, "no HIP device");
    }
    auto* s = new emsph_search();
    uint32_t const cap = static_cast<uint32_t>(std::max<size_t>(capacity, 1));
    s->cap = cap;
    s->stride = (cap + 255u) & ~255u;
    size_t const L = static_cast<size_t>(MAXN) * s->stride;
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&s->host), sizeof(Pinned)));
    hipError_t e;
    if ((e = s->pos.alloc(cap)) || (e = s->spos.alloc(cap)) || (e = s->idx.alloc(cap, true)) ||
        (e = s->count.alloc(s->stride)) || (e = s->lidx.alloc(L)) ||
        (e = s->keystate.alloc(1)) || (e = s->cells_total.alloc(1))) {
        delete s;
        return fail(EMSPH_ERR_OUT_OF_MEMORY, std::string("search allocation: ") +
                                               hipGetErrorString(e));
    }
    *out = s;
    return EMSPH_OK;
}

emsph_status emsph_search_destroy(emsph_search* s) {
    if (s) {
        (void)hipStreamSynchronize(s->stream);
        delete s;
    }
    return EMSPH_OK;
}

emsph_status emsph_search_set_points(emsph_search* s, const float* xy, size_t count) {
    if (!s || (count && !xy)) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "null argument"); }
    if (count > s->cap) { return fail(EMSPH_ERR_CAPACITY, "point count above capacity"); }
    s->n = static_cast<uint32_t>(count);
    s->ran = false;
    if (count) {
        HIP_TRY(hipMemcpyAsync(s->pos.p, xy, count * sizeof(V2), hipMemcpyHostToDevice,
                               s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return EMSPH_OK;
}

emsph_status emsph_search_run(emsph_search* s, float cell_size, int store_geometry,
                              float* elapsed_ms) {
    if (!s || !(cell_size > 0.0f)) { return fail(EMSPH_ERR_INVALID_ARGUMENT, "bad argument"); }
    if (store_geometry && !s->dist.p) {
        size_t const L = static_cast<size_t>(MAXN) * s->stride;
        HIP_TRY(s->dist.alloc(L));
        HIP_TRY(s->vec.alloc(L));
    }
    s->geometry = store_geometry != 0;
    s->cell = cell_size;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->stream));
    if (s->n) {
        CellsArgs g{};
        g.pos_in = s->pos.p;
        g.pos_out = s->spos.p;
        HIP_TRY(index_points(s->idx, s->pos.p, s->n, cell_size, s->stream, s->host, s->keystate.p, -1,
                             s->cells_total.p, nullptr, nullptr, g, nullptr));
        BuildArgs b{};
        b.c.cell = cell_size;
        b.c.kc = make_kernel_consts(cell_size * 0.5f);
        b.n = s->n;
        b.stride = s->stride;
        b.pos = s->spos.p;
        b.orig = s->idx.orig.p;
        b.rank = s->idx.rank.p;
        b.table = s->idx.table();
        b.solids = SolidView{};
        b.fl = Lists{s->count.p, s->lidx.p, nullptr, s->dist.p, s->vec.p};
        b.sl = b.fl;
        b.store_geometry = s->geometry ? 1 : 0;
        launch_build(BUILD_LISTS, b, s->stream);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(e1, s->stream));
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
    uint32_t const n = s->n;
    if (n == 0) {
        if (block_count) { *block_count = 0; }
        return EMSPH_OK;
    }
    hipStream_t const st = s->stream;
    HIP_TRY(hipMemcpyAsync(&s->host->cells, s->cells_total.p, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    uint32_t const nb = s->host->cells;
    if (block_count) { *block_count = nb; }
    if (grid_xy) {
        HIP_TRY(hipMemcpyAsync(grid_xy, s->idx.grid.p, n * 8, hipMemcpyDeviceToHost, st));
    }
    if (keys) {
        // keys0 holds the unsorted keys only when the sort ran an even number
        // of passes into it; recompute them in original order instead.
        KeysArgs ka{};
        ka.pos = s->pos.p;
        ka.n = n;
        ka.cell = s->cell;
        DBuf<uint64_t> k2;
        DBuf<uint32_t> v2;
        HIP_TRY(k2.alloc(n));
        HIP_TRY(v2.alloc(n));
        ka.keys = k2.p;
        ka.vals = v2.p;
        launch_keys(ka, st);
        HIP_TRY(hipMemcpyAsync(keys, k2.p, n * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (sorted_keys) {
        HIP_TRY(hipMemcpyAsync(sorted_keys, s->idx.sorted_keys, n * 8, hipMemcpyDeviceToHost, st));
    }
    if (sorted_idx) {
        HIP_TRY(hipMemcpyAsync(sorted_idx, s->idx.orig.p, n * 4, hipMemcpyDeviceToHost, st));
    }
    if (block_indices) {
        HIP_TRY(hipMemcpyAsync(block_indices, s->idx.block_index.p, n * 4, hipMemcpyDeviceToHost,
                               st));
    }
    if (block_begin) {
        HIP_TRY(hipMemcpyAsync(block_begin, s->idx.cell_begin.p, nb * 4, hipMemcpyDeviceToHost,
                               st));
    }
    if (block_end) {
        HIP_TRY(hipMemcpyAsync(block_end, s->idx.cell_end.p, nb * 4, hipMemcpyDeviceToHost, st));
    }
    if (counts || indices || distances || vectors_xy) {
        if ((distances || vectors_xy) && !s->geometry) {
            return fail(EMSPH_ERR_INVALID_ARGUMENT, "geometry not stored by emsph_search_run");
        }
        size_t const m = static_cast<size_t>(n) * MAXN;
        DBuf<uint8_t> dc;
        DBuf<uint32_t> di;
        DBuf<float> dd;
        DBuf<V2> dv;
        HIP_TRY(dc.alloc(n));
        HIP_TRY(di.alloc(m));
        if (distances) { HIP_TRY(dd.alloc(m)); }
        if (vectors_xy) { HIP_TRY(dv.alloc(m)); }
        Lists l{s->count.p, s->lidx.p, nullptr, s->dist.p, s->vec.p};
        launch_export_lists(l, n, s->stride, s->idx.orig.p, s->idx.orig.p, dc.p, di.p,
                            distances ? dd.p : nullptr, vectors_xy ? dv.p : nullptr, st);
        HIP_TRY(hipGetLastError());
        if (counts) { HIP_TRY(hipMemcpyAsync(counts, dc.p, n, hipMemcpyDeviceToHost, st)); }
        if (indices) { HIP_TRY(hipMemcpyAsync(indices, di.p, m * 4, hipMemcpyDeviceToHost, st)); }
        if (distances) {
            HIP_TRY(hipMemcpyAsync(distances, dd.p, m * 4, hipMemcpyDeviceToHost, st));
        }
        if (vectors_xy) {
            HIP_TRY(hipMemcpyAsync(vectors_xy, dv.p, m * 8, hipMemcpyDeviceToHost, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return EMSPH_OK;
}

}  // extern "C"
