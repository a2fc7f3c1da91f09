// Kernel argument blocks and host launchers for the 2D SPH hot path.
//
// Every kernel works on a batch of M >= 1 independent scenes packed one after
// another ("slots" [offset_s, offset_s + n_s) of scene s). A single scene is
// the batch M = 1. Fluid kernels walk scene-aligned tiles of up to 256 slots,
// so a workgroup never mixes scenes and can skip a converged scene as a whole
// and reduce its residual into that scene's slot with one atomic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_math.hpp"

namespace emsph {

constexpr int MAXN = 64;           // NEIGHBORHOOD_MAX_COUNT (sph_common/neighborhood.h:38)
constexpr int MAX_RED_SLOTS = 64;  // residual slots per scene and sub-step
constexpr int TILE = 256;          // fluid slots per tile (one workgroup)

// Per-scene constants of one sub-step.
struct StepConsts {
    KernelConsts kc;
    float cell;          // 2h: grid cell and neighbour cut-off
    // Smallest float x with sqrtf(x) >= cell: since the correctly rounded
    // square root is monotone, |v| < cell exactly when |v|^2 < cut2 (also in
    // Imath's tiny-length branch, |v|^2 < 2 FLT_MIN, when cell >= 1e-18);
    // 0 when cell is below that, and lengths are compared instead.
    float cut2;
    float rho0;          // target density
    float fluid_volume;  // m / rho0 (sph_common/volume.cpp:7-14)
    float mass;          // mass per particle
    float gravity;
    float dt;
    float omega;         // IISPH relaxation
    float pci_denom;     // PCISPH pressure correction denominator
    float viscosity;
    // Skin candidate lists (SkinView): a candidate is recorded at a rebuild
    // when |v|^2 < skin_r2 (at least (cell + h)^2), and the lists serve while
    // every particle's |x - x_rebuild|^2 < skin_lim2 (at most (h/2)^2 less a
    // margin for the rounding of both tests; both set on the host).
    float skin_r2;
    float skin_lim2;
};

// Solver-loop exit thresholds of one scene (sph2d_box/parameters.h:22-33).
struct Thresholds {
    float iisph_avg, iisph_max;
    float div_avg, div_max;
    float dens_avg, dens_max;
    int32_t iisph_max_iters;
    int32_t dfsph_max_iters;
};

// External force models of one scene (sph2d_box/forces.cpp:62-149), applied in
// flag order: FORCE_GRAVITY (forces.y -= m g, with StepConsts::gravity),
// FORCE_POLE, FORCE_DRAG, FORCE_ANTI_COUPLING. FORCE_GRAVITY alone with no
// host forces is the default, fused into the kernel that integrates the force.
constexpr uint32_t FORCE_GRAVITY = 1u;
constexpr uint32_t FORCE_POLE = 2u;
constexpr uint32_t FORCE_DRAG = 4u;
constexpr uint32_t FORCE_ANTI_COUPLING = 8u;
struct ForceDev {
    uint32_t flags;
    float pole_mag;
    float pole_x, pole_y;
    float drag_mult;    // -drag magnitude * drag diameter (forces.cpp:122)
    float anti_maxd;    // anti-coupling distance limit and magnitude
    float anti_mag;
    uint32_t pad;
};

// Frame loop of the adaptive time step (adaptive_time_step.h:35-122) kept on
// the device: the start of each sub-step picks its dt from the CFL condition
// in 1/min_den quanta of the frame; the end of the sub-step commits the
// remaining quanta and the time (so a sub-step redone after a short sort picks
// the same dt).
struct FrameDev {
    int64_t per;        // flicks per quantum (time_per_step / min_den)
    int64_t time;       // start of the current sub-step, flicks
    int64_t next_time;  // its end (committed at the end of the sub-step)
    int32_t remaining;  // quanta left before the current sub-step
    int32_t next_remaining;
    int32_t divisor;    // min_den / max_den: the most quanta one sub-step takes
    float cfl;          // the method's CFL factor
    int32_t on;
    int32_t pad;
};

struct SceneDev {
    StepConsts c;
    Thresholds thr;
    uint32_t offset;      // first fluid slot
    uint32_t n;           // fluid particles
    uint32_t soffset;     // first solid slot
    uint32_t ns;          // solid particles
    uint32_t first_tile;  // the scene's fluid tiles are [first_tile, first_tile + ntiles)
    uint32_t ntiles;
    float cfl;            // > 0: c.dt = min(dt_max, cfl * h / max |v|) at each sub-step
    float dt_max;
    ForceDev force;
    FrameDev frame;
    uint32_t retired;     // a pool scene removed but not yet committed: no solver iterations
    uint32_t pad_[3];
};

struct Tile {
    uint32_t scene;
    uint32_t start;
    uint32_t count;
    uint32_t pad;
};

// sum_j d_ij p_j and d_ii of one particle, gathered together by its
// neighbours in the IISPH beta sweep.
struct __align__(16) SdDii {
    V2 sdp;
    V2 dii;
};

// One residual reduction: quantised integer sum and float max (as bits; every
// reduced value is >= 0 so its bit pattern orders like the float).
struct Reduction {
    unsigned long long isum;
    unsigned int max_bits;
    unsigned int pad;
};

// Per-scene solver statistics of the last sub-step.
struct SceneStats {
    int32_t iterations;
    int32_t divergence_iterations;
    float error_average;
    float error_max;
    float divergence_error_average;
    float divergence_error_max;
    unsigned int vmax2_bits;
    unsigned int active;   // 1 while the current solver loop still iterates
    float dt;              // dt of the sub-step
    unsigned int pad;
    double kinetic_energy;    // sum of m |v|^2 / 2 after the sub-step (FP64)
    double potential_energy;  // sum of m g y after the sub-step (FP64)
};

// Workgroups of the key kernel (grid-stride), i.e. partial key ranges.
constexpr uint32_t KEY_PARTS = 2048;

// Key range of one sort and the flag raised when the digits that were sorted
// do not cover it (the work is then redone with the exact pass count), plus
// the solver-loop and neighbour-build counters the host reads back.
struct KeyState {
    unsigned long long kmin;
    unsigned long long kmax;
    unsigned int sort_short;
    unsigned int active_count;  // scenes still iterating (written by k_loop_*)
    unsigned int overflow_count;  // particles the last build deferred to its full-cap pass
    unsigned int max_count;       // largest fluid or solid list of the last build
    unsigned int cells;           // cells of the last indexing run
    unsigned int long_lists;      // particles of the last build with a list above 16
    unsigned int skin_flags;      // SKIN_* of the last candidate build or reuse
    unsigned int skin_disp;       // float bits: max |x - x_rebuild|^2 / skin_lim2 of the last reuse
    int frame_remaining;          // scene 0's frame quanta left after this sub-step (-1: off)
    unsigned int dense_crowded;   // a cell of the dense index exceeded DENSE_CELL_MAX
    unsigned int dense_out;       // a key fell outside its scene's dense table (k_keys)
    // Scenes still iterating in the solver loop, one bit per scene for the
    // first ACTIVE_BIT_SCENES scenes (set at the loop's start, cleared by the
    // exit test): the host's read-back sizes later sweeps to them.
    unsigned int active_bits[128];
};
constexpr uint32_t ACTIVE_BIT_SCENES = 128 * 32;

// The tiles a sweep launch covers when few scenes are still iterating: the
// launch's workgroup k (after the XCD remap) takes tile first[j] + k - start[j]
// for the j with start[j] <= k < start[j + 1] (start[n] = total, the grid).
// n = 0: every tile, in order. Passed by value (kernel arguments); built on
// the host from an earlier iteration's active bits, a superset of the scenes
// still iterating (a scene never re-enters the loop), so the kernels keep
// their per-scene active test.
constexpr int ACTIVE_TILES_MAX = 64;
struct ActiveTiles {
    uint32_t n;
    uint32_t first[ACTIVE_TILES_MAX];
    uint32_t start[ACTIVE_TILES_MAX + 1];
};

// Dense cell index: cells counted in a table of bins, scene s owning the
// 4^B_s bins [base_s, base_s + 4^B_s) indexed by Morton key (one extra bin,
// the last, collects keys outside their scene's table); an exclusive scan
// gives each cell's first sorted slot, and the particles are scattered there
// and put in increasing original index within their cell, which is the order
// of the stable sort by (scene, key). A key outside its table or a cell above
// DENSE_CELL_MAX particles flags the indexing run as short (it is redone by
// the sort); every slot still receives a valid particle index.
constexpr int DENSE_MAX_BITS = 12;
constexpr uint32_t DENSE_CELL_MAX = 256;

struct DenseScene {
    uint64_t limit;  // 4^B_s (0: the scene has no table)
    uint32_t base;   // first bin of the scene's table
    uint32_t pad;
};

// Bin of a key of scene d's table; bins (the table's total, i.e. the index of
// the extra bin) for a key outside it.
__host__ __device__ __forceinline__ uint64_t dense_bin(DenseScene const& d, uint64_t key,
                                                       uint64_t bins) {
    return key < d.limit ? d.base + key : bins;
}

// Neighbour-list capacities of the build kernel's LDS list (the full one is
// the reference's NEIGHBORHOOD_MAX_COUNT).
constexpr int BUILD_CAPS[3] = {16, 32, 64};

// Cells of one point set in sorted (scene, key) order: cell c is the slot
// range [begin[c], end[c]) of key key[c] in scene tag[c]; *count cells.
struct CellList {
    uint64_t const* key;
    uint32_t const* tag;
    uint32_t const* begin;
    uint32_t const* end;
    uint32_t const* count;
};

// Static set of cells (the solid cells): open addressing over 16-byte slots
// {key low, key high, scene, cell ordinal}, ordinal EMPTY_SLOT when free.
// Built once on the host with a load factor of at most 1/4, so a lookup of a
// cell that is not in the set mostly ends at its first slot.
constexpr uint32_t EMPTY_SLOT = 0xFFFFFFFFu;
// Top bit of a fluid window range's end: the same neighbour cell holds solids
// (slot counts stay below 2^31).
constexpr uint32_t WIN_SOLID_NEAR = 0x80000000u;
struct CellSet {
    uint4 const* slots;
    uint32_t mask;
};

// Static solid particles of all scenes, in their own sorted order.
struct SolidView {
    V2 const* pos;
    V2 const* vel;
    float const* vol;
    uint32_t const* rank;   // original solid index -> sorted slot
    uint32_t const* orig;   // sorted slot -> original solid index
};

// Skin candidate lists (Verlet lists made exact): at a rebuild sub-step each
// fluid particle records, from the 5x5 cells around its own, the fluid and
// solid particles closer than its scene's cut-off plus h (StepConsts::skin_r2).
// A later sub-step whose particles have all moved less than h/2 since
// (skin_lim2) finds every particle the reference would accept among those
// candidates: one that was farther than cut-off + h cannot have come closer
// than the cut-off. The build then applies the reference's own test to each
// candidate — the candidate's cell within one of the particle's cell in both
// axes (the 3x3 window, from the cells of this sub-step's positions) and
// |v|^2 < cut2 — so the lists are the reference's; their order comes from the
// sort by (distance, index) as before (the window test is made only for pairs
// whose separation is within the rounding of the cell assignment of a whole
// cell in some axis; all others are in neighbouring cells). Lists that would
// reach 64 entries
// (truncated in scan order by the reference) flag the sub-step for a redo by
// the full path, as do particles that moved too far.
constexpr int SKIN_FLUID = 48;  // fluid candidates per particle
constexpr int SKIN_SOLID = 32;  // solid candidates per particle
constexpr unsigned int SKIN_OVERFLOW = 1u;  // a candidate list did not fit: no reuse
constexpr unsigned int SKIN_STALE = 2u;     // a particle moved too far: redo by the full path
constexpr unsigned int SKIN_FULL = 4u;      // a reused list reached 64 entries: redo likewise
constexpr int SKIN_MIN_RUN = 6;    // reuses a candidate build must serve to pay for itself
constexpr int SKIN_PAUSE = 64;     // sub-steps on the full path alone after a shorter run
struct SkinView {
    int reuse;              // 1: this sub-step's lists come from the candidates
    uint32_t* cf;           // [SKIN_FLUID][stride] fluid candidates (sorted slots)
    uint8_t* nf;            // per slot: fluid candidate count
    uint32_t* cs;           // [SKIN_SOLID][stride] solid candidates (solid sorted slots)
    uint8_t* ns;            // per slot: solid candidate count
    V2* xref;               // per slot: position at the rebuild
    unsigned int* tmax;     // per tile: largest displacement ratio (float bits) of the gather
};

// Neighbour lists in slot-major layout: entry k of slot t at [k * stride + t].
struct Lists {
    uint8_t* count;
    uint32_t* idx;   // neighbour sorted slot
    V2* gw;          // GradW(-vector_to) of the pair
    float* dist;     // optional (PCISPH, parity export)
    V2* vec;         // optional (PCISPH, parity export)
};

// Everything a fluid kernel may touch.
struct Batch {
    SceneDev const* scenes;
    Tile const* tiles;
    uint32_t ntiles;
    uint32_t nscenes;
    uint32_t stride;
    SceneStats* stats;
    Reduction* red;         // [scene][MAX_RED_SLOTS]
    KeyState* keystate;
    // sorted-space fluid fields
    V2* pos;
    V2* vel;
    uint32_t const* orig;   // sorted slot -> original index
    uint32_t const* rank;   // original index -> sorted slot
    SolidView solids;
    // Cell windows: for the cell of each fluid slot, the [begin, end) slot
    // ranges of its 3x3 neighbour cells (scan order j-outer, i-inner) in the
    // fluid (win) and solid (swin) index; cell_of maps slot -> cell.
    uint32_t const* cell_of;
    uint2 const* win;
    uint2 const* swin;
    uint2* overflow;        // {slot, scene} of particles deferred to the full-cap pass
    double2* tile_energy;   // per fluid tile: FP64 kinetic, potential energy
    Reduction* tile_red;    // per fluid tile: residual partial of the last sweep
    unsigned int* tile_vmax;  // per fluid tile: max |v|^2 bits of the final kernel
    Lists fl;
    Lists sl;
    float* rho;
    float* rho_star;
    float* aii;
    V2* dii;
    float const* p_in;
    float* p_out;
    // Second buffer of a solver loop's ping-pong: a scene that ran an odd number
    // of iterations left its result there (converged scenes stop writing, so
    // scenes of one batch end in different buffers). Read by the final kernels.
    float* p_alt;
    // IISPH gathers: {p, rho} of every particle (ping-ponged like the
    // pressures; pr_alt as p_alt) and {sum_j d_ij p_j, d_ii}, so a neighbour's
    // fields are two loads.
    V2 const* pr_in;
    V2* pr_out;
    V2 const* pr_alt;
    SdDii* sd;
    // IISPH_AP / IISPH_PSEUDO_AP pressure accelerations (displacements)
    V2* pacc;
    float* alpha;
    float* kappa;
    // original-order state (gathered in by the cells stage, scattered back at the end)
    V2* pos_orig;
    V2* vel_orig;
    float* kv_orig;
    float* kd_orig;
    float* kvs;
    float* kds;
    // PCISPH
    V2* vext;
    V2* xstar;               // PCISPH: the predicted positions written (init, force)
    V2 const* xstar_in;      // PCISPH: the predicted positions of the iteration
    V2 const* xstar_alt;     // PCISPH: the other x* buffer (the final kernel)
    V2 const* vstar_in;
    V2* vstar_out;
    V2 const* vstar_alt;
    V2* force;
    // External forces per sorted slot when a scene's force model is not the
    // default gravity (or a host callback supplies forces); nullptr otherwise,
    // and the kernels that integrate the force use gravity directly.
    V2* fext;
    int store_geometry;
    // Tuning choices (results do not depend on them): neighbour batch depth of
    // the sweeps (2, 4 or 8) and the XCD-aware workgroup -> tile remap.
    int nb;
    int xcd_remap;
    SkinView skin;
};

// BUILD_AP: density and the external-force velocity update (IISPH_AP and
// IISPH_PSEUDO_AP), without IISPH's d_ii.
enum BuildMode : int { BUILD_LISTS = 0, BUILD_IISPH = 1, BUILD_DFSPH = 2, BUILD_AP = 3 };
enum LoopKind : int { LOOP_IISPH = 0, LOOP_DIVERGENCE = 1, LOOP_DENSITY = 2, LOOP_PCISPH = 3 };

//------------------------------------------------------------------------------
// sort.hip
size_t scan_scratch_words(size_t n);
size_t radix_counts_words(size_t n);
void launch_exclusive_scan(uint32_t const* in, uint32_t* out, uint32_t n, uint32_t* scratch,
                           uint32_t* total, hipStream_t s);
// Same, then every input element set to zero (in != out).
void launch_exclusive_scan_clear(uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch,
                                 uint32_t* total, hipStream_t s);
// Exclusive scan of n bin counts giving each bin's first slot (first) and the
// number of nonempty bins before it (cells), clearing the counts; *total and
// *cell_total receive the sums. scratch holds 2 * scan_scratch_words(n) words.
void launch_cell_scan_clear(uint32_t* counts, uint32_t* first, uint32_t* cells, uint32_t n,
                            uint32_t* scratch, uint32_t* total, uint32_t* cell_total,
                            hipStream_t s);
int radix_pass_count(uint64_t kmin, uint64_t kmax);
// Stable LSD sort of (key, val): key_passes 8-bit digits of the key, then
// scene_passes 8-bit digits of scene_of[val] (scene the most significant).
// Returns 1 when the result is in keys1/vals1, else 0.
int launch_radix_sort(uint64_t* keys0, uint32_t* vals0, uint64_t* keys1, uint32_t* vals1,
                      uint32_t n, int key_passes, uint32_t const* scene_of, int scene_passes,
                      uint32_t* counts, uint32_t* scan_scratch, hipStream_t s);

//------------------------------------------------------------------------------
// sph_kernels.hip: spatial index of a point set (fluid or solid).
struct IndexArgs {
    V2 const* pos;             // original order
    uint32_t n;
    SceneDev const* scenes;    // cell size of scene s is scenes[s].c.cell
    uint32_t const* scene_of;  // original index -> scene
    uint64_t* keys;
    uint32_t* vals;
    int2* grid;                // optional
    KeyState* keystate;
    unsigned long long* key_partials;  // 2 * KEY_PARTS: per-workgroup key min, max
    // Optional start of a sub-step (per-scene CFL dt, cleared statistics),
    // done by the key-range kernel when step_stats is set.
    SceneDev* step_scenes;
    SceneStats* step_stats;
    uint32_t step_nscenes;
    // cells
    int key_passes;
    int check_keys;
    uint64_t const* sorted_keys;
    uint32_t const* sorted_vals;
    uint32_t* heads;
    uint32_t const* head_scan;
    uint32_t* block_index;
    uint32_t* cell_begin;
    uint32_t* cell_end;
    uint64_t* cell_key;
    uint32_t* cell_tag;
    uint32_t* orig;
    uint32_t* rank;
    // dense cell index (dense_bins = the scenes' table bins; 0: the radix path)
    uint64_t dense_bins;
    DenseScene const* dscene;  // [scenes] table of each scene
    uint32_t* dcount;       // [dense_bins + 1] per-bin counts (zero between runs)
    uint32_t* doffset;      // [dense_bins + 2] exclusive scan of dcount, then n
    uint32_t* dcell;        // [dense_bins + 1] cells (nonempty bins) before each bin
    uint32_t* drank;        // per point (original order): rank within its bin
    // optional: each scene's largest key (written at its last sorted slot)
    uint64_t* scene_kmax;
    // gathers into sorted order (optional)
    V2 const* vel_in;
    V2* pos_out;
    V2* vel_out;
    float const* kv_in;
    float const* kd_in;
    float* kv_out;
    float* kd_out;
};

// Cell windows of one index against itself (win) and optionally another
// index (owin), e.g. fluid cells against the solid cells.
struct WindowArgs {
    CellList self;           // windows of these cells against themselves -> win
    CellList other;          // ranges of the other cells
    CellSet other_set;       // lookup of the other cells; slots == nullptr: none
    uint32_t max_cells;      // launch bound
    // Dense index of the self cells (dense_bins != 0): a neighbour cell's
    // range is [doffset[bin], doffset[bin + 1]) for keys inside its scene's
    // table (bin = dense_bin).
    uint64_t dense_bins;
    DenseScene const* dscene;
    uint32_t const* doffset;
    // With the dense index: one bit per bin set where the other index (the
    // solids) has a cell with that key, so most cells skip the set lookups.
    uint32_t const* other_bits;
    uint2* win;
    uint2* owin;
    unsigned int* report;    // optional: receives the cell count
};
void launch_cell_windows(WindowArgs const& w, hipStream_t s);

// Keys, grid coordinates and the key range (KeyState::kmin/kmax, with the
// sub-step start of every scene); with the dense index (a.dense_bins set) the
// range is folded by launch_dense_scatter, which must follow.
void launch_keys(IndexArgs const& a, hipStream_t s);

// Skin candidate lists at a rebuild (after the index and the build of the
// sub-step, dense index only): every fluid slot's fluid and solid candidates
// into b.skin (cf, nf, cs, ns, xref). w: the sub-step's windows arguments
// (its dense tables, and the solid cells with their bitmap and set).
void launch_skin_candidates(Batch const& b, WindowArgs const& w, hipStream_t s);
// Start of a sub-step that reuses the candidates: each scene's sub-step start
// (as the key-range kernel does) and the build counters cleared; the key
// range of the last indexing run stays.
void launch_skin_begin(Batch const& b, hipStream_t s);
// Positions and velocities (and the DFSPH stiffnesses when kv_in is set) into
// the last rebuild's sorted order, and the displacement test against
// b.skin.xref (its largest ratio per tile, then folded into
// KeyState::skin_disp).
void launch_skin_gather(Batch const& b, float const* kv_in, float const* kd_in, hipStream_t s);
// Dense index: scatter of (key, index) to doffset[bin] + drank into
// sorted_keys / sorted_vals (bin = dense_bin of the point's scene); its first
// workgroup also folds the key range that launch_keys left as partials.
void launch_dense_scatter(IndexArgs const& a, hipStream_t s);
void launch_cell_heads(IndexArgs const& a, hipStream_t s);
void launch_cells_finalize(IndexArgs const& a, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

// Per-sub-step resets: stats, residual slots, vmax2 of every scene; the CFL dt
// of scenes that use one.
void launch_step_begin(Batch const& b, hipStream_t s);
// Per-scene energy sums over the tile partials written by the final kernel.
void launch_scene_totals(Batch const& b, int energy, hipStream_t s);
// scene_of of every slot covered by the tiles.
void launch_scene_of(Tile const* tiles, uint32_t ntiles, uint32_t* scene_of, hipStream_t s);
// Pool repacking: segments {src, dst, n} of the state arrays into the scratch
// arrays, total slots from seg[0].dst on.
void launch_repack(uint3 const* seg, uint32_t nseg, uint32_t total, V2 const* pos, V2 const* vel,
                   float const* kv, float const* kd, V2* pos_o, V2* vel_o, float* kv_o,
                   float* kd_o, hipStream_t s);
// Start a solver loop: every non-empty scene active; KeyState::active_count set.
void launch_loop_begin(Batch const& b, hipStream_t s);
// After iteration `it` of loop `kind`: each active scene applies the
// reference's exit test to its residual slot and drops out when it passes.
void launch_loop_decide(Batch const& b, int kind, int it, int slot, hipStream_t s);

// cap: LDS list capacity (one of BUILD_CAPS). Below 64 a particle whose list
// would exceed it is deferred and rebuilt by a full-capacity pass.
// ev (optional, kernel timing): start/stop events of the build kernel and of
// its deferred pass, [4].
void launch_build(int mode, int cap, Batch const& b, hipStream_t s,
                  hipEvent_t const* ev = nullptr);
// a_ii, rho* and the pressure loop's first iteration (pressures +0) in one
// sweep, into pr_out.
void launch_iisph_aii_beta0(Batch const& b, hipStream_t s);
void launch_iisph_sdp(Batch const& b, ActiveTiles const& at, hipStream_t s);
void launch_iisph_beta_p(Batch const& b, ActiveTiles const& at, hipStream_t s);
// IISPH_AP (pseudo = 0) and IISPH_PSEUDO_AP (pseudo = 1): diagonals and
// density stars with the loop start and the first iteration (pressures +0);
// pressure accelerations of active scenes; one in-place Jacobi iteration of
// active scenes; and the final accelerations with the integration (every
// scene), writing the state back in original order.
void launch_ap_diag0(Batch const& b, int pseudo, hipStream_t s);
void launch_ap_accel(Batch const& b, ActiveTiles const& at, hipStream_t s);
void launch_ap_iterate(Batch const& b, int pseudo, ActiveTiles const& at, hipStream_t s);
void launch_ap_final(Batch const& b, int pseudo, hipStream_t s);
// Apply pressure accelerations (pressure.cpp:9-80) with p_in as the pressure,
// for active scenes (or all scenes when all_scenes != 0).
// with_gravity: first integrate the external gravity force (dynamics.cpp:24-37).
// finalize: then x += dt v and scatter state to original order (dynamics.cpp:46-51).
// all_scenes = 0: scenes still iterating only, over the tiles of `at` (every
// tile when at.n = 0); all_scenes = 1: every tile.
void launch_apply(Batch const& b, int with_gravity, int finalize, int all_scenes, hipStream_t s,
                  ActiveTiles const& at = ActiveTiles{});
void launch_finalize(Batch const& b, int dfsph, hipStream_t s);
void launch_dfsph_div(Batch const& b, ActiveTiles const& at, hipStream_t s);
void launch_dfsph_dens(Batch const& b, ActiveTiles const& at, hipStream_t s);
void launch_pci_init(Batch const& b, hipStream_t s);
// colors.cpp:9-20 target_density_colors from the densities of the last
// sub-step, RGBA per particle in original order.
void launch_colors(Batch const& b, float4* out, hipStream_t s);
// colors.cpp:57-70 convert_colors.
void launch_convert_colors(float4 const* in, uchar4* out, uint32_t n, hipStream_t s);
// External forces of every fluid slot into b.fext (forces.cpp:11-149): the
// host-supplied forces already in b.fext (host_forces != 0) or zero, then each
// scene's force models. apply != 0 (IISPH, IISPH_AP, IISPH_PSEUDO_AP): instead
// of only storing, also integrate them into the velocities (dynamics.cpp:24-37).
void launch_external_forces(Batch const& b, int apply, int host_forces, hipStream_t s);
void launch_pci_density(Batch const& b, ActiveTiles const& at, hipStream_t s);
void launch_pci_force(Batch const& b, ActiveTiles const& at, hipStream_t s);
void launch_pci_finalize(Batch const& b, hipStream_t s);
void launch_vmax2(Batch const& b, hipStream_t s);
void launch_scatter_f32(float const* src_sorted, uint32_t const* orig, float* dst, uint32_t n,
                        hipStream_t s);
// src_sorted[t * src_stride] -> dst[orig[t]] (src_stride in V2 units).
void launch_scatter_v2(V2 const* src_sorted, uint32_t src_stride, uint32_t const* orig, V2* dst,
                       uint32_t n, hipStream_t s);
// Solid volumes (solids.cpp:52-74) from the solid-solid lists (sorted solid slots).
void launch_solid_volumes(Batch const& b, uint32_t n, hipStream_t s);
// Neighbour lists exported in original order, particle-major [n][64]; indices
// are made scene-local by subtracting the neighbour scene's offset.
void launch_export_lists(Lists const& l, uint32_t n, uint32_t stride, uint32_t const* orig_self,
                         uint32_t const* orig_other, uint32_t const* other_scene_of,
                         SceneDev const* scenes, int solid_other, uint8_t* counts, uint32_t* idx,
                         float* dist, V2* vec, hipStream_t s);

}  // namespace emsph
