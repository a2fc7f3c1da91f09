// Kernel argument blocks and host launchers for the 2D SPH hot path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_math.hpp"

namespace emsph {

constexpr int MAXN = 64;          // NEIGHBORHOOD_MAX_COUNT (sph_common/neighborhood.h:38)
constexpr int MAX_RED_SLOTS = 64; // per-sub-step residual slots (one per solver iteration)

// Cell table of one particle array (cell key -> ordinal -> [begin, end)).
struct CellTable {
    uint64_t const* keys;
    uint32_t const* vals;
    uint32_t const* begin;
    uint32_t const* end;
    uint32_t mask;
};

// Static solid particles, held in their own sorted order.
struct SolidView {
    V2 const* pos;
    V2 const* vel;
    float const* vol;
    uint32_t const* rank;   // original solid index -> sorted slot
    uint32_t const* orig;   // sorted slot -> original solid index
    CellTable table;
    uint32_t n;
};

// Neighbour lists in slot-major layout: entry k of particle t at [k * stride + t].
struct Lists {
    uint8_t* count;
    uint32_t* idx;   // neighbour sorted slot
    V2* gw;          // GradW(-vector_to) of the pair
    float* dist;     // optional (PCISPH, parity export)
    V2* vec;         // optional (PCISPH, parity export)
};

// One residual reduction: quantised integer sum and float max (as bits; every
// reduced value is >= 0 so its bit pattern orders like the float).
struct Reduction {
    unsigned long long isum;
    unsigned int max_bits;
    unsigned int pad;
};

// Per-sub-step scalars shared by the kernels.
struct StepConsts {
    KernelConsts kc;
    float cell;          // 2h: grid cell and neighbour cut-off
    float rho0;          // target density
    float fluid_volume;  // m / rho0 (sph_common/volume.cpp:7-14)
    float mass;          // mass per particle
    float gravity;
    float dt;
    float omega;         // IISPH relaxation
    float pci_denom;     // PCISPH pressure correction denominator
    float viscosity;
};

struct BuildArgs {
    StepConsts c;
    uint32_t n;
    uint32_t stride;
    V2 const* pos;        // sorted fluid positions
    V2* vel;              // sorted fluid velocities (IISPH gravity update in place)
    uint32_t const* orig; // sorted slot -> original index
    uint32_t const* rank; // original index -> sorted slot
    CellTable table;
    SolidView solids;
    Lists fl;
    Lists sl;
    float* rho;           // densities (IISPH, DFSPH)
    V2* dii;              // IISPH d_ii
    float* alpha;         // DFSPH shared coefficient
    int store_geometry;   // write dist/vec for parity export
};

enum BuildMode : int { BUILD_LISTS = 0, BUILD_IISPH = 1, BUILD_DFSPH = 2 };

// Fluid sweeps.
struct SweepArgs {
    StepConsts c;
    uint32_t n;
    uint32_t stride;
    V2 const* pos;
    V2* vel;
    SolidView solids;
    Lists fl;
    Lists sl;
    float const* rho;
    float* rho_star;
    float* aii;
    V2 const* dii;
    V2* sdp;
    float const* p_in;
    float* p_out;
    float const* alpha;
    float* kappa;
    Reduction* red;
    unsigned int* vmax2_bits;
    // finalisation (scatter back to original order)
    uint32_t const* orig;
    V2* pos_orig;
    V2* vel_orig;
    float const* kappa_v_sorted;
    float const* kappa_d_sorted;
    float* kappa_v_orig;
    float* kappa_d_orig;
    // PCISPH
    V2* vext;
    V2* xstar;
    V2 const* vstar_in;
    V2* vstar_out;
    V2* force;
};

// sort.hip
size_t scan_scratch_words(size_t n);
size_t radix_counts_words(size_t n);
void launch_exclusive_scan(uint32_t const* in, uint32_t* out, uint32_t n, uint32_t* scratch,
                           uint32_t* total, hipStream_t s);
int radix_pass_count(uint64_t kmin, uint64_t kmax);
int launch_radix_sort(uint64_t* keys0, uint32_t* vals0, uint64_t* keys1, uint32_t* vals1,
                      uint32_t n, int passes, uint32_t* counts, uint32_t* scan_scratch,
                      hipStream_t s);

// sph_kernels.hip
// Key range of one sort and the flag raised when the digits that were sorted
// do not cover it (the sub-step is then redone with the exact pass count).
struct KeyState {
    unsigned long long kmin;
    unsigned long long kmax;
    unsigned int sort_short;
    unsigned int pad;
};

struct KeysArgs {
    V2 const* pos;
    uint32_t n;
    float cell;
    uint64_t* keys;
    uint32_t* vals;
    int2* grid;                   // optional
    KeyState* keystate;           // min/max accumulated here (reset before the launch)
    Reduction* red_reset;         // optional: zeroed (MAX_RED_SLOTS entries)
    unsigned int* vmax2_reset;    // optional: zeroed
};
void launch_keys(KeysArgs const& a, hipStream_t s);

struct CellsArgs {
    KeyState* keystate;           // checked against `passes` by the first thread
    int passes;
    uint64_t const* sorted_keys;
    uint32_t const* sorted_vals;
    uint32_t n;
    uint32_t const* head_scan;    // exclusive scan of run-head flags
    uint32_t* heads;              // run-head flags (input of the scan)
    uint32_t* block_index;        // per slot cell ordinal
    uint32_t* cell_begin;
    uint32_t* cell_end;
    uint64_t* table_keys;
    uint32_t* table_vals;
    uint32_t mask;
    uint32_t* orig;               // sorted slot -> original index
    uint32_t* rank;               // original index -> sorted slot
    // gathers into sorted order (optional)
    V2 const* pos_in;
    V2 const* vel_in;
    V2* pos_out;
    V2* vel_out;
    float const* kv_in;
    float const* kd_in;
    float* kv_out;
    float* kd_out;
};
void launch_cell_heads(CellsArgs const& a, hipStream_t s);
void launch_cells_finalize(CellsArgs const& a, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

void launch_build(int mode, BuildArgs const& a, hipStream_t s);
void launch_iisph_aii(SweepArgs const& a, hipStream_t s);
void launch_iisph_sdp(SweepArgs const& a, hipStream_t s);
void launch_iisph_beta_p(SweepArgs const& a, int slot, hipStream_t s);
// Apply pressure accelerations (pressure.cpp:9-80) with p_in as the pressure.
// with_gravity: first integrate the external gravity force (dynamics.cpp:24-37).
// finalize: then x += dt v and scatter state to original order (dynamics.cpp:46-51).
void launch_apply(SweepArgs const& a, int with_gravity, int finalize, hipStream_t s);
void launch_finalize(SweepArgs const& a, int dfsph, hipStream_t s);
void launch_dfsph_div(SweepArgs const& a, int slot, hipStream_t s);
void launch_dfsph_dens(SweepArgs const& a, int slot, hipStream_t s);
void launch_pci_init(SweepArgs const& a, hipStream_t s);
void launch_pci_density(SweepArgs const& a, int slot, hipStream_t s);
void launch_pci_force(SweepArgs const& a, hipStream_t s);
void launch_pci_finalize(SweepArgs const& a, hipStream_t s);
void launch_vmax2(V2 const* vel, uint32_t n, unsigned int* out, hipStream_t s);
void launch_scatter_f32(float const* src_sorted, uint32_t const* orig, float* dst, uint32_t n,
                        hipStream_t s);
void launch_scatter_v2(V2 const* src_sorted, uint32_t const* orig, V2* dst, uint32_t n,
                       hipStream_t s);
// Solid volumes (solids.cpp:52-74) from the solid-solid lists.
void launch_solid_volumes(KernelConsts const& kc, uint32_t n, uint32_t stride,
                          Lists const& l, float* vol_sorted, hipStream_t s);
// Neighbour lists exported in original order, particle-major [n][64].
void launch_export_lists(Lists const& l, uint32_t n, uint32_t stride, uint32_t const* orig_self,
                         uint32_t const* orig_other, uint8_t* counts, uint32_t* idx,
                         float* dist, V2* vec, hipStream_t s);

}  // namespace emsph
