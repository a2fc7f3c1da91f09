/* emerald_sph.h - C ABI of the MI355X (gfx950) 2D SPH step.
 *
 * Drop-in for the CPU step of the reference's sph2d_box (paths relative to
 * the reference root). The reference has no FFI layer; its boundary is the
 * in-process C++ API below, which this ABI replaces entry by entry:
 *
 *   Parameters / parse defaults   emerald/sph2d_box/parameters.h:12-43
 *   Simulation_config ctor        emerald/sph2d_box/config.cpp:16-132
 *   emit_solid_box                emerald/sph2d_box/solids.cpp:13-50
 *   world_walls_initial_solid_state + compute_neighbor_data_and_volumes
 *                                 emerald/sph2d_box/solids.cpp:76-213
 *   <method>_sub_step             iisph.cpp:23-151, pcisph.cpp:36-164,
 *                                 dfsph_p.cpp:265-374
 *   <method>_simulation_step      iisph.cpp:168-188, pcisph.cpp:166-204,
 *                                 dfsph_p.cpp:376-396 via
 *                                 adaptive_time_step.h:35-122
 *   EZ_EXAMPLE_SIM::Method        emerald/sph2d_box/simulation.h:24-44
 *   neighbour search              emerald/sph_common/neighborhood.h:198-282,
 *                                 neighborhoods_from_state.cpp:9-67
 *
 * Conventions: plain C types; positions/velocities are interleaved xy floats
 * (the reference's Imath V2f arrays); every call returns an emsph_status;
 * no exception crosses the ABI. A context owns device-resident state and one
 * HIP stream; distinct contexts may be driven from distinct threads, one
 * context is not re-entrant. Host buffers are copied in or out only by the
 * set/get entry points.
 */
#ifndef EMERALD_SPH_H
#define EMERALD_SPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum emsph_status {
    EMSPH_OK = 0,
    EMSPH_ERR_INVALID_ARGUMENT = 1,
    EMSPH_ERR_OUT_OF_MEMORY = 2,
    EMSPH_ERR_HIP = 3,
    EMSPH_ERR_UNKNOWN_METHOD = 4, /* simulation.cpp:26-38 aborts here */
    EMSPH_ERR_CAPACITY = 5,
    EMSPH_ERR_NO_DEVICE = 6
} emsph_status;

/* EZ_EXAMPLE_SIM::Method (simulation.h:26-32). IISPH_AP (iisph_ap.cpp) and
 * IISPH_PSEUDO_AP (iisph_pseudo_ap.cpp, the reference's default method) use the
 * iisph_* parameters; their frame step is (60, 4, CFL 0.4) like DFSPH_P. */
typedef enum emsph_method {
    EMSPH_PCISPH = 0,
    EMSPH_IISPH = 1,
    EMSPH_IISPH_AP = 2,
    EMSPH_IISPH_PSEUDO_AP = 3,
    EMSPH_DFSPH_P = 4
} emsph_method;

/* Parameters (parameters.h:12-43); time_per_step in flicks (1/705600000 s). */
typedef struct emsph_parameters {
    uint64_t seed;
    int64_t time_per_step;
    int32_t sub_steps;
    float length;
    float support;
    float gravity;
    float viscosity;
    float target_density;
    int32_t iisph_max_pressure_iterations;
    float iisph_error_average_threshold;
    float iisph_error_max_threshold;
    float iisph_omega;
    int32_t dfsph_max_correction_iterations;
    float dfsph_density_error_average_threshold;
    float dfsph_density_error_max_threshold;
    float dfsph_divergence_error_average_threshold;
    float dfsph_divergence_error_max_threshold;
    int32_t num_batch_steps;
    int32_t method; /* emsph_method */
} emsph_parameters;

/* Simulation_config (config.h:7-18). */
typedef struct emsph_config {
    emsph_parameters params;
    float seconds_per_sub_step;
    float mass_per_particle;
    float draw_radius;
    float pressure_correction_denom;
} emsph_config;

/* Per-sub-step solver statistics (the reference prints none of these; its
 * commented-out prints are at iisph.cpp:121-136). */
typedef struct emsph_step_stats {
    int32_t iterations;            /* IISPH/PCISPH pressure or DFSPH density iterations */
    int32_t divergence_iterations; /* DFSPH divergence iterations */
    float error_average;           /* last residual average (quantised, as the reference) */
    float error_max;               /* last residual max */
    float divergence_error_average;
    float divergence_error_max;
    float max_velocity_sq;         /* max |v|^2 after the sub-step */
    uint32_t cell_count;           /* occupied grid cells (all scenes of a batch) */
    uint32_t sort_passes;          /* 8-bit radix passes of the key sort */
    uint32_t max_neighbors;        /* longest fluid or solid neighbour list */
    float dt;                      /* dt of the sub-step */
    uint32_t reserved;
    double kinetic_energy;         /* sum of m |v|^2 / 2 after the sub-step, FP64 */
    double potential_energy;       /* sum of m g y after the sub-step, FP64 */
} emsph_step_stats;

typedef struct emsph_frame_stats {
    int32_t sub_steps;
    int32_t total_iterations;
    int32_t total_divergence_iterations;
    float last_dt;
    emsph_step_stats last;
} emsph_frame_stats;

typedef enum emsph_field {
    EMSPH_FIELD_DENSITY = 0,
    EMSPH_FIELD_DENSITY_STAR = 1,
    EMSPH_FIELD_PRESSURE = 2,
    EMSPH_FIELD_AII = 3,
    EMSPH_FIELD_DII = 4,          /* 2 floats per particle */
    EMSPH_FIELD_SUM_DIJ_PJ = 5,   /* 2 floats per particle */
    EMSPH_FIELD_ALPHA = 6,
    EMSPH_FIELD_KAPPA_DENSITY = 7,
    EMSPH_FIELD_KAPPA_DIVERGENCE = 8,
    EMSPH_FIELD_PRESSURE_FORCE = 9, /* 2 floats per particle (PCISPH) */
    /* 2 floats per particle: IISPH_AP pressure accelerations (pressure.cpp:
     * 83-178), IISPH_PSEUDO_AP pseudo-pressure displacements, of the final
     * pressures */
    EMSPH_FIELD_PRESSURE_ACCELERATION = 10,
    /* 2 floats per particle: the external force of the sub-step
     * (compute_all_external_forces, forces.cpp:11-61) */
    EMSPH_FIELD_EXTERNAL_FORCE = 11
} emsph_field;

typedef struct emsph_context emsph_context;

/* Last error message of the calling thread (never NULL). */
const char* emsph_last_error(void);
/* Library build identification (gfx target, flags). */
const char* emsph_build_info(void);
/* Number of visible HIP devices (0 when none; never fails). */
int emsph_device_count(void);

/* Parameters defaults (parameters.h:12-43). */
emsph_status emsph_default_parameters(emsph_parameters* out);
/* Simulation_config ctor: hex calibration block -> mass per particle, seconds
 * per sub-step, draw radius and the PCISPH correction denominator
 * (config.cpp:16-132). */
emsph_status emsph_config_from_parameters(const emsph_parameters* params, emsph_config* out);

/* solids.cpp:13-20 / :22-50. Returns the count. */
size_t emsph_estimate_solid_box_emission_count(float min_x, float min_y, float max_x,
                                               float max_y, float R);
size_t emsph_emit_solid_box(float min_x, float min_y, float max_x, float max_y, float R,
                            size_t max_count, float* out_xy);
/* world_walls_initial_solid_state positions (solids.cpp:139-206): the four
 * walls, plus the 16 random obstacle blocks when random_obstacles != 0.
 * out_xy may be NULL to query the count. Returns the count. */
size_t emsph_world_walls_solid_positions(const emsph_parameters* params, int random_obstacles,
                                         float* out_xy, size_t capacity);
/* dam_break_initial_state positions (initial_state.cpp:14-97): hex block
 * [0, L/2] x [0, 3.25 L/4] minus fluid within 2R of any solid. out_xy may be
 * NULL to query the count. Returns the count. */
size_t emsph_dam_break_positions(const emsph_parameters* params, const float* solid_xy,
                                 size_t solid_count, float* out_xy, size_t capacity);
/* Jittered square lattice block of nx * ny particles (the benchmark's dam-break
 * input; SplitMix64 draws, see host_scene.cpp). out_xy may be NULL to query the
 * count. Returns nx * ny. No counterpart in the reference. */
size_t emsph_lattice_positions(uint32_t nx, uint32_t ny, double spacing, uint64_t seed,
                               double jitter, float* out_xy);

/* Create a context on the current HIP device with static solids (their
 * neighbourhoods and volumes are computed on the device, solids.cpp:76-137)
 * and room for up to fluid_capacity fluid particles. */
emsph_status emsph_create(const emsph_config* config, const float* solid_xy, size_t solid_count,
                          size_t fluid_capacity, emsph_context** out);
emsph_status emsph_destroy(emsph_context* ctx);
/* Store neighbour distances/vectors for emsph_get_neighborhood (off by default). */
emsph_status emsph_set_store_geometry(emsph_context* ctx, int enable);

/* Upload/download the fluid State (state.h:13-17); vel may be NULL (zero). */
emsph_status emsph_set_state(emsph_context* ctx, const float* pos_xy, const float* vel_xy,
                             size_t count);
emsph_status emsph_get_state(emsph_context* ctx, float* pos_xy, float* vel_xy);
size_t emsph_fluid_count(const emsph_context* ctx);
size_t emsph_solid_count(const emsph_context* ctx);
/* Solid volumes in original solid order (solids.cpp:52-74). */
emsph_status emsph_get_solid_volumes(emsph_context* ctx, float* out);
/* A per-particle field of the last sub-step, original particle order. */
emsph_status emsph_get_field(emsph_context* ctx, int field, float* out);
/* Neighbourhood of the last sub-step, original order, [count][64] layout:
 * which = 0 fluid, 1 solid. dist/vec need emsph_set_store_geometry(1) before
 * the sub-step; pass NULL to skip them. */
emsph_status emsph_get_neighborhood(emsph_context* ctx, int which, uint8_t* counts,
                                    uint32_t* indices, float* distances, float* vectors_xy);

/* External forces (forces.h:13-21, forces.cpp:11-149). Each sub-step the
 * reference zeroes a per-particle force array, lets the User_forces_function
 * accumulate into it (default_gravity_forces by default) and integrates it
 * into the velocities (IISPH, IISPH_AP, IISPH_PSEUDO_AP: at time t from the
 * state at the start of the sub-step; DFSPH_P: at t + dt after the divergence
 * solve; PCISPH: at t into v_ext). Here the force of a particle is built in
 * this order from zero:
 *   1. the host callback set by emsph_set_user_forces, if any;
 *   2. the device force models of emsph_force_model.flags, in flag order:
 *      EMSPH_FORCE_GRAVITY          accumulate_gravity_forces (forces.cpp:62-69)
 *      EMSPH_FORCE_POLE             accumulate_constant_pole_attraction_forces (:85-114)
 *      EMSPH_FORCE_DRAG             accumulate_simple_drag_forces (:116-124)
 *      EMSPH_FORCE_ANTI_COUPLING    accumulate_anti_coupling_repulsive_forces
 *                                   (:126-149; over the fluid neighbour list)
 * The default model is EMSPH_FORCE_GRAVITY with the parameters' gravity and no
 * callback: the reference's default_gravity_forces. */
enum {
    EMSPH_FORCE_GRAVITY = 1,
    EMSPH_FORCE_POLE = 2,
    EMSPH_FORCE_DRAG = 4,
    EMSPH_FORCE_ANTI_COUPLING = 8
};
typedef struct emsph_force_model {
    uint32_t flags;
    float gravity;                 /* forces.y -= mass * gravity */
    float pole_magnitude;          /* attraction towards (pole_x, pole_y) plus the swirl */
    float pole_x, pole_y;
    float drag_magnitude;          /* forces += (-drag_magnitude * drag_diameter) * v */
    float drag_diameter;
    float anti_coupling_max_distance;  /* per fluid neighbour closer than this ... */
    float anti_coupling_magnitude;     /* ... += magnitude * (cos a_i, sin a_i) */
} emsph_force_model;
/* The default force model of a config (gravity from its parameters). */
emsph_status emsph_default_force_model(const emsph_config* config, emsph_force_model* out);
emsph_status emsph_set_force_model(emsph_context* ctx, const emsph_force_model* model);
emsph_status emsph_get_force_model(emsph_context* ctx, emsph_force_model* out);
/* User_forces_function (forces.h:13-21): called once per sub-step on the
 * calling thread with host arrays in original particle order; forces_xy is
 * zeroed before the call and the callback accumulates into it. Honoured by a
 * device-to-host copy of the state, the call, and a host-to-device copy of the
 * forces at the point of the sub-step where the reference calls it (slow: two
 * synchronisations per sub-step, and the sub-step's fixed segments are not
 * replayed from graphs while a callback is set). fn = NULL removes it. */
typedef void (*emsph_user_forces_fn)(void* user, size_t count, float time_seconds, float dt,
                                     float mass_per_particle, float radius_per_particle,
                                     float* forces_xy, const float* pos_xy, const float* vel_xy);
emsph_status emsph_set_user_forces(emsph_context* ctx, emsph_user_forces_fn fn, void* user);
/* State::colors (state.h:13-17): computed at the end of every emsph_frame_step
 * at the frame's end time, as compute_colors (colors.cpp:22-40) is called by
 * adaptive_time_step.h:105-110 and pcisph.cpp:186-192, from the densities of
 * the last sub-step. The default is target_density_colors (colors.cpp:9-20,
 * on the device); a host User_colors_function (colors.h:13-21) replaces it:
 * it gets the state and densities in original order and writes RGBA floats
 * (the array arrives zeroed). fn = NULL restores the default.
 * emsph_compute_colors computes them on request (callers of emsph_sub_step);
 * emsph_get_colors returns RGBA floats [count][4] and emsph_get_colors_u8 the
 * bytes of convert_colors (colors.cpp:57-70). */
typedef void (*emsph_user_colors_fn)(void* user, size_t count, float time_seconds,
                                     float mass_per_particle, float radius_per_particle,
                                     float* colors_rgba, const float* pos_xy, const float* vel_xy,
                                     const float* densities);
emsph_status emsph_set_user_colors(emsph_context* ctx, emsph_user_colors_fn fn, void* user);
emsph_status emsph_compute_colors(emsph_context* ctx, float time_seconds);
emsph_status emsph_get_colors(emsph_context* ctx, float* rgba);
emsph_status emsph_get_colors_u8(emsph_context* ctx, uint8_t* rgba);

/* Frame-start initialisation of the reference's std_adaptive_time_step and
 * DFSPH temps (constant volumes, zeroed DFSPH stiffness warm starts). */
emsph_status emsph_frame_init(emsph_context* ctx, int method);
/* One sub-step of fixed dt: the drop-in for <method>_sub_step. */
emsph_status emsph_sub_step(emsph_context* ctx, int method, float time_seconds, float dt,
                            emsph_step_stats* stats);
/* One frame: the drop-in for <method>_simulation_step (adaptive CFL sub-steps
 * in 1/1800 s quanta for IISPH/DFSPH; fixed sub-steps for PCISPH). */
emsph_status emsph_frame_step(emsph_context* ctx, int method, int64_t global_time_flicks,
                              emsph_frame_stats* stats);
/* Run n_sub_steps fixed-dt sub-steps back to back and report the elapsed
 * device time measured with HIP events on the context's stream. */
emsph_status emsph_run_sub_steps(emsph_context* ctx, int method, float dt, int n_sub_steps,
                                 float* elapsed_ms, emsph_step_stats* last);
/* Per-kernel-class device timing (HIP events around each launch; adds
 * synchronisation, for measurement only). names/ms/calls arrays of size max. */
/* Per-sub-step dt from the CFL condition (cfl.cpp:12-21): with factor > 0
 * every sub-step uses dt = min(dt_max, factor * h / sqrt(max |v|^2)) over the
 * velocities the previous sub-step left (or set_state gave), evaluated on the
 * device, where dt_max is the dt passed to the step call (emsph_sub_step,
 * emsph_run_sub_steps) or config.seconds_per_sub_step (batches). 0 turns it
 * off (the default). The dt used is reported in emsph_step_stats.dt. This is
 * the reference's cfl_maximum_time_step applied per sub-step rather than its
 * per-frame quantisation (adaptive_time_step.h:35-122, which emsph_frame_step
 * keeps). */
emsph_status emsph_set_cfl(emsph_context* ctx, float factor);
/* Replay the fixed segments of each sub-step (sort, cells, neighbour build,
 * the first and last sweeps) from captured HIP graphs (on by default; used
 * once the sort pass count is known and kernel timing is off). Results are
 * identical either way. */
emsph_status emsph_set_graphs(emsph_context* ctx, int enable);
/* Skin candidate lists for the neighbour search (on by default for contexts
 * and batches; pools do not use them). No
 * counterpart in the reference, whose lists every sub-step builds again from
 * the 3x3 cell windows (neighborhood.h:198-282); the lists here are the same
 * (the same neighbours, in the same (distance, index) order), only found
 * another way: a rebuild sub-step (the full keys, sort, cells and build, plus
 * one kernel) records for each particle the particles within the cut-off plus
 * h; the following sub-steps skip the keys, sort and cells and test only those
 * candidates with the reference's acceptance (3x3 cell window of this
 * sub-step's cells, |v| < cut-off), until some particle could have moved h/2
 * since the rebuild (predicted from the last two sub-steps; a sub-step where
 * one did is redone by the full path). Used with the dense cell index only;
 * paused for 64 sub-steps after candidates that served fewer than 6 sub-steps
 * (motion too fast to repay their build); turned off by itself when a
 * candidate list would exceed 48 fluid or 32 solid candidates, or a reused
 * list would reach the 64-entry cap (where the reference truncates in scan
 * order). */
emsph_status emsph_set_skin_lists(emsph_context* ctx, int enable);
/* Whether skin lists are on (they turn off by themselves, see above) and the
 * sub-steps since creation that built candidates and that reused them. */
emsph_status emsph_get_skin_lists(emsph_context* ctx, int* enabled, int64_t* rebuilds,
                                  int64_t* reuses);

/* Execution choices that do not change any result. */
typedef struct emsph_tuning {
    int32_t persistent_loop; /* reserved: must be 0 (a one-launch cooperative pressure
                                solve is not available; setting it returns
                                EMSPH_ERR_INVALID_ARGUMENT) */
    int32_t graphs;          /* as emsph_set_graphs */
    int32_t neighbor_batch;  /* neighbour-list entries each thread of a sweep has in flight
                                (2, 4 or 8; sums stay in list order) */
    int32_t xcd_remap;       /* workgroups take tiles so that each of the 8 XCDs walks one
                                contiguous run of Morton-sorted space (1) or in launch order */
    int32_t reserved[4];
} emsph_tuning;
emsph_status emsph_get_tuning(emsph_context* ctx, emsph_tuning* out);
emsph_status emsph_set_tuning(emsph_context* ctx, const emsph_tuning* tuning);
/* Autotuner: times every candidate tuning (neighbour batch depth x XCD remap
 * with graphs, then the best of those without graphs) on trial_sub_steps
 * sub-steps of this context's own state (restored after each trial and at the
 * end, so the simulation does not advance), three trials each in turn, and
 * keeps the candidate with the fastest best trial. times_ms (up to 8 entries,
 * may be NULL) receives each candidate's best time and n_candidates their
 * count. No counterpart in the reference. */
emsph_status emsph_autotune(emsph_context* ctx, int method, float dt, int trial_sub_steps,
                            emsph_tuning* chosen, double* times_ms, int* n_candidates);
emsph_status emsph_set_kernel_timing(emsph_context* ctx, int enable);
int emsph_get_kernel_timing(emsph_context* ctx, const char** names, double* total_ms,
                            int64_t* calls, int max);
/* Running totals since creation or the last reset, into out[4]: out[0]
 * sub-steps, out[1] pressure (IISPH/PCISPH) or density (DFSPH) iterations,
 * out[2] DFSPH divergence iterations, out[3] sub-steps redone because the sort
 * run with the previous sub-step's digit count was short (the key range grew).
 * Measurement only; no counterpart in the reference. */
emsph_status emsph_get_iteration_totals(emsph_context* ctx, int reset, int64_t* out);

/* Batches: many independent scenes (own parameters, solids, particle count and
 * dt) stepped together, one launch per sub-step stage for all of them. Every
 * scene's result is bit-identical to the same scene run alone: scenes never
 * see each other's particles, and each scene's solver loop ends at its own
 * iteration (the exit test runs per scene on the device). The batch sub-step
 * uses each scene's config.seconds_per_sub_step as its dt. No counterpart in
 * the reference (it steps one scene at a time). */
typedef struct emsph_batch emsph_batch;
emsph_status emsph_batch_create(int n_scenes, const emsph_config* configs,
                                const float* const* solid_xy, const size_t* solid_counts,
                                const size_t* fluid_counts, emsph_batch** out);
emsph_status emsph_batch_destroy(emsph_batch* b);
int emsph_batch_scene_count(const emsph_batch* b);
/* Scene state (exactly fluid_counts[scene] particles; vel may be NULL). */
emsph_status emsph_batch_set_state(emsph_batch* b, int scene, const float* pos_xy,
                                   const float* vel_xy);
emsph_status emsph_batch_get_state(emsph_batch* b, int scene, float* pos_xy, float* vel_xy);
emsph_status emsph_batch_get_field(emsph_batch* b, int scene, int field, float* out);
emsph_status emsph_batch_set_cfl(emsph_batch* b, float factor);  /* as emsph_set_cfl, per scene */
emsph_status emsph_batch_set_graphs(emsph_batch* b, int enable);  /* as emsph_set_graphs */
/* As emsph_set_skin_lists / emsph_get_skin_lists, for every scene of the batch. */
emsph_status emsph_batch_set_skin_lists(emsph_batch* b, int enable);
emsph_status emsph_batch_get_skin_lists(emsph_batch* b, int* enabled, int64_t* rebuilds,
                                        int64_t* reuses);
/* As emsph_set_kernel_timing / emsph_get_kernel_timing, over the batch's
 * launches (one launch per stage covers every scene). Measurement only. */
emsph_status emsph_batch_set_kernel_timing(emsph_batch* b, int enable);
int emsph_batch_get_kernel_timing(emsph_batch* b, const char** names, double* total_ms,
                                  int64_t* calls, int max);
/* As emsph_get_iteration_totals, summed over the batch's sub-steps (out[1],
 * out[2]: iterations launched, the maximum over the scenes of each sub-step). */
emsph_status emsph_batch_get_iteration_totals(emsph_batch* b, int reset, int64_t* out);
emsph_status emsph_batch_set_tuning(emsph_batch* b, const emsph_tuning* tuning);
/* Per-scene device force model (as emsph_set_force_model; host callbacks are
 * for contexts only). */
emsph_status emsph_batch_set_force_model(emsph_batch* b, int scene,
                                         const emsph_force_model* model);
emsph_status emsph_batch_autotune(emsph_batch* b, int method, int trial_sub_steps,
                                  emsph_tuning* chosen, double* times_ms, int* n_candidates);
emsph_status emsph_batch_frame_init(emsph_batch* b, int method);
/* One sub-step of every scene; stats (array of n_scenes, may be NULL). */
emsph_status emsph_batch_sub_step(emsph_batch* b, int method, emsph_step_stats* stats);
emsph_status emsph_batch_run_sub_steps(emsph_batch* b, int method, int n_sub_steps,
                                       float* elapsed_ms);

/* Pools: a batch whose scenes come and go, for serving. Device memory for
 * fluid_capacity fluid and solid_capacity solid particles is allocated at
 * creation (the admission budget); one solver (method) steps every live scene
 * per sub-step, each with its own parameters, solids and dt (its
 * config.seconds_per_sub_step, or the CFL dt of emsph_pool_set_cfl). Adds and
 * removes are staged and take effect at emsph_pool_commit, between sub-steps;
 * emsph_pool_add fails with EMSPH_ERR_CAPACITY when the pool would exceed its
 * capacity after the staged changes. A scene's results do not depend on the
 * other live scenes: they are bit-identical to the scene stepped alone. Scene
 * ids are positive and never reused. One pool has one HIP stream and is not
 * re-entrant; distinct pools may run from distinct threads and overlap on the
 * device. No counterpart in the reference. */
typedef struct emsph_pool emsph_pool;
emsph_status emsph_pool_create(size_t fluid_capacity, size_t solid_capacity, int max_scenes,
                               int method, emsph_pool** out);
emsph_status emsph_pool_destroy(emsph_pool* pool);
/* Stage a scene (state copied now; vel may be NULL). *id receives its id. */
emsph_status emsph_pool_add(emsph_pool* pool, const emsph_config* config, const float* solid_xy,
                            size_t solid_count, const float* pos_xy, const float* vel_xy,
                            size_t count, int* id);
/* Stage the removal of a live scene (or drop a staged add). */
emsph_status emsph_pool_remove(emsph_pool* pool, int id);
emsph_status emsph_pool_commit(emsph_pool* pool);
/* Live scene ids in slot order (the order of emsph_pool_sub_step's stats);
 * returns the live count. ids may be NULL. */
int emsph_pool_live(const emsph_pool* pool, int* ids, int max);
emsph_status emsph_pool_set_cfl(emsph_pool* pool, float factor);
emsph_status emsph_pool_set_tuning(emsph_pool* pool, const emsph_tuning* tuning);
/* One sub-step of every live scene; stats: one per live scene (may be NULL).
 * A scene removed with emsph_pool_remove rides along until the commit in the
 * sort and the neighbour build, but takes no solver iterations. */
emsph_status emsph_pool_sub_step(emsph_pool* pool, emsph_step_stats* stats);
/* As emsph_batch_get_iteration_totals, over the pool's sub-steps. */
emsph_status emsph_pool_get_iteration_totals(emsph_pool* pool, int reset, int64_t* out);
emsph_status emsph_pool_get_state(emsph_pool* pool, int id, float* pos_xy, float* vel_xy);
/* Device bytes currently held by the library (all contexts, batches, pools). */
int64_t emsph_device_bytes(void);

/* State snapshots: a fluid State (state.h:13-17: positions, velocities and
 * optionally colors) and its time in flicks in a self-describing binary
 * format, for checkpoints and served results. No counterpart in the reference.
 * Layout (little-endian): a 48-byte header - magic "EMSPHST\0", u32 version
 * (1), u32 flags (bit 0 velocities, bit 1 colors), u64 count, i64 time in
 * flicks, i32 method (or -1), u32 reserved, u64 FNV-1a hash of the payload -
 * then positions [count][2], velocities [count][2] and colors [count][4] as
 * IEEE float bit patterns (round trips are exact). Decoding checks the magic,
 * version, size and hash and fails with EMSPH_ERR_INVALID_ARGUMENT. The
 * encode/decode functions need no device. */
typedef struct emsph_snapshot_header {
    uint64_t count;
    int64_t time_flicks;
    int32_t method;
    uint32_t flags;
    uint64_t bytes;  /* size of the whole snapshot */
} emsph_snapshot_header;
/* Bytes of a snapshot of count particles with the given flags. */
size_t emsph_snapshot_bytes(size_t count, uint32_t flags);
/* vel_xy and colors_rgba may be NULL (not stored). */
emsph_status emsph_snapshot_encode(const float* pos_xy, const float* vel_xy,
                                   const float* colors_rgba, size_t count, int64_t time_flicks,
                                   int32_t method, void* out, size_t capacity, size_t* written);
emsph_status emsph_snapshot_info(const void* buf, size_t len, emsph_snapshot_header* info);
/* Outputs may be NULL; absent velocities/colors are returned as zeros. */
emsph_status emsph_snapshot_decode(const void* buf, size_t len, float* pos_xy, float* vel_xy,
                                   float* colors_rgba);
/* Writes buf to path atomically (a temporary file renamed into place). */
emsph_status emsph_snapshot_write_file(const char* path, const void* buf, size_t len);
/* A context's state (and its colors once a frame computed them) to a file,
 * and back (the count must fit the context's capacity; colors are not
 * restored: they are an output of the next frame). */
emsph_status emsph_save_snapshot(emsph_context* ctx, const char* path, int64_t time_flicks,
                                 int32_t method);
emsph_status emsph_load_snapshot(emsph_context* ctx, const char* path, int64_t* time_flicks,
                                 int32_t* method);
/* A live pool scene's state to a file (the served form's result record). */
emsph_status emsph_pool_save_snapshot(emsph_pool* pool, int id, const char* path,
                                      int64_t time_flicks);

/* Stand-alone z-index sort and neighbour search on one point set (no solids,
 * no dynamics): the reference's compute_all_neighbhorhoods fluid part. */
typedef struct emsph_search emsph_search;
emsph_status emsph_search_create(size_t capacity, emsph_search** out);
emsph_status emsph_search_destroy(emsph_search* s);
emsph_status emsph_search_set_points(emsph_search* s, const float* xy, size_t count);
/* Keys, sort, cells and 64-capped neighbour lists on the device. */
emsph_status emsph_search_run(emsph_search* s, float cell_size, int store_geometry,
                              float* elapsed_ms);
/* Results in the reference's layouts: grid coords [n][2], z keys [n], sorted
 * (key, index) pairs split in two arrays, block indices [n], block ranges
 * [blocks], neighbour counts [n] and indices/distances/vectors [n][64]. Any
 * output may be NULL. Returns the block count through block_count. */
emsph_status emsph_search_get(emsph_search* s, int32_t* grid_xy, uint64_t* keys,
                              uint64_t* sorted_keys, uint32_t* sorted_idx,
                              uint32_t* block_indices, uint32_t* block_begin,
                              uint32_t* block_end, size_t* block_count, uint8_t* counts,
                              uint32_t* indices, float* distances, float* vectors_xy);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The neighbour lists of emsph_search_run +
 * emsph_search_get (counts and indices, original order, [count][64], entries
 * past a count zero) for count <= 65536 points already on the device, written
 * into caller-owned device buffers; every intermediate buffer lives in a
 * 32 MiB device array of the library that each GPU holds until the process
 * ends. No hipMalloc, hipFree or other allocation is made. Work is enqueued on
 * `stream` (a hipStream_t; NULL: the null stream) and is not waited for. Calls
 * that overlap on one GPU share that array and race: serialise them. count
 * above 65536 fails with EMSPH_ERR_CAPACITY. */
emsph_status emsph_search_lists_noalloc(const float* xy_dev, size_t count, float cell_size,
                                        uint8_t* counts_dev, uint32_t* indices_dev,
                                        void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EMERALD_SPH_H */
