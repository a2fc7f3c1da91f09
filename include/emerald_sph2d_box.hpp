// emerald_sph2d_box.hpp - C++ host API of the gfx950 2D SPH step, mirroring
// the reference's sph2d_box interface over the C ABI of emerald_sph.h.
//
//   Parameters, parse_parameters          sph2d_box/parameters.h:12-43, parameters.cpp:10-44
//   Simulation_config                     sph2d_box/config.h:7-18
//   State, Solid_state, Temp_data         sph2d_box/state.h:13-69
//   world_walls_initial_solid_state       sph2d_box/solids.cpp:139-213
//   dam_break_initial_state               sph2d_box/initial_state.cpp:14-97
//   iisph/pcisph/dfsph_p_simulation_step  iisph.h:13-19, pcisph.h:12-18, dfsph_p.h:15-21
//   EZ_EXAMPLE_SIM                        sph2d_box/simulation.h:24-44
//
//   User_forces_function, default_gravity_forces   sph2d_box/forces.h:13-21, forces.cpp:62-83
//   User_colors_function, default_target_density_colors  sph2d_box/colors.h:13-21,
//                                          colors.cpp:42-55
//
// Differences that follow from device residency: Temp_data owns the device
// context (the reference's caller-owned scratch that persists across calls);
// default_gravity_forces and the other forces.cpp models are device force
// models (Forces::device), while a host User_forces_function runs through a
// per-sub-step copy round trip; default_target_density_colors is computed on
// the device (an empty User_colors_function stands for it). Errors throw
// std::runtime_error where the reference aborts (EMLD_FAIL).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "emerald_sph.h"

namespace emerald_amd::sph2d_box {

struct V2f {
    float x = 0.0f;
    float y = 0.0f;
};
static_assert(sizeof(V2f) == 8, "V2f is two packed floats");

using flicks = int64_t;  // 1 / 705,600,000 s (util/flicks.h:88-116)
constexpr flicks k_flicks_one_ninetieth_of_second = 705600000 / 90;

inline void check(emsph_status s, char const* where) {
    if (s != EMSPH_OK) {
        throw std::runtime_error(std::string(where) + ": " + emsph_last_error());
    }
}

struct Parameters {
    uint64_t seed = 1;
    flicks time_per_step = 3 * k_flicks_one_ninetieth_of_second;
    int sub_steps = 30;
    float length = 1.0f;
    float support = 0.025f;
    float gravity = 9.81f;
    float viscosity = 0.01f;
    float target_density = 1000.0f;
    struct {
        int max_pressure_iterations = 30;
        float error_average_threshold = 0.005f;
        float error_max_threshold = 0.05f;
        float omega = 0.5f;
    } iisph;
    struct {
        int max_correction_iterations = 15;
        float density_error_average_threshold = 0.0035f;
        float density_error_max_threshold = 0.03f;
        float divergence_error_average_threshold = 0.0035f;
        float divergence_error_max_threshold = 0.03f;
    } dfsph;
    int num_batch_steps = 100;
    std::string method = "iisph_pseudo_ap";

    emsph_parameters to_c() const {
        emsph_parameters p{};
        p.seed = seed;
        p.time_per_step = time_per_step;
        p.sub_steps = sub_steps;
        p.length = length;
        p.support = support;
        p.gravity = gravity;
        p.viscosity = viscosity;
        p.target_density = target_density;
        p.iisph_max_pressure_iterations = iisph.max_pressure_iterations;
        p.iisph_error_average_threshold = iisph.error_average_threshold;
        p.iisph_error_max_threshold = iisph.error_max_threshold;
        p.iisph_omega = iisph.omega;
        p.dfsph_max_correction_iterations = dfsph.max_correction_iterations;
        p.dfsph_density_error_average_threshold = dfsph.density_error_average_threshold;
        p.dfsph_density_error_max_threshold = dfsph.density_error_max_threshold;
        p.dfsph_divergence_error_average_threshold = dfsph.divergence_error_average_threshold;
        p.dfsph_divergence_error_max_threshold = dfsph.divergence_error_max_threshold;
        p.num_batch_steps = num_batch_steps;
        p.method = -1;
        return p;
    }
};

// parameters.cpp:10-44 flags: --seed --sub_steps --length --support --gravity
// --viscosity --density -n/--num_steps --method.
inline Parameters parse_parameters(int argc, char* argv[]) {
    Parameters p;
    for (int i = 1; i + 1 < argc; i += 2) {
        std::string const k = argv[i];
        char const* v = argv[i + 1];
        if (k == "--seed") { p.seed = std::strtoull(v, nullptr, 10); }
        else if (k == "--sub_steps") { p.sub_steps = std::atoi(v); }
        else if (k == "--length") { p.length = std::strtof(v, nullptr); }
        else if (k == "--support") { p.support = std::strtof(v, nullptr); }
        else if (k == "--gravity") { p.gravity = std::strtof(v, nullptr); }
        else if (k == "--viscosity") { p.viscosity = std::strtof(v, nullptr); }
        else if (k == "--density") { p.target_density = std::strtof(v, nullptr); }
        else if (k == "-n" || k == "--num_steps") { p.num_batch_steps = std::atoi(v); }
        else if (k == "--method") { p.method = v; }
        else { throw std::runtime_error("unknown option " + k); }
    }
    return p;
}

inline int method_id(std::string const& m) {
    if (m == "pcisph") { return EMSPH_PCISPH; }
    if (m == "iisph") { return EMSPH_IISPH; }
    if (m == "iisph_ap") { return EMSPH_IISPH_AP; }
    if (m == "iisph_pseudo_ap") { return EMSPH_IISPH_PSEUDO_AP; }
    if (m == "dfsph_p") { return EMSPH_DFSPH_P; }
    throw std::runtime_error("Unknown simulation method: " + m);  // simulation.cpp:37
}

struct Simulation_config {
    Parameters params;
    float seconds_per_sub_step = 1.0f;
    float mass_per_particle = 1.0f;
    float draw_radius = 1.0f;
    float pressure_correction_denom = 1.0f;

    Simulation_config() = default;
    explicit Simulation_config(Parameters const& in) : params(in) {
        emsph_parameters const p = in.to_c();
        emsph_config c{};
        check(emsph_config_from_parameters(&p, &c), "Simulation_config");
        seconds_per_sub_step = c.seconds_per_sub_step;
        mass_per_particle = c.mass_per_particle;
        draw_radius = c.draw_radius;
        pressure_correction_denom = c.pressure_correction_denom;
    }
    emsph_config to_c() const {
        emsph_config c{};
        c.params = params.to_c();
        c.seconds_per_sub_step = seconds_per_sub_step;
        c.mass_per_particle = mass_per_particle;
        c.draw_radius = draw_radius;
        c.pressure_correction_denom = pressure_correction_denom;
        return c;
    }
};

struct C4f {
    float r = 0.0f, g = 0.0f, b = 0.0f, a = 0.0f;
};
static_assert(sizeof(C4f) == 16, "C4f is four packed floats");

struct State {
    std::vector<V2f> positions;
    std::vector<V2f> velocities;
    std::vector<C4f> colors;
};

// forces.h:13-21: (count, time, dt, mass per particle, radius per particle,
// forces, positions, velocities); the forces arrive zeroed.
using User_forces_function =
  std::function<void(size_t, float, float, float, float, V2f*, V2f const*, V2f const*)>;
// colors.h:13-21: (count, time, mass, radius, colors, positions, velocities,
// densities).
using User_colors_function = std::function<void(size_t, float, float, float, C4f*, V2f const*,
                                                V2f const*, float const*)>;

// The forces of a step: device force models (flags 0: none) and an optional
// host callback, accumulated callback first (emsph_set_force_model,
// emsph_set_user_forces).
struct Forces {
    emsph_force_model device{};
    User_forces_function host;
};
// forces.cpp:72-83 as the device gravity model.
inline Forces default_gravity_forces(float gravity) {
    Forces f;
    f.device.flags = EMSPH_FORCE_GRAVITY;
    f.device.gravity = gravity;
    return f;
}
// A host User_forces_function in place of gravity (keeps the device's
// gravity value for the energy diagnostic).
inline Forces host_forces(User_forces_function fn, float gravity) {
    Forces f;
    f.device.gravity = gravity;
    f.host = std::move(fn);
    return f;
}
// colors.cpp:42-55: the default colours, computed on the device.
inline User_colors_function default_target_density_colors(float /*target_density*/) {
    return {};
}

namespace detail {
inline void forces_trampoline(void* user, size_t n, float t, float dt, float m, float r,
                              float* f, float const* p, float const* v) {
    (*static_cast<User_forces_function const*>(user))(n, t, dt, m, r, reinterpret_cast<V2f*>(f),
                                                      reinterpret_cast<V2f const*>(p),
                                                      reinterpret_cast<V2f const*>(v));
}
inline void colors_trampoline(void* user, size_t n, float t, float m, float r, float* c,
                              float const* p, float const* v, float const* d) {
    (*static_cast<User_colors_function const*>(user))(n, t, m, r, reinterpret_cast<C4f*>(c),
                                                      reinterpret_cast<V2f const*>(p),
                                                      reinterpret_cast<V2f const*>(v), d);
}
}  // namespace detail

struct Solid_state {
    std::vector<V2f> positions;
};

// Caller-owned scratch that persists across calls: here the device context
// (resident state, neighbour lists, solids and their volumes).
struct Temp_data {
    struct Deleter {
        void operator()(emsph_context* c) const { emsph_destroy(c); }
    };
    std::unique_ptr<emsph_context, Deleter> ctx;
    size_t capacity = 0;
    Solid_state const* solids = nullptr;

    void ensure(Simulation_config const& config, Solid_state const& solid_state, size_t n) {
        if (ctx && solids == &solid_state && capacity >= n) { return; }
        emsph_config const c = config.to_c();
        emsph_context* raw = nullptr;
        check(emsph_create(&c, reinterpret_cast<float const*>(solid_state.positions.data()),
                           solid_state.positions.size(), n, &raw),
              "emsph_create");
        ctx.reset(raw);
        capacity = n;
        solids = &solid_state;
    }
};

inline Solid_state world_walls_initial_solid_state(Parameters const& params,
                                                   bool random_obstacles = true) {
    emsph_parameters const p = params.to_c();
    size_t const n = emsph_world_walls_solid_positions(&p, random_obstacles ? 1 : 0, nullptr, 0);
    Solid_state s;
    s.positions.resize(n);
    emsph_world_walls_solid_positions(&p, random_obstacles ? 1 : 0,
                                      reinterpret_cast<float*>(s.positions.data()), n);
    return s;
}

inline State dam_break_initial_state(Parameters const& params, Solid_state const& solid_state) {
    emsph_parameters const p = params.to_c();
    auto const* sp = reinterpret_cast<float const*>(solid_state.positions.data());
    size_t const ns = solid_state.positions.size();
    size_t const n = emsph_dam_break_positions(&p, sp, ns, nullptr, 0);
    State s;
    s.positions.resize(n);
    emsph_dam_break_positions(&p, sp, ns, reinterpret_cast<float*>(s.positions.data()), n);
    s.velocities.assign(n, V2f{});
    return s;
}

// Installs a step's forces and colour callback on the context (they must
// outlive the step: the callbacks are called during it).
inline void set_callbacks(emsph_context* ctx, Forces const& forces,
                          User_colors_function const& colors) {
    check(emsph_set_force_model(ctx, &forces.device), "emsph_set_force_model");
    check(emsph_set_user_forces(ctx, forces.host ? detail::forces_trampoline : nullptr,
                                const_cast<User_forces_function*>(&forces.host)),
          "emsph_set_user_forces");
    check(emsph_set_user_colors(ctx, colors ? detail::colors_trampoline : nullptr,
                                const_cast<User_colors_function*>(&colors)),
          "emsph_set_user_colors");
}

inline State simulation_step(int method, flicks global_time, Simulation_config const& config,
                             State&& state, Solid_state const& solid_state, Temp_data& temp,
                             Forces const& forces, User_colors_function const& colors,
                             emsph_frame_stats* stats = nullptr) {
    size_t const n = state.positions.size();
    temp.ensure(config, solid_state, n);
    set_callbacks(temp.ctx.get(), forces, colors);
    check(emsph_set_state(temp.ctx.get(), reinterpret_cast<float const*>(state.positions.data()),
                          reinterpret_cast<float const*>(state.velocities.data()), n),
          "emsph_set_state");
    check(emsph_frame_step(temp.ctx.get(), method, global_time, stats), "emsph_frame_step");
    check(emsph_get_state(temp.ctx.get(), reinterpret_cast<float*>(state.positions.data()),
                          reinterpret_cast<float*>(state.velocities.data())),
          "emsph_get_state");
    state.colors.resize(n);  // compute_colors (colors.cpp:22-40) at the frame's end
    if (n) {
        check(emsph_get_colors(temp.ctx.get(), reinterpret_cast<float*>(state.colors.data())),
              "emsph_get_colors");
    }
    set_callbacks(temp.ctx.get(), default_gravity_forces(forces.device.gravity), {});
    return std::move(state);
}

// The reference's signatures (iisph.h:13-19, pcisph.h:12-18, dfsph_p.h:15-21,
// iisph_ap.h:11-17, iisph_pseudo_ap.h:14-20).
#define EMSPH_SIMULATION_STEP(name, id)                                                     \
    inline State name(flicks t, Simulation_config const& c, State&& s, Solid_state const& ss, \
                      Temp_data& temp, Forces const& forces,                                \
                      User_colors_function const& colors = {}) {                            \
        return simulation_step(id, t, c, std::move(s), ss, temp, forces, colors);             \
    }                                                                                        \
    inline State name(flicks t, Simulation_config const& c, State&& s, Solid_state const& ss, \
                      Temp_data& temp) {                                                     \
        return name(t, c, std::move(s), ss, temp, default_gravity_forces(c.params.gravity));  \
    }
EMSPH_SIMULATION_STEP(iisph_simulation_step, EMSPH_IISPH)
EMSPH_SIMULATION_STEP(pcisph_simulation_step, EMSPH_PCISPH)
EMSPH_SIMULATION_STEP(dfsph_p_simulation_step, EMSPH_DFSPH_P)
EMSPH_SIMULATION_STEP(iisph_ap_simulation_step, EMSPH_IISPH_AP)
EMSPH_SIMULATION_STEP(iisph_pseudo_ap_simulation_step, EMSPH_IISPH_PSEUDO_AP)
#undef EMSPH_SIMULATION_STEP

// simulation.h:24-44. The state stays resident on the device between steps;
// fetch_state() copies it back (the reference's GUI reads it every frame).
struct EZ_EXAMPLE_SIM {
    enum Method { PCISPH, IISPH, IISPH_AP, IISPH_PSEUDO_AP, DFSPH_P };

    Simulation_config config;
    Solid_state solid_state;
    State state;
    Temp_data temp_data;
    Forces user_forces;
    User_colors_function user_colors;
    flicks time = 0;
    Method method = IISPH_PSEUDO_AP;

    // simulation.cpp:14-39 (forces and colours default to
    // default_gravity_forces / default_target_density_colors, as the
    // reference's bins pass them).
    EZ_EXAMPLE_SIM(Parameters const& params,
                   Solid_state (*solid_initial_state)(Parameters const&, bool),
                   State (*fluid_initial_state)(Parameters const&, Solid_state const&),
                   std::optional<Forces> in_user_forces = std::nullopt,
                   User_colors_function in_user_colors = {})
      : config(params),
        solid_state(solid_initial_state(params, true)),
        state(fluid_initial_state(params, solid_state)),
        user_forces(in_user_forces.value_or(default_gravity_forces(params.gravity))),
        user_colors(std::move(in_user_colors)) {
        method = static_cast<Method>(method_id(params.method));
        temp_data.ensure(config, solid_state, state.positions.size());
        set_callbacks(temp_data.ctx.get(), user_forces, user_colors);
        check(emsph_set_state(temp_data.ctx.get(),
                              reinterpret_cast<float const*>(state.positions.data()),
                              reinterpret_cast<float const*>(state.velocities.data()),
                              state.positions.size()),
              "emsph_set_state");
    }

    emsph_frame_stats step(std::optional<Method> override_method = std::nullopt) {
        emsph_frame_stats fs{};
        check(emsph_frame_step(temp_data.ctx.get(), override_method.value_or(method), time, &fs),
              "EZ_EXAMPLE_SIM::step");
        time += config.params.time_per_step;
        return fs;
    }

    State const& fetch_state() {
        check(emsph_get_state(temp_data.ctx.get(),
                              reinterpret_cast<float*>(state.positions.data()),
                              reinterpret_cast<float*>(state.velocities.data())),
              "emsph_get_state");
        state.colors.resize(state.positions.size());
        if (time > 0 && !state.colors.empty()) {
            check(emsph_get_colors(temp_data.ctx.get(),
                                   reinterpret_cast<float*>(state.colors.data())),
                  "emsph_get_colors");
        }
        return state;
    }
};

}  // namespace emerald_amd::sph2d_box
