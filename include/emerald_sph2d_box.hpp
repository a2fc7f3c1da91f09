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
// Differences that follow from device residency: Temp_data owns the device
// context (the reference's caller-owned scratch that persists across calls);
// the built-in gravity force stands for default_gravity_forces; the colour
// callback (display only) is not carried. Errors throw std::runtime_error
// where the reference aborts (EMLD_FAIL).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
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

struct State {
    std::vector<V2f> positions;
    std::vector<V2f> velocities;
};

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

inline State simulation_step(int method, flicks global_time, Simulation_config const& config,
                             State&& state, Solid_state const& solid_state, Temp_data& temp,
                             emsph_frame_stats* stats = nullptr) {
    size_t const n = state.positions.size();
    temp.ensure(config, solid_state, n);
    check(emsph_set_state(temp.ctx.get(), reinterpret_cast<float const*>(state.positions.data()),
                          reinterpret_cast<float const*>(state.velocities.data()), n),
          "emsph_set_state");
    check(emsph_frame_step(temp.ctx.get(), method, global_time, stats), "emsph_frame_step");
    check(emsph_get_state(temp.ctx.get(), reinterpret_cast<float*>(state.positions.data()),
                          reinterpret_cast<float*>(state.velocities.data())),
          "emsph_get_state");
    return std::move(state);
}

inline State iisph_simulation_step(flicks t, Simulation_config const& c, State&& s,
                                   Solid_state const& ss, Temp_data& temp) {
    return simulation_step(EMSPH_IISPH, t, c, std::move(s), ss, temp);
}
inline State pcisph_simulation_step(flicks t, Simulation_config const& c, State&& s,
                                    Solid_state const& ss, Temp_data& temp) {
    return simulation_step(EMSPH_PCISPH, t, c, std::move(s), ss, temp);
}
inline State dfsph_p_simulation_step(flicks t, Simulation_config const& c, State&& s,
                                     Solid_state const& ss, Temp_data& temp) {
    return simulation_step(EMSPH_DFSPH_P, t, c, std::move(s), ss, temp);
}

// simulation.h:24-44. The state stays resident on the device between steps;
// fetch_state() copies it back (the reference's GUI reads it every frame).
struct EZ_EXAMPLE_SIM {
    enum Method { PCISPH, IISPH, IISPH_AP, IISPH_PSEUDO_AP, DFSPH_P };

    Simulation_config config;
    Solid_state solid_state;
    State state;
    Temp_data temp_data;
    flicks time = 0;
    Method method = IISPH_PSEUDO_AP;

    EZ_EXAMPLE_SIM(Parameters const& params,
                   Solid_state (*solid_initial_state)(Parameters const&, bool),
                   State (*fluid_initial_state)(Parameters const&, Solid_state const&))
      : config(params),
        solid_state(solid_initial_state(params, true)),
        state(fluid_initial_state(params, solid_state)) {
        method = static_cast<Method>(method_id(params.method));
        temp_data.ensure(config, solid_state, state.positions.size());
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
        return state;
    }
};

}  // namespace emerald_amd::sph2d_box
