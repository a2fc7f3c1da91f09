// sph2d_box_headless on the GPU path: the reference's headless driver
// (sph2d_box/bin/headless_main.cpp:12-33) written against the C++ host API of
// include/emerald_sph2d_box.hpp. Same flags, same per-frame print.
#include <chrono>
#include <cstdio>
#include <exception>

#include "../../include/emerald_sph2d_box.hpp"

int main(int argc, char* argv[]) {
    using namespace emerald_amd::sph2d_box;
    try {
        auto const params = parse_parameters(argc, argv);
        EZ_EXAMPLE_SIM sim{params, world_walls_initial_solid_state, dam_break_initial_state};
        std::printf("Initial dam break state particle count: %zu, solids: %zu\n",
                    sim.state.positions.size(), sim.solid_state.positions.size());
        for (int step = 0; step < params.num_batch_steps; ++step) {
            auto const start = std::chrono::high_resolution_clock::now();
            auto const fs = sim.step();
            auto const end = std::chrono::high_resolution_clock::now();
            std::printf("%s Step: %d, sub_steps: %d, iterations: %d, duration: %f ms\n",
                        params.method.c_str(), step, fs.sub_steps, fs.total_iterations,
                        std::chrono::duration<double, std::milli>{end - start}.count());
        }
        auto const& st = sim.fetch_state();
        double ke = 0.0;
        for (auto const& v : st.velocities) { ke += 0.5 * (double(v.x) * v.x + double(v.y) * v.y); }
        std::printf("final kinetic energy per unit mass: %.9g\n", ke);
    } catch (std::exception const& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
