// Exercises include/emerald_sph2d_box.hpp (the C++ mirror of sph2d_box) the
// way the reference's callers use sph2d_box: EZ_EXAMPLE_SIM with a host
// User_forces_function lambda (gravity written by hand) and a host
// User_colors_function lambda, one frame; and a second simulation with the
// default device gravity. Writes both final states as snapshots (argv[1],
// argv[2]) for the Python test to compare, and prints one summary line.
#include <cstdio>
#include <vector>

#include "../../include/emerald_sph2d_box.hpp"

int main(int argc, char* argv[]) {
    using namespace emerald_amd::sph2d_box;
    if (argc < 3) {
        std::fprintf(stderr, "usage: api_check host.emsph device.emsph\n");
        return 2;
    }
    try {
        Parameters params;
        params.method = "iisph";
        int calls = 0;
        size_t colored = 0;
        float const g = params.gravity;
        User_forces_function gravity_by_hand = [&](size_t n, float, float, float m, float,
                                                   V2f* f, V2f const*, V2f const*) {
            ++calls;
            for (size_t i = 0; i < n; ++i) { f[i].y -= m * g; }
        };
        User_colors_function speed_colors = [&](size_t n, float, float, float, C4f* c,
                                                V2f const*, V2f const* v, float const*) {
            colored = n;
            for (size_t i = 0; i < n; ++i) { c[i] = C4f{v[i].x, v[i].y, 0.25f, 1.0f}; }
        };
        EZ_EXAMPLE_SIM host_sim{params, world_walls_initial_solid_state, dam_break_initial_state,
                                host_forces(gravity_by_hand, g), speed_colors};
        EZ_EXAMPLE_SIM device_sim{params, world_walls_initial_solid_state,
                                  dam_break_initial_state};
        auto const fs = host_sim.step();
        auto const fd = device_sim.step();
        State const& a = host_sim.fetch_state();
        State const& b = device_sim.fetch_state();
        for (int k = 0; k < 2; ++k) {
            State const& s = k ? b : a;
            size_t const n = s.positions.size();
            std::vector<unsigned char> buf(emsph_snapshot_bytes(n, 3u));
            size_t written = 0;
            check(emsph_snapshot_encode(&s.positions[0].x, &s.velocities[0].x, &s.colors[0].r, n,
                                        host_sim.time, EMSPH_IISPH, buf.data(), buf.size(),
                                        &written),
                  "encode");
            check(emsph_snapshot_write_file(argv[1 + k], buf.data(), written), "write");
        }
        bool colors_ok = colored == a.positions.size() && a.colors.size() == colored &&
                         a.colors[0].r == a.velocities[0].x && a.colors[0].b == 0.25f;
        std::printf("api_check sub_steps %d %d callbacks %d colors %s\n", fs.sub_steps,
                    fd.sub_steps, calls, colors_ok ? "ok" : "bad");
        return (calls == fs.sub_steps && colors_ok) ? 0 : 1;
    } catch (std::exception const& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
