// Host-side scene inputs and configuration of the 2D SPH step (run once per
// scene on the CPU; no part of the per-sub-step path).
//
//   emsph_default_parameters        sph2d_box/parameters.h:12-43
//   emsph_config_from_parameters    sph2d_box/config.cpp:16-132
//   emsph_emit_solid_box            sph2d_box/solids.cpp:13-50
//   emsph_world_walls_solid_positions  sph2d_box/solids.cpp:139-206
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <unordered_map>
#include <vector>

#include "../../include/emerald_sph.h"

namespace {

constexpr int64_t kFlicksPerSecond = 705600000;
constexpr float kPi = 3.14159274101257324218750f;  // sph_common/kernels.h:12

// Cubic spline W, dW and GradW in the float operation order of kernels.h.
float spline_W(float r, float h) {
    float const q = r / h;
    float const k = h * h * h * kPi;
    if (q <= 1.0f) { return (1.0f + q * q * (q * 0.75f - 1.5f)) / k; }
    if (q <= 2.0f) {
        float const t = 2.0f - q;
        return ((t * t * t) / 4.0f) / k;
    }
    return 0.0f;
}
float spline_dW(float r, float h) {
    float const q = r / h;
    float const k = h * h * h * kPi;
    if (q <= 1.0f) { return (-3.0f * r / (h * h) + 2.25f * r * r / (h * h * h)) / k; }
    if (q <= 2.0f) {
        float const t = 2.0f - (r / h);
        return (-3.0f * (t * t) / (4.0f * h)) / k;
    }
    return 0.0f;
}
float imath_len(float x, float y) {
    float const l2 = x * x + y * y;
    if (l2 < 2.0f * std::numeric_limits<float>::min()) {
        float ax = x >= 0.0f ? x : -x;
        float ay = y >= 0.0f ? y : -y;
        float mx = ax < ay ? ay : ax;
        if (mx == 0.0f) { return 0.0f; }
        ax /= mx;
        ay /= mx;
        return mx * std::sqrt(ax * ax + ay * ay);
    }
    return std::sqrt(l2);
}
bool safe_div_v2(float nx, float ny, float d) {
    if (d == 0.0f) { return false; }
    float const ad = std::fabs(d);
    if (ad >= 1.0f) { return true; }
    float const mr = ad / std::numeric_limits<float>::min();
    return mr > std::fabs(nx) && mr > std::fabs(ny);
}

}  // namespace

extern "C" {

emsph_status emsph_default_parameters(emsph_parameters* out) {
    if (!out) { return EMSPH_ERR_INVALID_ARGUMENT; }
    std::memset(out, 0, sizeof(*out));
    out->seed = 1;
    out->time_per_step = 3 * (kFlicksPerSecond / 90);  // 3 * k_flicks_one_ninetieth_of_second
    out->sub_steps = 30;
    out->length = 1.0f;
    out->support = 0.025f;
    out->gravity = 9.81f;
    out->viscosity = 0.01f;
    out->target_density = 1000.0f;
    out->iisph_max_pressure_iterations = 30;
    out->iisph_error_average_threshold = 0.005f;
    out->iisph_error_max_threshold = 0.05f;
    out->iisph_omega = 0.5f;
    out->dfsph_max_correction_iterations = 15;
    out->dfsph_density_error_average_threshold = 0.0035f;
    out->dfsph_density_error_max_threshold = 0.03f;
    out->dfsph_divergence_error_average_threshold = 0.0035f;
    out->dfsph_divergence_error_max_threshold = 0.03f;
    out->num_batch_steps = 100;
    out->method = EMSPH_IISPH_PSEUDO_AP;
    return EMSPH_OK;
}

size_t emsph_estimate_solid_box_emission_count(float min_x, float min_y, float max_x,
                                               float max_y, float R) {
    float const sx = max_x - min_x;
    float const sy = max_y - min_y;
    auto const nx = static_cast<size_t>(std::ceil(sx / (2.0f * R)));
    auto const ny = static_cast<size_t>(std::ceil(sy / (std::sqrt(3.0f) * R)));
    return nx * ny;
}

// Hex-packed rows: every other row is shifted by R; positions accumulate in
// float exactly as the reference's loop does.
size_t emsph_emit_solid_box(float min_x, float min_y, float max_x, float max_y, float R,
                            size_t max_count, float* out_xy) {
    if (!out_xy || max_count == 0) { return 0; }
    float const dx = 2.0f * R;
    float const dy = std::sqrt(3.0f) * R;
    float const x0 = min_x + R;
    float const y0 = min_y + R;
    float const x_stop = max_x - R;
    float const y_stop = max_y - R;
    bool shifted_row = false;
    size_t count = 0;
    for (float y = y0; y <= y_stop; y += dy) {
        float x = shifted_row ? x0 : x0 + R;
        x += R;
        for (; x <= x_stop; x += dx) {
            out_xy[2 * count] = x;
            out_xy[2 * count + 1] = y;
            if (++count == max_count) { return count; }
        }
        shifted_row = !shifted_row;
    }
    return count;
}

emsph_status emsph_config_from_parameters(const emsph_parameters* params, emsph_config* out) {
    if (!params || !out) { return EMSPH_ERR_INVALID_ARGUMENT; }
    if (!(params->support > 0.0f) || params->sub_steps <= 0) {
        return EMSPH_ERR_INVALID_ARGUMENT;
    }
    out->params = *params;
    float const H = params->support;
    float const R = 0.99f * (H * 0.5f);
    out->draw_radius = R;
    // A fully packed hex block of +-15 R; the point nearest the origin is the
    // sample particle (config.cpp:38-73).
    float const lo = -15.0f * R;
    float const hi = 15.0f * R;
    size_t const cap = emsph_estimate_solid_box_emission_count(lo, lo, hi, hi, R);
    std::vector<float> xy(2 * cap);
    size_t const n = emsph_emit_solid_box(lo, lo, hi, hi, R, cap, xy.data());
    float bx = 0.0f;
    float by = 0.0f;
    float best = std::numeric_limits<float>::max();
    for (size_t i = 0; i < n; ++i) {
        float const r2 = xy[2 * i] * xy[2 * i] + xy[2 * i + 1] * xy[2 * i + 1];
        if (r2 < best) {
            best = r2;
            bx = xy[2 * i];
            by = xy[2 * i + 1];
        }
    }
    float muchness = 0.0f;
    float gsx = 0.0f;
    float gsy = 0.0f;
    float gdot = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float const ddx = bx - xy[2 * i];
        float const ddy = by - xy[2 * i + 1];
        float const len = imath_len(ddx, ddy);
        float const w = spline_W(len, H);
        if (w > 0.0f) { muchness += w; }
        float gx = 0.0f;
        float gy = 0.0f;
        if (safe_div_v2(ddx, ddy, len)) {
            float const dw = spline_dW(len, H);
            gx = (ddx / len) * dw;
            gy = (ddy / len) * dw;
        }
        gsx += gx;
        gsy += gy;
        gdot += gx * gx + gy * gy;
    }
    out->mass_per_particle = params->target_density / muchness;
    int64_t const per = params->time_per_step / params->sub_steps;
    out->seconds_per_sub_step =
      static_cast<float>(static_cast<double>(per) / static_cast<double>(kFlicksPerSecond));
    float const ratio = out->seconds_per_sub_step / muchness;
    float const beta = 2.0f * (ratio * ratio);
    out->pressure_correction_denom = beta * ((gsx * gsx + gsy * gsy) + gdot);
    return EMSPH_OK;
}

size_t emsph_world_walls_solid_positions(const emsph_parameters* params, int random_obstacles,
                                         float* out_xy, size_t capacity) {
    if (!params) { return 0; }
    float const H = params->support;
    float const R = 0.99f * (H * 0.5f);
    float const L = params->length;
    float const b = H * 3.0f;
    struct Box {
        float x0, y0, x1, y1;
    };
    std::vector<Box> boxes{
      {-b, -b, 0.0f, L + b},
      {L, -b, L + b, L + b},
      {-0.5f * b, -b, L + 0.5f * b, 0.0f},
      {-0.5f * b, L, L + 0.5f * b, L + b},
    };
    if (random_obstacles) {
        // solids.cpp:158-176. The offsets are drawn in source order
        // (x offset, y offset, half-size offset), as a left-to-right argument
        // evaluation (the reference's clang build) produces.
        std::mt19937_64 gen{1919};
        std::uniform_real_distribution<float> dist{-0.06f, 0.06f};
        int const N = 4;
        for (int by = 0; by < N; ++by) {
            float const byf = float(by + 1) / (N + 1);
            for (int bx = 0; bx < N; ++bx) {
                float const bxf = float(bx + 1) / (N + 1);
                float const ux = bxf + dist(gen);
                float const uy = byf + dist(gen);
                float const roff = dist(gen);
                float const r = 0.04f + std::abs(roff);
                boxes.push_back({L * (ux - r), L * (uy - r), L * (ux + r), L * (uy + r)});
            }
        }
    }
    size_t total = 100;
    for (auto const& bb : boxes) {
        total += emsph_estimate_solid_box_emission_count(bb.x0, bb.y0, bb.x1, bb.y1, R);
    }
    std::vector<float> xy(2 * total);
    size_t remaining = total;
    size_t at = 0;
    for (auto const& bb : boxes) {
        size_t const k =
          emsph_emit_solid_box(bb.x0, bb.y0, bb.x1, bb.y1, R, remaining, xy.data() + 2 * at);
        remaining -= k;
        at += k;
    }
    if (out_xy) {
        size_t const m = std::min(at, capacity);
        std::memcpy(out_xy, xy.data(), 2 * m * sizeof(float));
    }
    return at;
}

// initial_state.cpp:14-97: a hex block [0, L/2] x [0, 3.25 L / 4] of radius R,
// minus every fluid particle within 2R of a solid (searched in the 3x3 cells
// of size 2H around it, as the reference does).
size_t emsph_dam_break_positions(const emsph_parameters* params, const float* solid_xy,
                                 size_t solid_count, float* out_xy, size_t capacity) {
    if (!params) { return 0; }
    float const H = params->support;
    float const R = 0.99f * (H * 0.5f);
    float const L = params->length;
    float const x1 = L / 2.0f;
    float const y1 = 3.25f * L / 4.0f;
    size_t const cap = 100 + emsph_estimate_solid_box_emission_count(0.0f, 0.0f, x1, y1, R);
    std::vector<float> xy(2 * cap);
    size_t const n = emsph_emit_solid_box(0.0f, 0.0f, x1, y1, R, cap, xy.data());
    float const cell = 2.0f * H;
    auto cell_of = [cell](float x, float y) {
        int32_t const gx = static_cast<int32_t>(std::floor(x / cell));
        int32_t const gy = static_cast<int32_t>(std::floor(y / cell));
        return std::pair<int32_t, int32_t>{gx, gy};
    };
    auto pack = [](int32_t gx, int32_t gy) {
        return (static_cast<uint64_t>(static_cast<uint32_t>(gx)) << 32) |
               static_cast<uint32_t>(gy);
    };
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    for (size_t b = 0; b < solid_count; ++b) {
        auto const g = cell_of(solid_xy[2 * b], solid_xy[2 * b + 1]);
        grid[pack(g.first, g.second)].push_back(static_cast<uint32_t>(b));
    }
    size_t kept = 0;
    for (size_t i = 0; i < n; ++i) {
        float const px = xy[2 * i];
        float const py = xy[2 * i + 1];
        auto const g = cell_of(px, py);
        bool kill = false;
        for (int32_t j = g.second - 1; !kill && j <= g.second + 1; ++j) {
            for (int32_t k = g.first - 1; !kill && k <= g.first + 1; ++k) {
                auto const it = grid.find(pack(k, j));
                if (it == grid.end()) { continue; }
                for (uint32_t b : it->second) {
                    float const d = imath_len(solid_xy[2 * b] - px, solid_xy[2 * b + 1] - py);
                    if (d < 2 * R) {
                        kill = true;
                        break;
                    }
                }
            }
        }
        if (!kill) {
            if (out_xy && kept < capacity) {
                out_xy[2 * kept] = px;
                out_xy[2 * kept + 1] = py;
            }
            ++kept;
        }
    }
    return kept;
}

// Square lattice block of nx * ny particles at ((i + 1/2) s, (j + 1/2) s),
// each coordinate jittered by U(-1, 1) * jitter * s with u from SplitMix64
// (state seed + k * 0x9E3779B97F4A7C15 for draw k = 1, 2, ...; u = (x >> 40)
// * 2^-24; particle p = j * nx + i takes draws 2p + 1 (x) and 2p + 2 (y)),
// evaluated in double and rounded to float: the synthetic dam-break input of
// the benchmark workloads. No counterpart in the reference (its own dam break
// is dam_break_initial_state, a hex block).
size_t emsph_lattice_positions(uint32_t nx, uint32_t ny, double spacing, uint64_t seed,
                               double jitter, float* out_xy) {
    size_t const n = static_cast<size_t>(nx) * ny;
    if (!out_xy) { return n; }
    double const js = jitter * spacing;
    auto draw = [seed](uint64_t k) {
        uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        return static_cast<double>(z >> 40) * (1.0 / 16777216.0);
    };
    for (uint32_t j = 0; j < ny; ++j) {
        for (uint32_t i = 0; i < nx; ++i) {
            size_t const p = static_cast<size_t>(j) * nx + i;
            double const ux = draw(2 * p + 1);
            double const uy = draw(2 * p + 2);
            double const bx = (static_cast<double>(i) + 0.5) * spacing;
            double const by = (static_cast<double>(j) + 0.5) * spacing;
            out_xy[2 * p] = static_cast<float>(bx + (2.0 * ux - 1.0) * js);
            out_xy[2 * p + 1] = static_cast<float>(by + (2.0 * uy - 1.0) * js);
        }
    }
    return n;
}

}  // extern "C"
