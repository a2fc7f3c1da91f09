// Emerald 2D SPH oracle: a CPU restatement of the reference's sph2d_box step.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP path: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
// product library (emerald_amd/libemerald_sph.so) never links or calls it.
//
// It restates, in plain C++17, the per-particle arithmetic and ordering of the
// reference CPU path (paths relative to the reference root):
//   emerald/z_index/z_index.{h,cpp}
//   emerald/sph_common/{common.h,common.cpp,kernels.h,neighborhood.h,
//                       neighborhood.cpp,density.cpp,divergence.cpp,pressure.cpp,
//                       dynamics.cpp,volume.cpp,cfl.cpp}
//   emerald/sph2d_box/{config.cpp,solids.cpp,initial_state.cpp,
//                      neighborhoods_from_state.cpp,forces.cpp,adaptive_time_step.h,
//                      iisph.cpp,iisph_ops.cpp,pcisph.cpp,pcisph_ops.cpp,
//                      dfsph_p.cpp,dfsph_p_ops.cpp}
//   emerald/util/{safe_divide.h,functions.h,flicks.h}
// Each function cites the file:line it follows. The TBB loops
// (sph_common/common.h:24-38, grain 512) become OpenMP static chunks of 512.
// Every reduction on the path is an integer sum, a max, or a per-element write,
// so thread count never changes a result.
//
// The scalar type is a template parameter: float reproduces the reference, and
// double is the "FP64 build" used for error statistics.
//
// Imath 2.5 (Vec2 length/dot/operators) is absent from this image; its
// semantics are restated in struct V2 below. Vec2::length() and lengthTiny()
// follow Imath's published algorithm; the tiny branch (|v|^2 < 2*FLT_MIN,
// coincident particles) is checked against the exact length and its scaling
// properties by the tests, since no reference test covers it.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off -fopenmp, no fast-math).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include <omp.h>
#include <parallel/algorithm>

namespace orc {

//------------------------------------------------------------------------------
// Imath Vec2 semantics (component-wise operators, dot = x*x' + y*y').
template <class T>
struct V2 {
    T x = T(0);
    T y = T(0);
    V2() = default;
    V2(T a, T b) : x(a), y(b) {}
    V2 operator+(V2 const& o) const { return {x + o.x, y + o.y}; }
    V2 operator-(V2 const& o) const { return {x - o.x, y - o.y}; }
    V2 operator-() const { return {-x, -y}; }
    V2 operator*(T a) const { return {x * a, y * a}; }
    V2 operator/(T a) const { return {x / a, y / a}; }
    V2& operator+=(V2 const& o) { x += o.x; y += o.y; return *this; }
    V2& operator-=(V2 const& o) { x -= o.x; y -= o.y; return *this; }
    T dot(V2 const& o) const { return x * o.x + y * o.y; }
    T length2() const { return dot(*this); }
    // Imath Vec2<T>::lengthTiny
    T length_tiny() const {
        T ax = (x >= T(0)) ? x : -x;
        T ay = (y >= T(0)) ? y : -y;
        T mx = ax;
        if (mx < ay) { mx = ay; }
        if (mx == T(0)) { return T(0); }
        ax /= mx;
        ay /= mx;
        return mx * std::sqrt(ax * ax + ay * ay);
    }
    // Imath Vec2<T>::length
    T length() const {
        T l2 = dot(*this);
        if (l2 < T(2) * std::numeric_limits<T>::min()) { return length_tiny(); }
        return std::sqrt(l2);
    }
};
template <class T>
V2<T> operator*(T a, V2<T> const& v) {
    return {a * v.x, a * v.y};
}

//------------------------------------------------------------------------------
// util/safe_divide.h:14-99
template <class T>
bool is_safe_divide(T numer, T denom) {
    if (denom == 0) { return false; }
    if (std::abs(denom) >= 1) { return true; }
    T const mr = std::abs(denom) / std::numeric_limits<T>::min();
    return mr > std::abs(numer);
}
template <class T>
bool is_safe_divide(V2<T> numer, T denom) {
    if (denom == 0) { return false; }
    if (std::abs(denom) >= 1) { return true; }
    T const mr = std::abs(denom) / std::numeric_limits<T>::min();
    return (mr > std::abs(numer.x)) && (mr > std::abs(numer.y));
}

// util/functions.h:209 (sqr), :46 (mix)
template <class T>
T sqr(T a) {
    return a * a;
}
template <class T>
T mix(T a, T b, T t) {
    return (a * (T(1) - t)) + (b * t);
}

//------------------------------------------------------------------------------
// Flicks: util/flicks.h:88-116. 705,600,000 ticks per second.
constexpr int64_t FLICKS_PER_SECOND = 705600000;
inline double to_seconds(int64_t f) {
    return static_cast<double>(f) / static_cast<double>(FLICKS_PER_SECOND);
}
inline int64_t to_flicks(double s) {
    return static_cast<int64_t>(s * static_cast<double>(FLICKS_PER_SECOND));
}

//------------------------------------------------------------------------------
// z_index/z_index.h:11-24 Single_axis_index<2,B>: bit b of i goes to bit 2b.
// z_index/z_index.cpp:17-48 Z_index_2: low and high 16-bit halves through the
// LUT, giving a full 64-bit interleave of two 32-bit coordinates.
inline uint64_t single_axis_index_2(uint32_t i) {
    uint64_t out = 0;
    for (int b = 0; b < 32; ++b) {
        out |= static_cast<uint64_t>((i >> b) & 1u) << (2 * b);
    }
    return out;
}
inline uint64_t z_index(uint32_t x, uint32_t y) {
    return single_axis_index_2(x) | (single_axis_index_2(y) << 1);
}

//------------------------------------------------------------------------------
// sph_common/kernels.h:12-64. 2D cubic spline with the 3D normalisation.
template <class T>
struct K {
    static T pi() { return static_cast<T>(3.14159274101257324218750f); }
    static T wa(T q) { return T(1) + q * q * (q * (T(3) / T(4)) - (T(3) / T(2))); }
    static T wb(T q_in) {
        T const q = T(2) - q_in;
        return (q * q * q) / T(4);
    }
    static T W(T r, T h) {
        T const q = r / h;
        T const k = h * h * h * pi();
        if (q <= 1) { return wa(q) / k; }
        if (q <= 2) { return wb(q) / k; }
        return T(0);
    }
    static T dwa(T r, T h) {
        return T(-3) * r / (h * h) + (T(9) / T(4)) * r * r / (h * h * h);
    }
    static T dwb(T r, T h) { return T(-3) * sqr(T(2) - (r / h)) / (T(4) * h); }
    static T dW(T r, T h) {
        T const q = r / h;
        T const k = h * h * h * pi();
        if (q <= 1) { return dwa(r, h) / k; }
        if (q <= 2) { return dwb(r, h) / k; }
        return T(0);
    }
    static V2<T> GradW(V2<T> const& dP, T h) {
        T const r = dP.length();
        if (is_safe_divide(dP, r)) { return (dP / r) * dW(r, h); }
        return {T(0), T(0)};
    }
};

//------------------------------------------------------------------------------
// Neighbourhoods: sph_common/neighborhood.h:105-282, neighborhood.cpp:10-210.
// Indices are kept as uint32_t; for counts below 2^32 their values equal the
// reference's size_t indices.
constexpr int MAXN = 64;  // NEIGHBORHOOD_MAX_COUNT, neighborhood.h:38

template <class T>
struct Neighborhood {
    std::vector<uint8_t> counts;
    std::vector<uint32_t> indices;  // [n * 64]
    std::vector<T> distances;
    std::vector<V2<T>> vectors_to;
    std::vector<T> kernels;
    std::vector<V2<T>> kernel_gradients;
    void resize(size_t n) {
        counts.assign(n, 0);
        indices.resize(n * MAXN);
        distances.resize(n * MAXN);
        vectors_to.resize(n * MAXN);
        kernels.resize(n * MAXN);
        kernel_gradients.resize(n * MAXN);
    }
    // Only the arrays create_neighborhoods writes.
    void resize_lists(size_t n) {
        counts.assign(n, 0);
        indices.resize(n * MAXN);
        distances.resize(n * MAXN);
        vectors_to.resize(n * MAXN);
    }
};

// The spatial index of one particle array: grid coords, z keys, sorted
// (key, index) pairs, per-slot block ordinals, block ranges and the key map.
template <class T>
struct SpatialIndex {
    std::vector<int32_t> grid;  // [n * 2]
    std::vector<uint64_t> keys;
    std::vector<std::pair<uint64_t, uint32_t>> pairs;
    std::vector<uint32_t> block_indices;
    std::vector<std::pair<uint32_t, uint32_t>> blocks;
    std::unordered_map<uint64_t, std::pair<uint32_t, uint32_t>> block_map;
};

// common.cpp:43-67 compute_grid_coord(s): floor((p - min) / cell) per axis.
template <class T>
inline void grid_coord(V2<T> p, T cell, int32_t* out) {
    V2<T> const q = (p - V2<T>{T(0), T(0)}) / cell;
    out[0] = static_cast<int32_t>(std::floor(q.x));
    out[1] = static_cast<int32_t>(std::floor(q.y));
}

// neighborhoods_from_state.cpp:14-40 and solids.cpp:80-108: grid, keys,
// index pairs, sort, block indices, blocks, block map.
template <class T>
void build_spatial_index(size_t n, V2<T> const* pos, T cell, SpatialIndex<T>& si) {
    si.grid.resize(n * 2);
    si.keys.resize(n);
    si.pairs.resize(n);
    si.block_indices.resize(n);
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        grid_coord(pos[i], cell, &si.grid[2 * i]);
        // common.cpp:69-76: int32 coordinates convert to uint32.
        si.keys[i] = z_index(static_cast<uint32_t>(si.grid[2 * i]),
                             static_cast<uint32_t>(si.grid[2 * i + 1]));
        si.pairs[i] = {si.keys[i], static_cast<uint32_t>(i)};  // neighborhood.cpp:10-16
    }
    // neighborhood.cpp:18-26: lexicographic sort of unique (key, index) pairs
    // (tbb::parallel_sort there; a parallel sort here: unique pairs have one
    // sorted order).
    __gnu_parallel::sort(si.pairs.begin(), si.pairs.end());
    si.blocks.clear();
    si.block_map.clear();
    if (n == 0) { return; }
    // neighborhood.cpp:33-74: block ordinal per sorted slot (tbb::parallel_scan
    // there): key changes counted per chunk, chunk offsets, then the ordinals.
    si.block_indices[0] = 0;
    {
        int const nt = std::max(1, omp_get_max_threads());
        std::vector<uint32_t> part(static_cast<size_t>(nt) + 1, 0);
        size_t const per = (n + nt - 1) / nt;
#pragma omp parallel for schedule(static, 1)
        for (int t = 0; t < nt; ++t) {
            size_t const lo = std::max<size_t>(1, t * per);
            size_t const hi = std::min(n, (t + 1) * per);
            uint32_t c = 0;
            for (size_t i = lo; i < hi; ++i) { c += si.pairs[i - 1].first != si.pairs[i].first; }
            part[t + 1] = c;
        }
        for (int t = 0; t < nt; ++t) { part[t + 1] += part[t]; }
#pragma omp parallel for schedule(static, 1)
        for (int t = 0; t < nt; ++t) {
            size_t const lo = std::max<size_t>(1, t * per);
            size_t const hi = std::min(n, (t + 1) * per);
            uint32_t sum = part[t];
            for (size_t i = lo; i < hi; ++i) {
                sum += si.pairs[i - 1].first != si.pairs[i].first;
                si.block_indices[i] = sum;
            }
        }
    }
    size_t const block_count = static_cast<size_t>(si.block_indices[n - 1]) + 1;
    si.blocks.resize(block_count);
    // neighborhood.cpp:76-106: each run head scans to its run end.
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        uint32_t const b = si.block_indices[i];
        if (i == 0 || si.block_indices[i - 1] != b) {
            si.blocks[b].first = static_cast<uint32_t>(i);
            size_t j = i + 1;
            while (j < n && si.block_indices[j] == b) { ++j; }
            si.blocks[b].second = static_cast<uint32_t>(j);
        }
    }
    // neighborhood.cpp:108-124: key -> [begin, end).
    si.block_map.reserve(2 * block_count);
    for (size_t b = 0; b < block_count; ++b) {
        si.block_map.emplace(si.pairs[si.blocks[b].first].first, si.blocks[b]);
    }
}

struct Cand {
    uint32_t index;
    float pad;
};

// neighborhood.h:198-282 create_neighborhoods with the "true" accept function
// (neighborhood.cpp:126-152). Scan cells j-outer, i-inner; walk each cell's
// sorted slots; skip self only when the arrays are the same; stop at 64
// accepted in scan order; then sort by (distance, index).
template <class T>
void create_neighborhoods(size_t n, T max_distance, Neighborhood<T>& nb,
                          V2<T> const* self_pos, int32_t const* self_grid,
                          V2<T> const* other_pos, SpatialIndex<T> const& other,
                          bool same_array) {
#pragma omp parallel for schedule(static, 512)
    for (size_t pi = 0; pi < n; ++pi) {
        struct Nb {
            uint32_t index;
            T distance;
            V2<T> vec;
        };
        Nb list[MAXN];
        V2<T> const pos = self_pos[pi];
        int32_t const gx = self_grid[2 * pi];
        int32_t const gy = self_grid[2 * pi + 1];
        int count = 0;
        for (int32_t j = gy - 1; count < MAXN && j <= gy + 1; ++j) {
            for (int32_t i = gx - 1; count < MAXN && i <= gx + 1; ++i) {
                auto const it = other.block_map.find(
                  z_index(static_cast<uint32_t>(i), static_cast<uint32_t>(j)));
                if (it == other.block_map.end()) { continue; }
                for (uint32_t sip = it->second.first;
                     count < MAXN && sip != it->second.second; ++sip) {
                    uint32_t const oi = other.pairs[sip].second;
                    if (same_array && oi == pi) { continue; }
                    Nb& e = list[count];
                    e.index = oi;
                    e.vec = other_pos[oi] - pos;
                    e.distance = e.vec.length();
                    if (e.distance >= max_distance) { continue; }
                    ++count;
                }
            }
        }
        // neighborhood.h:166-186 neighbor_order: distance, then index.
        std::sort(list, list + count, [](Nb const& a, Nb const& b) {
            if (a.distance < b.distance) { return true; }
            if (a.distance > b.distance) { return false; }
            return a.index < b.index;
        });
        nb.counts[pi] = static_cast<uint8_t>(count);
        for (int k = 0; k < count; ++k) {
            nb.indices[pi * MAXN + k] = list[k].index;
            nb.distances[pi * MAXN + k] = list[k].distance;
            nb.vectors_to[pi * MAXN + k] = list[k].vec;
        }
    }
}

// neighborhood.cpp:154-176
template <class T>
void compute_neighbor_distances_and_vectors_to(size_t n, Neighborhood<T>& nb,
                                               V2<T> const* self_pos,
                                               V2<T> const* other_pos) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < nb.counts[i]; ++k) {
            V2<T> const d = other_pos[nb.indices[i * MAXN + k]] - self_pos[i];
            nb.vectors_to[i * MAXN + k] = d;
            nb.distances[i * MAXN + k] = d.length();
        }
    }
}

// neighborhood.cpp:178-210. Gradients are evaluated at -vector_to.
template <class T>
void compute_neighbor_kernels(size_t n, T h, Neighborhood<T>& nb) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < nb.counts[i]; ++k) {
            nb.kernels[i * MAXN + k] = K<T>::W(nb.distances[i * MAXN + k], h);
            nb.kernel_gradients[i * MAXN + k] =
              K<T>::GradW(-nb.vectors_to[i * MAXN + k], h);
        }
    }
}

//------------------------------------------------------------------------------
// Parameters (sph2d_box/parameters.h:12-43) and Simulation_config (config.h).
// User_forces_function (forces.h:13-21) as a C callback on float arrays:
// (user, count, t, dt, mass, radius, forces_xy, pos_xy, vel_xy).
using UserForces = void (*)(void*, size_t, float, float, float, float, float*, float const*,
                            float const*);

template <class T>
struct Config {
    T support = T(0.025);
    T gravity = T(9.81);
    T viscosity = T(0.01);
    T target_density = T(1000);
    int64_t time_per_step = 23520000;  // 3 * (1/90 s) in flicks
    int sub_steps = 30;
    int iisph_max_iterations = 30;
    T iisph_error_average_threshold = T(0.005f);
    T iisph_error_max_threshold = T(0.05f);
    T iisph_omega = T(0.5f);
    int dfsph_max_iterations = 15;
    T dfsph_density_avg = T(0.0035f);
    T dfsph_density_max = T(0.03f);
    T dfsph_divergence_avg = T(0.0035f);
    T dfsph_divergence_max = T(0.03f);
    T seconds_per_sub_step = T(1);
    T mass_per_particle = T(1);
    T pressure_correction_denom = T(1);
    T draw_radius = T(0);
    // External forces (forces.cpp:11-149): an optional user callback, then the
    // built-in models of `force_flags` in flag order (see external_forces).
    uint32_t force_flags = 1;
    T pole_magnitude = T(0), pole_x = T(0), pole_y = T(0);
    T drag_magnitude = T(0), drag_diameter = T(0);
    T anti_max_distance = T(0), anti_magnitude = T(0);
    UserForces user_forces = nullptr;
    void* user_data = nullptr;
};

template <class T>
struct Solids {
    std::vector<V2<T>> pos;
    std::vector<V2<T>> vel;
    std::vector<T> volumes;
    SpatialIndex<T> index;
};

template <class T>
struct Temp {
    SpatialIndex<T> index;
    Neighborhood<T> nb;
    Neighborhood<T> snb;
    std::vector<V2<T>> external_forces, diis, sum_dij_pjs, velocity_external_forces,
      position_stars, velocity_stars, pressure_forces, pressure_accelerations;
    std::vector<T> densities, density_stars, aiis, pressures, new_pressures,
      fluid_volumes, alphas, divergence_kappas, density_kappas;
};

struct Stats {
    int32_t iterations;            // pressure (IISPH/PCISPH) or density (DFSPH) iterations
    int32_t divergence_iterations; // DFSPH only
    float error_average;           // last residual average
    float error_max;               // last residual max
    float divergence_error_average;
    float divergence_error_max;
};

//------------------------------------------------------------------------------
// sph_common/volume.cpp:7-14
template <class T>
void compute_constant_volumes(size_t n, T rho0, T mass, T* vol) {
    T const v = mass / rho0;
    for (size_t i = 0; i < n; ++i) { vol[i] = v; }
}

// solids.cpp:52-74 compute_volumes (the 1.7 numerator is the reference's).
template <class T>
void compute_solid_volumes(size_t n, T h, T* vol, Neighborhood<T> const& nb) {
    T const w0 = K<T>::W(T(0), h);
    for (size_t i = 0; i < n; ++i) {
        if (!nb.counts[i]) {
            vol[i] = T(1) / w0;
            continue;
        }
        T m = w0;
        for (int k = 0; k < nb.counts[i]; ++k) { m += nb.kernels[i * MAXN + k]; }
        vol[i] = T(1.7f) / m;
    }
}

// solids.cpp:76-137 compute_neighbor_data_and_volumes.
template <class T>
void init_solids(Config<T> const& cfg, Solids<T>& s) {
    size_t const n = s.pos.size();
    T const cell = cfg.support * T(2);
    build_spatial_index(n, s.pos.data(), cell, s.index);
    Neighborhood<T> nb;
    nb.resize(n);
    create_neighborhoods(n, cell, nb, s.pos.data(), s.index.grid.data(), s.pos.data(),
                         s.index, true);
    compute_neighbor_kernels(n, cfg.support, nb);
    s.volumes.resize(n);
    compute_solid_volumes(n, cfg.support, s.volumes.data(), nb);
    s.vel.assign(n, V2<T>{});
}

// neighborhoods_from_state.cpp:9-101
template <class T>
void compute_all_neighborhoods(Config<T> const& cfg, std::vector<V2<T>> const& pos,
                               Solids<T> const& s, Temp<T>& t) {
    size_t const n = pos.size();
    T const cell = cfg.support * T(2);
    build_spatial_index(n, pos.data(), cell, t.index);
    t.nb.resize(n);
    create_neighborhoods(n, cell, t.nb, pos.data(), t.index.grid.data(), pos.data(),
                         t.index, true);
    t.snb.resize(n);
    create_neighborhoods(n, cell, t.snb, pos.data(), t.index.grid.data(), s.pos.data(),
                         s.index, false);
    compute_neighbor_kernels(n, cfg.support, t.nb);
    compute_neighbor_kernels(n, cfg.support, t.snb);
}

//------------------------------------------------------------------------------
// density.cpp:10-73. Fluid pass sets, solid pass adds a separately summed term.
template <class T>
void compute_densities(size_t n, T h, T rho0, T* rho, T const* fvol,
                       Neighborhood<T> const& nb, T const* svol,
                       Neighborhood<T> const& snb) {
    T const w0 = K<T>::W(T(0), h);
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        T d = fvol[i] * rho0 * w0;
        for (int k = 0; k < nb.counts[i]; ++k) {
            T const mass = rho0 * fvol[nb.indices[i * MAXN + k]];
            d += nb.kernels[i * MAXN + k] * mass;
        }
        rho[i] = d;
        if (snb.counts[i]) {
            T ds = T(0);
            for (int k = 0; k < snb.counts[i]; ++k) {
                T const mass = rho0 * svol[snb.indices[i * MAXN + k]];
                ds += snb.kernels[i * MAXN + k] * mass;
            }
            rho[i] += ds;
        }
    }
}

// divergence.cpp:5-74
template <class T>
void compute_divergences(size_t n, T rho0, T* div, T const* fvol, V2<T> const* fvel,
                         Neighborhood<T> const& nb, T const* svol, V2<T> const* svel,
                         Neighborhood<T> const& snb) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        V2<T> const vi = fvel[i];
        if (!nb.counts[i]) {
            div[i] = T(0);
        } else {
            T d = T(0);
            for (int k = 0; k < nb.counts[i]; ++k) {
                uint32_t const j = nb.indices[i * MAXN + k];
                T const m = rho0 * fvol[j];
                d += m * ((vi - fvel[j]).dot(nb.kernel_gradients[i * MAXN + k]));
            }
            div[i] = d;
        }
        if (snb.counts[i]) {
            T d = T(0);
            for (int k = 0; k < snb.counts[i]; ++k) {
                uint32_t const j = snb.indices[i * MAXN + k];
                T const m = rho0 * svol[j];
                d += m * ((vi - svel[j]).dot(snb.kernel_gradients[i * MAXN + k]));
            }
            div[i] += d;
        }
    }
}

// density.cpp:111-145 predict_density_stars (divergence, then clamp).
template <class T>
void predict_density_stars(size_t n, T dt, T rho0, T* rho_star, T const* rho,
                           T const* fvol, V2<T> const* fvel, Neighborhood<T> const& nb,
                           T const* svol, V2<T> const* svel,
                           Neighborhood<T> const& snb) {
    compute_divergences(n, rho0, rho_star, fvol, fvel, nb, svol, svel, snb);
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        rho_star[i] = std::max(T(0), rho[i] + (dt * rho_star[i]));
    }
}

// pressure.cpp:9-80 apply_pressures_in_place: two separate v updates.
template <class T>
void apply_pressures_in_place(size_t n, T dt, T rho0, V2<T>* vel, T const* rho,
                              T const* p, T const* fvol, Neighborhood<T> const& nb,
                              T const* svol, Neighborhood<T> const& snb) {
    auto p_over_rho2 = [](T pp, T rr) -> T {
        T const d = sqr(rr);
        return is_safe_divide(pp, d) ? pp / d : T(0);
    };
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        if (nb.counts[i]) {
            T const pi = p_over_rho2(p[i], rho[i]);
            V2<T> a{};
            for (int k = 0; k < nb.counts[i]; ++k) {
                uint32_t const j = nb.indices[i * MAXN + k];
                T const mass = rho0 * fvol[j];
                T const pj = p_over_rho2(p[j], rho[j]);
                a += -mass * (pi + pj) * nb.kernel_gradients[i * MAXN + k];
            }
            vel[i] += dt * a;
        }
        if (snb.counts[i]) {
            T const pi = p_over_rho2(p[i], rho[i]);
            V2<T> a{};
            for (int k = 0; k < snb.counts[i]; ++k) {
                uint32_t const j = snb.indices[i * MAXN + k];
                T const mass = rho0 * svol[j];
                a += -mass * pi * snb.kernel_gradients[i * MAXN + k];
            }
            vel[i] += dt * a;
        }
    }
}

// dynamics.cpp:24-37 integrate_velocities_in_place (external forces).
template <class T>
void integrate_velocities_forces(size_t n, T dt, T rho0, V2<T>* vel, T const* vol,
                                 V2<T> const* f) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        T const denom = rho0 * vol[i];
        V2<T> const numer = dt * f[i];
        if (is_safe_divide(numer, denom)) { vel[i] += numer / denom; }
    }
}

// common.h:75-81 accumulate: dst += mult * src (dynamics.cpp:46-51).
template <class T>
void accumulate(size_t n, T mult, V2<T>* dst, V2<T> const* src) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) { dst[i] += mult * src[i]; }
}

// util/random.h:112-134 Lehmer_rand_gen_64: a 128-bit state multiplied by
// 0xda942042e4dd58b5 twice at seeding and once per draw, which returns the high
// 64 bits; range [0, 2^64 - 1].
struct Lehmer64 {
    using result_type = uint64_t;
    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return std::numeric_limits<uint64_t>::max(); }
    explicit Lehmer64(uint64_t seed) : state(seed) {
        state *= 0xda942042e4dd58b5ull;
        state *= 0xda942042e4dd58b5ull;
    }
    uint64_t operator()() {
        state *= 0xda942042e4dd58b5ull;
        return static_cast<uint64_t>(state >> 64);
    }
    unsigned __int128 state;
};

// The angle forces.cpp:134-141 draws for particle i: a generator seeded with i
// and std::uniform_real_distribution<float>{-pi, pi} (the host C++ library's,
// as in the reference).
inline float anti_coupling_angle(uint64_t i) {
    std::uniform_real_distribution<float> angle_dist{static_cast<float>(-M_PI),
                                                     static_cast<float>(M_PI)};
    Lehmer64 gen{i};
    return angle_dist(gen);
}

// forces.cpp:11-61 compute_all_external_forces: zeroed forces; the user
// callback (default_gravity_forces, forces.cpp:72-83, unless replaced); then
// the built-in models of cfg.force_flags in this order: gravity
// (accumulate_gravity_forces :62-69), constant pole attraction (:85-114),
// simple drag (:116-124), anti-coupling repulsion over the fluid neighbour
// distances (:126-149). Flag 1 alone is the reference's default.
template <class T>
void external_forces(size_t n, T t, T dt, Config<T> const& cfg, V2<T> const* pos,
                     V2<T> const* vel, Neighborhood<T> const& nb, V2<T>* f) {
    for (size_t i = 0; i < n; ++i) { f[i] = V2<T>{T(0), T(0)}; }
    if (cfg.user_forces) {
        std::vector<float> fx(2 * n, 0.0f), px(2 * n), vx(2 * n);
        for (size_t i = 0; i < n; ++i) {
            px[2 * i] = static_cast<float>(pos[i].x);
            px[2 * i + 1] = static_cast<float>(pos[i].y);
            vx[2 * i] = static_cast<float>(vel[i].x);
            vx[2 * i + 1] = static_cast<float>(vel[i].y);
        }
        cfg.user_forces(cfg.user_data, n, static_cast<float>(t), static_cast<float>(dt),
                        static_cast<float>(cfg.mass_per_particle),
                        static_cast<float>(cfg.draw_radius), fx.data(), px.data(), vx.data());
        for (size_t i = 0; i < n; ++i) { f[i] = V2<T>{T(fx[2 * i]), T(fx[2 * i + 1])}; }
    }
    uint32_t const flags = cfg.force_flags;
    if (flags & 1u) {
        T const mg = cfg.mass_per_particle * cfg.gravity;
        for (size_t i = 0; i < n; ++i) { f[i].y -= mg; }
    }
    if (flags & 2u) {
        T const mag = cfg.pole_magnitude;
        V2<T> const pole{cfg.pole_x, cfg.pole_y};
        T const tz = T(0.125f) * mag;  // torque (0, 0, 0.125 magnitude)
        for (size_t i = 0; i < n; ++i) {
            V2<T> const r = pos[i] - pole;
            // Imath Vec2::normalized: zero when the length is zero.
            T const l = r.length();
            V2<T> const rn = (l == T(0)) ? V2<T>{T(0), T(0)} : V2<T>{r.x / l, r.y / l};
            f[i] -= mag * rn;
            // Imath Vec3::cross of the torque with (rn, 0), components x and y.
            V2<T> const turn{T(0) * T(0) - tz * rn.y, tz * rn.x - T(0) * T(0)};
            f[i] += turn;
        }
    }
    if (flags & 4u) {
        T const mult = -cfg.drag_magnitude * cfg.drag_diameter;
        for (size_t i = 0; i < n; ++i) { f[i] += mult * vel[i]; }
    }
    if (flags & 8u) {
        for (size_t i = 0; i < n; ++i) {
            int const c = nb.counts[i];
            if (!c) { continue; }
            T const* d = &nb.distances[i * MAXN];
            for (int j = 0; j < c; ++j) {
                if (!(d[j] < cfg.anti_max_distance)) { break; }
                float const angle = anti_coupling_angle(i);
                f[i] += cfg.anti_magnitude * V2<T>{static_cast<T>(std::cos(angle)),
                                                   static_cast<T>(std::sin(angle))};
            }
        }
    }
}

// common.h:113-148, common.cpp:78-92
template <class T>
T max_vector_squared_magnitude(size_t n, V2<T> const* v) {
    T m = T(0);
    for (size_t i = 0; i < n; ++i) { m = std::max(m, v[i].dot(v[i])); }
    return m;
}
template <class T>
T max_density_error(size_t n, T rho0, T const* rho) {
    T m = T(0);
    for (size_t i = 0; i < n; ++i) { m = std::max(m, std::max(T(0), rho[i] - rho0)); }
    return m;
}

// common.h:209-232 compute_transformed_scalar_average_and_max: int64 sum of
// (int64)(value * discretization), max via compare-exchange, average in double.
template <class T, class F>
std::pair<T, T> transformed_average_and_max(size_t n, T disc, T const* v, F&& f) {
    if (n < 1) { return {T(0), T(0)}; }
    int64_t sum = 0;
    T mx = std::numeric_limits<T>::lowest();
    for (size_t i = 0; i < n; ++i) {
        T const x = f(v[i]);
        if (mx < x) { mx = x; }
        sum += static_cast<int64_t>(x * disc);
    }
    double const num = static_cast<double>(sum);
    double const den = static_cast<double>(1000 * n);
    return {static_cast<T>(num / den), mx};
}

// cfl.cpp:12-21
template <class T>
int64_t cfl_maximum_time_step(size_t n, T h, T factor, V2<T> const* vel) {
    double const numer = factor * h;
    double const denom = std::sqrt(static_cast<double>(max_vector_squared_magnitude(n, vel)));
    double const q = is_safe_divide(numer, denom) ? numer / denom : 1.0;
    return to_flicks(q);
}

//------------------------------------------------------------------------------
// IISPH ops: sph2d_box/iisph_ops.cpp
// :20-90 d_ii = safe(-dt^2 * sum m_j gradW / rho_i^2), fluid then solid (+=).
template <class T>
void iisph_compute_diis(size_t n, T dt, T rho0, V2<T>* dii, T const* rho, T const* fvol,
                        Neighborhood<T> const& nb, T const* svol,
                        Neighborhood<T> const& snb) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        for (int pass = 0; pass < 2; ++pass) {
            Neighborhood<T> const& g = pass ? snb : nb;
            T const* vol = pass ? svol : fvol;
            if (!g.counts[i]) {
                if (!pass) { dii[i] = V2<T>{}; }
                continue;
            }
            T const density = rho[i];
            V2<T> s{};
            for (int k = 0; k < g.counts[i]; ++k) {
                T const mass = rho0 * vol[g.indices[i * MAXN + k]];
                s += mass * g.kernel_gradients[i * MAXN + k];
            }
            V2<T> const numer = -sqr(dt) * s;
            T const denom = sqr(density);
            V2<T> const d = is_safe_divide(numer, denom) ? numer / denom : V2<T>{};
            if (pass) {
                dii[i] += d;
            } else {
                dii[i] = d;
            }
        }
    }
}

// :94-221 a_ii and density stars.
template <class T>
void iisph_compute_aiis_and_density_stars(size_t n, T dt, T rho0, T* aii, T* rho_star,
                                          V2<T> const* dii, T const* rho, T const* fvol,
                                          V2<T> const* fvel, Neighborhood<T> const& nb,
                                          T const* svol, V2<T> const* svel,
                                          Neighborhood<T> const& snb) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        // fluid pass
        if (!nb.counts[i]) {
            rho_star[i] = rho[i];
            aii[i] = T(0);
        } else {
            V2<T> const d = dii[i];
            V2<T> const vi = fvel[i];
            T a = T(0);
            T dd = T(0);
            for (int k = 0; k < nb.counts[i]; ++k) {
                uint32_t const j = nb.indices[i * MAXN + k];
                V2<T> const gw = nb.kernel_gradients[i * MAXN + k];
                T const mass = fvol[j] * rho0;
                V2<T> const numer = -sqr(dt) * mass * gw;
                T const denom = sqr(rho[j]);
                if (is_safe_divide(numer, denom)) {
                    V2<T> const dij = numer / denom;
                    V2<T> const dji = -dij;
                    a += mass * ((d - dji).dot(gw));
                }
                dd += dt * mass * ((vi - fvel[j]).dot(gw));
            }
            aii[i] = a;
            rho_star[i] = rho[i] + dd;
        }
        // solid pass
        if (snb.counts[i]) {
            V2<T> const d = dii[i];
            V2<T> const vi = fvel[i];
            T a = T(0);
            T dd = T(0);
            for (int k = 0; k < snb.counts[i]; ++k) {
                uint32_t const j = snb.indices[i * MAXN + k];
                V2<T> const gw = snb.kernel_gradients[i * MAXN + k];
                T const mass = svol[j] * rho0;
                a += mass * (d.dot(gw));
                dd += dt * mass * ((vi - svel[j]).dot(gw));
            }
            aii[i] += a;
            rho_star[i] += dd;
        }
    }
}

// :225-258 sum_j d_ij p_j, fluid only.
template <class T>
void iisph_compute_sum_dij_pjs(size_t n, T dt, T rho0, V2<T>* out, T const* fvol,
                               T const* rho, T const* p, Neighborhood<T> const& nb) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        V2<T> s{};
        for (int k = 0; k < nb.counts[i]; ++k) {
            uint32_t const j = nb.indices[i * MAXN + k];
            T const mass = fvol[j] * rho0;
            V2<T> const numer = -sqr(dt) * mass * p[j] * nb.kernel_gradients[i * MAXN + k];
            T const denom = sqr(rho[j]);
            if (is_safe_divide(numer, denom)) { s += numer / denom; }
        }
        out[i] = s;
    }
}

// :261-353 betas (written into the new-pressure buffer).
template <class T>
void iisph_compute_betas(size_t n, T dt, T rho0, T* beta_out, T const* rho, T const* p,
                         V2<T> const* dii, V2<T> const* sdp, T const* fvol,
                         Neighborhood<T> const& nb, T const* svol,
                         Neighborhood<T> const& snb) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        if (!nb.counts[i]) {
            beta_out[i] = T(0);
        } else {
            T const pi = p[i];
            V2<T> const si = sdp[i];
            T b = T(0);
            for (int k = 0; k < nb.counts[i]; ++k) {
                uint32_t const j = nb.indices[i * MAXN + k];
                V2<T> const gw = nb.kernel_gradients[i * MAXN + k];
                T const mass = rho0 * fvol[j];
                V2<T> const numer = -sqr(dt) * mass * gw;
                T const denom = sqr(rho[j]);
                V2<T> const dij = is_safe_divide(numer, denom) ? numer / denom : V2<T>{};
                V2<T> const dji = -dij;
                V2<T> const sjk = sdp[j] - (dji * pi);
                b += mass * ((si - (dii[j] * p[j]) - sjk).dot(gw));
            }
            beta_out[i] = b;
        }
        if (snb.counts[i]) {
            V2<T> const si = sdp[i];
            T b = T(0);
            for (int k = 0; k < snb.counts[i]; ++k) {
                uint32_t const j = snb.indices[i * MAXN + k];
                T const mass = rho0 * svol[j];
                b += mass * (si.dot(snb.kernel_gradients[i * MAXN + k]));
            }
            beta_out[i] += b;
        }
    }
}

// :357-425 new pressures with quantised uint64 error sum and float max.
template <class T>
std::pair<T, T> iisph_compute_new_pressures(size_t n, T omega, T rho0, T* beta_in_p_out,
                                            T const* p_old, T const* alphas,
                                            T const* rho_star) {
    if (!n) { return {T(0), T(0)}; }
    uint64_t isum = 0;
    T emax = T(0);
#pragma omp parallel for schedule(static, 512) reduction(+ : isum) reduction(max : emax)
    for (size_t i = 0; i < n; ++i) {
        T const alpha = alphas[i];
        T const po = p_old[i];
        if (alpha == T(0)) {
            beta_in_p_out[i] = po;
            continue;
        }
        T const numer = rho0 - rho_star[i] - beta_in_p_out[i];
        if (!is_safe_divide(numer, alpha)) {
            beta_in_p_out[i] = po;
            continue;
        }
        T const err = std::max(T(0), -(numer - (alpha * po)) / rho0);
        emax = std::max(emax, err);
        isum += static_cast<uint64_t>(err * T(1000));
        T const pn = numer / alpha;
        beta_in_p_out[i] = std::max(T(0), mix(po, pn, omega));
    }
    double const num = static_cast<double>(isum);
    double const den = static_cast<double>(1000 * n);
    return {static_cast<T>(num / den), emax};
}

//------------------------------------------------------------------------------
// iisph.cpp:23-151 iisph_sub_step
template <class T>
void iisph_sub_step(T t, T dt, Config<T> const& cfg, std::vector<V2<T>>& pos,
                    std::vector<V2<T>>& vel, Solids<T> const& s, Temp<T>& tp, Stats* st) {
    size_t const n = pos.size();
    T const rho0 = cfg.target_density;
    compute_all_neighborhoods(cfg, pos, s, tp);
    tp.external_forces.resize(n);
    external_forces(n, t, dt, cfg, pos.data(), vel.data(), tp.nb, tp.external_forces.data());
    tp.densities.resize(n);
    compute_densities(n, cfg.support, rho0, tp.densities.data(), tp.fluid_volumes.data(),
                      tp.nb, s.volumes.data(), tp.snb);
    integrate_velocities_forces(n, dt, rho0, vel.data(), tp.fluid_volumes.data(),
                                tp.external_forces.data());
    tp.diis.resize(n);
    iisph_compute_diis(n, dt, rho0, tp.diis.data(), tp.densities.data(),
                       tp.fluid_volumes.data(), tp.nb, s.volumes.data(), tp.snb);
    tp.aiis.resize(n);
    tp.density_stars.resize(n);
    iisph_compute_aiis_and_density_stars(n, dt, rho0, tp.aiis.data(), tp.density_stars.data(),
                                         tp.diis.data(), tp.densities.data(),
                                         tp.fluid_volumes.data(), vel.data(), tp.nb,
                                         s.volumes.data(), s.vel.data(), tp.snb);
    tp.pressures.assign(n, T(0));
    tp.new_pressures.resize(n);
    tp.sum_dij_pjs.resize(n);
    int iter = 0;
    std::pair<T, T> err{T(0), T(0)};
    for (; iter < cfg.iisph_max_iterations; ++iter) {
        iisph_compute_sum_dij_pjs(n, dt, rho0, tp.sum_dij_pjs.data(), tp.fluid_volumes.data(),
                                  tp.densities.data(), tp.pressures.data(), tp.nb);
        iisph_compute_betas(n, dt, rho0, tp.new_pressures.data(), tp.densities.data(),
                            tp.pressures.data(), tp.diis.data(), tp.sum_dij_pjs.data(),
                            tp.fluid_volumes.data(), tp.nb, s.volumes.data(), tp.snb);
        err = iisph_compute_new_pressures(n, cfg.iisph_omega, rho0, tp.new_pressures.data(),
                                          tp.pressures.data(), tp.aiis.data(),
                                          tp.density_stars.data());
        std::swap(tp.pressures, tp.new_pressures);
        if (iter > 1 && err.first <= cfg.iisph_error_average_threshold &&
            err.second <= cfg.iisph_error_max_threshold) {
            break;
        }
    }
    if (st) {
        // The reference's loop variable: the break leaves iter at the index of
        // the last executed iteration; running out leaves it at the maximum.
        st->iterations = (iter < cfg.iisph_max_iterations) ? iter + 1 : iter;
        st->error_average = static_cast<float>(err.first);
        st->error_max = static_cast<float>(err.second);
    }
    apply_pressures_in_place(n, dt, rho0, vel.data(), tp.densities.data(), tp.pressures.data(),
                             tp.fluid_volumes.data(), tp.nb, s.volumes.data(), tp.snb);
    accumulate(n, dt, pos.data(), vel.data());
}

//------------------------------------------------------------------------------
// PCISPH ops: sph2d_box/pcisph_ops.cpp
// :18-38 densities = m * (W0 + sum W)
template <class T>
void pcisph_compute_densities(size_t n, T mass, T h, T* rho, Neighborhood<T> const& nb) {
    T const w0 = K<T>::W(T(0), h);
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        if (!nb.counts[i]) {
            rho[i] = mass * w0;
            continue;
        }
        T w = w0;
        for (int k = 0; k < nb.counts[i]; ++k) { w += nb.kernels[i * MAXN + k]; }
        rho[i] = mass * w;
    }
}
// :40-63
template <class T>
void pcisph_accumulate_density_from_solids(size_t n, T rho0, T* rho, T const* svol,
                                           Neighborhood<T> const& snb) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        if (!snb.counts[i]) { continue; }
        T vf = T(0);
        for (int k = 0; k < snb.counts[i]; ++k) {
            vf += svol[snb.indices[i * MAXN + k]] * snb.kernels[i * MAXN + k];
        }
        rho[i] += rho0 * vf;
    }
}
// :65-76
template <class T>
void pcisph_update_pressures(size_t n, T rho0, T denom, T* p, T const* rho) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        T const e = std::max(rho[i] - rho0, T(0));
        p[i] += e / denom;
    }
}
// :78-142 pressure + artificial viscosity forces.
template <class T>
void pcisph_accumulate_pressure_forces(size_t n, T mass, T h, T visc, V2<T>* F,
                                       Neighborhood<T> const& nb, V2<T> const* vel,
                                       T const* p, T const* rho) {
    T const H = h;
    T const M = mass;
    T const N2 = T(0.01f) * H * H;
    T const alpha = T(1);
    T const beta = T(2);
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        if (!nb.counts[i]) { continue; }
        V2<T> const v = vel[i];
        T const density = rho[i];
        T const por = p[i] / sqr(density);
        V2<T> f{};
        for (int k = 0; k < nb.counts[i]; ++k) {
            uint32_t const j = nb.indices[i * MAXN + k];
            V2<T> const r = nb.vectors_to[i * MAXN + k];
            V2<T> const gw = nb.kernel_gradients[i * MAXN + k];
            T const od = rho[j];
            f += (por + (p[j] / sqr(od))) * gw;
            V2<T> const dv = vel[j] - v;
            T const dvr = r.dot(dv);
            if (dvr < 0) {
                T const dist = nb.distances[i * MAXN + k];
                T const mu = H * dvr / (dist + N2);
                T const den_avg = (density + od) / T(2);
                f += (((-alpha * visc * mu) + (beta * sqr(mu))) / den_avg) * gw;
            }
        }
        F[i] += -(M * M) * f;
    }
}
// :144-178
template <class T>
void pcisph_accumulate_pressure_forces_from_solids(size_t n, T mass, T rho0, V2<T>* F,
                                                   T const* svol, Neighborhood<T> const& snb,
                                                   T const* p, T const* rho) {
    T const M = mass;
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        if (!snb.counts[i]) { continue; }
        T const por = p[i] / sqr(rho[i]);
        V2<T> g{};
        for (int k = 0; k < snb.counts[i]; ++k) {
            g += svol[snb.indices[i * MAXN + k]] * snb.kernel_gradients[i * MAXN + k];
        }
        F[i] += (-M * rho0 * por) * g;
    }
}

// pcisph.cpp:36-164 pcisph_sub_step
template <class T>
void pcisph_sub_step(T t, T dt, Config<T> const& cfg, std::vector<V2<T>>& pos,
                     std::vector<V2<T>>& vel, Solids<T> const& s, Temp<T>& tp, Stats* st) {
    size_t const n = pos.size();
    T const m = cfg.mass_per_particle;
    compute_all_neighborhoods(cfg, pos, s, tp);
    tp.external_forces.resize(n);
    external_forces(n, t, dt, cfg, pos.data(), vel.data(), tp.nb, tp.external_forces.data());
    tp.velocity_external_forces = vel;
    accumulate(n, dt / m, tp.velocity_external_forces.data(), tp.external_forces.data());
    tp.pressures.assign(n, T(0));
    tp.pressure_forces.resize(n);
    tp.densities.resize(n);
    tp.position_stars.resize(n);
    tp.velocity_stars.resize(n);
    int it = 0;
    T last_err = T(0);
    for (; it < 6; ++it) {
        tp.velocity_stars = tp.velocity_external_forces;
        if (it > 0) {
            accumulate(n, dt / m, tp.velocity_stars.data(), tp.pressure_forces.data());
        }
        tp.position_stars = pos;
        accumulate(n, dt, tp.position_stars.data(), tp.velocity_stars.data());
        // neighborhoods_from_state.cpp:103-128
        compute_neighbor_distances_and_vectors_to(n, tp.nb, tp.position_stars.data(),
                                                  tp.position_stars.data());
        compute_neighbor_distances_and_vectors_to(n, tp.snb, tp.position_stars.data(),
                                                  s.pos.data());
        compute_neighbor_kernels(n, cfg.support, tp.nb);
        compute_neighbor_kernels(n, cfg.support, tp.snb);
        for (size_t i = 0; i < n; ++i) { tp.pressure_forces[i] = V2<T>{}; }
        pcisph_compute_densities(n, m, cfg.support, tp.densities.data(), tp.nb);
        pcisph_accumulate_density_from_solids(n, cfg.target_density, tp.densities.data(),
                                              s.volumes.data(), tp.snb);
        pcisph_update_pressures(n, cfg.target_density, cfg.pressure_correction_denom,
                                tp.pressures.data(), tp.densities.data());
        pcisph_accumulate_pressure_forces(n, m, cfg.support, cfg.viscosity,
                                          tp.pressure_forces.data(), tp.nb,
                                          tp.velocity_stars.data(), tp.pressures.data(),
                                          tp.densities.data());
        pcisph_accumulate_pressure_forces_from_solids(n, m, cfg.target_density,
                                                      tp.pressure_forces.data(),
                                                      s.volumes.data(), tp.snb,
                                                      tp.pressures.data(), tp.densities.data());
        if (it > 1) {
            T const me = max_density_error(n, cfg.target_density, tp.densities.data());
            last_err = me / cfg.target_density;
            if (last_err <= T(0.01f)) { break; }
        }
    }
    if (st) {
        st->iterations = (it < 6) ? it + 1 : it;
        st->error_average = 0.0f;
        st->error_max = static_cast<float>(last_err);
    }
    vel = tp.velocity_external_forces;
    accumulate(n, dt / m, vel.data(), tp.pressure_forces.data());
    accumulate(n, dt, pos.data(), vel.data());
}

//------------------------------------------------------------------------------
// DFSPH_P: sph2d_box/dfsph_p_ops.cpp and dfsph_p.cpp
// dfsph_p_ops.cpp:19-84 shared coefficients (alpha).
template <class T>
void dfsph_p_compute_shared_coeffs(size_t n, T rho0, T* alpha, T const* rho, T const* fvol,
                                   Neighborhood<T> const& nb, T const* svol,
                                   Neighborhood<T> const& snb) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        int const fc = nb.counts[i];
        int const sc = snb.counts[i];
        if (!(fc + sc)) {
            alpha[i] = T(0);
            continue;
        }
        V2<T> s{};
        T ss = T(0);
        for (int k = 0; k < fc; ++k) {
            T const m = fvol[nb.indices[i * MAXN + k]] * rho0;
            V2<T> const mg = m * nb.kernel_gradients[i * MAXN + k];
            s += mg;
            ss += mg.dot(mg);
        }
        for (int k = 0; k < sc; ++k) {
            T const m = svol[snb.indices[i * MAXN + k]] * rho0;
            s += m * snb.kernel_gradients[i * MAXN + k];
        }
        T const numer = sqr(rho[i]);
        T const denom = s.dot(s) + ss;
        alpha[i] = is_safe_divide(numer, denom) ? numer / denom : T(0);
    }
}
// :87-106
template <class T>
void dfsph_p_density_pseudo_pressures(size_t n, T dt, T rho0, T* kappa, T const* rho_star,
                                      T const* alpha) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        T const dd = rho_star[i] - rho0;
        if (dd <= T(0)) {
            kappa[i] = T(0);
        } else {
            T const num = dd * alpha[i];
            T const den = sqr(dt);
            kappa[i] = is_safe_divide(num, den) ? num / den : T(0);
        }
    }
}
// :109-126
template <class T>
void dfsph_p_divergence_pseudo_pressures(size_t n, T dt, T* kappa, T const* div,
                                         T const* alpha) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        T const d = div[i];
        if (d <= T(0)) {
            kappa[i] = T(0);
        } else {
            T const num = d * alpha[i];
            kappa[i] = is_safe_divide(num, dt) ? num / dt : T(0);
        }
    }
}

// dfsph_p.cpp:61-134
template <class T>
void dfsph_p_correct_density_error(T dt, Config<T> const& cfg, std::vector<V2<T>>& vel,
                                   Solids<T> const& s, Temp<T>& tp, Stats* st) {
    size_t const n = vel.size();
    T const rho0 = cfg.target_density;
    apply_pressures_in_place(n, dt, rho0, vel.data(), tp.densities.data(),
                             tp.density_kappas.data(), tp.fluid_volumes.data(), tp.nb,
                             s.volumes.data(), tp.snb);
    int iter = 0;
    std::pair<T, T> e{T(0), T(0)};
    for (; iter < cfg.dfsph_max_iterations; ++iter) {
        predict_density_stars(n, dt, rho0, tp.density_stars.data(), tp.densities.data(),
                              tp.fluid_volumes.data(), vel.data(), tp.nb, s.volumes.data(),
                              s.vel.data(), tp.snb);
        dfsph_p_density_pseudo_pressures(n, dt, rho0, tp.density_kappas.data(),
                                         tp.density_stars.data(), tp.alphas.data());
        apply_pressures_in_place(n, dt, rho0, vel.data(), tp.densities.data(),
                                 tp.density_kappas.data(), tp.fluid_volumes.data(), tp.nb,
                                 s.volumes.data(), tp.snb);
        if (iter > 1) {
            e = transformed_average_and_max(n, T(1000), tp.density_stars.data(),
                                            [rho0](T ds) {
                                                T const err = ds - rho0;
                                                return err <= T(0) ? T(0) : (err / rho0);
                                            });
            if (e.first <= cfg.dfsph_density_avg && e.second <= cfg.dfsph_density_max) {
                break;
            }
        }
    }
    if (st) {
        st->iterations = (iter < cfg.dfsph_max_iterations) ? iter + 1 : iter;
        st->error_average = static_cast<float>(e.first);
        st->error_max = static_cast<float>(e.second);
    }
}

// dfsph_p.cpp:136-212
template <class T>
void dfsph_p_correct_divergence_error(T dt, Config<T> const& cfg, std::vector<V2<T>>& vel,
                                      Solids<T> const& s, Temp<T>& tp, Stats* st) {
    size_t const n = vel.size();
    T const rho0 = cfg.target_density;
    apply_pressures_in_place(n, dt, rho0, vel.data(), tp.densities.data(),
                             tp.divergence_kappas.data(), tp.fluid_volumes.data(), tp.nb,
                             s.volumes.data(), tp.snb);
    int iter = 0;
    std::pair<T, T> e{T(0), T(0)};
    for (; iter < cfg.dfsph_max_iterations; ++iter) {
        compute_divergences(n, rho0, tp.density_stars.data(), tp.fluid_volumes.data(),
                            vel.data(), tp.nb, s.volumes.data(), s.vel.data(), tp.snb);
        dfsph_p_divergence_pseudo_pressures(n, dt, tp.divergence_kappas.data(),
                                            tp.density_stars.data(), tp.alphas.data());
        apply_pressures_in_place(n, dt, rho0, vel.data(), tp.densities.data(),
                                 tp.divergence_kappas.data(), tp.fluid_volumes.data(), tp.nb,
                                 s.volumes.data(), tp.snb);
        if (iter > 0) {
            e = transformed_average_and_max(n, T(1000), tp.density_stars.data(),
                                            [dt, rho0](T d) {
                                                return d <= T(0) ? T(0) : (dt * d / rho0);
                                            });
            if (e.first <= cfg.dfsph_divergence_avg && e.second <= cfg.dfsph_divergence_max) {
                break;
            }
        }
    }
    if (st) {
        st->divergence_iterations = (iter < cfg.dfsph_max_iterations) ? iter + 1 : iter;
        st->divergence_error_average = static_cast<float>(e.first);
        st->divergence_error_max = static_cast<float>(e.second);
    }
}

//------------------------------------------------------------------------------
// IISPH_AP (iisph_ap.cpp, iisph_ap_ops.cpp) and IISPH_PSEUDO_AP
// (iisph_pseudo_ap.cpp, iisph_pseudo_ap_ops.cpp): the pressure-acceleration
// form of IISPH; the pseudo variant solves for displacements with a
// fictitious time step of 1 (no dt^2 in the diagonal or in A p).

// iisph_ap_ops.cpp:19-106 iisph_ap_density_stars_and_diagonals (pseudo =
// false) and iisph_pseudo_ap_ops.cpp:19-110 ..._pseudo_diagonals (pseudo =
// true, whose diagonal has no -dt^2 factor).
template <class T>
void iisph_ap_density_stars_and_diagonals(size_t n, bool pseudo, T dt, T rho0, T* rho_star,
                                          T* diag, T const* rho, T const* fvol,
                                          V2<T> const* fvel, Neighborhood<T> const& nb,
                                          T const* svol, V2<T> const* svel,
                                          Neighborhood<T> const& snb) {
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        int const fc = nb.counts[i];
        int const sc = snb.counts[i];
        if (!(fc + sc)) {
            rho_star[i] = rho[i];
            diag[i] = T(0);
            continue;
        }
        T const self_mass = fvol[i] * rho0;
        T const self_rho = rho[i];
        V2<T> const vi = fvel[i];
        V2<T> sum_m_gradw{T(0), T(0)};
        T sum_m_gradw_dot_gradw = T(0);
        T divergence = T(0);
        for (int k = 0; k < fc; ++k) {
            uint32_t const j = nb.indices[i * MAXN + k];
            V2<T> const gw = nb.kernel_gradients[i * MAXN + k];
            T const mass = fvol[j] * rho0;
            sum_m_gradw += mass * gw;
            sum_m_gradw_dot_gradw += mass * gw.dot(gw);
            divergence += mass * (vi - fvel[j]).dot(gw);
        }
        for (int k = 0; k < sc; ++k) {
            uint32_t const j = snb.indices[i * MAXN + k];
            V2<T> const gw = snb.kernel_gradients[i * MAXN + k];
            T const mass = svol[j] * rho0;
            sum_m_gradw += mass * gw;
            divergence += mass * (vi - svel[j]).dot(gw);
        }
        T const inner = sum_m_gradw.dot(sum_m_gradw) + self_mass * sum_m_gradw_dot_gradw;
        T const numer = pseudo ? -inner : -sqr(dt) * inner;
        T const denom = sqr(self_rho);
        diag[i] = is_safe_divide(numer, denom) ? numer / denom : T(0);
        rho_star[i] = self_rho + dt * divergence;
    }
}

// sph_common/pressure.cpp:83-178 compute_pressure_accelerations; also
// iisph_pseudo_ap_ops.cpp:113-225 (the displacements, the same arithmetic).
// The fluid pass sets, the boundary pass adds.
template <class T>
void compute_pressure_accelerations(size_t n, T rho0, V2<T>* acc, T const* p, T const* rho,
                                    T const* fvol, Neighborhood<T> const& nb, T const* svol,
                                    Neighborhood<T> const& snb) {
    auto por = [&](size_t i) {
        T const d = sqr(rho[i]);
        return is_safe_divide(p[i], d) ? p[i] / d : T(0);
    };
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        int const fc = nb.counts[i];
        if (!fc) {
            acc[i] = V2<T>{T(0), T(0)};
            continue;
        }
        T const pi = por(i);
        V2<T> a{T(0), T(0)};
        for (int k = 0; k < fc; ++k) {
            uint32_t const j = nb.indices[i * MAXN + k];
            T const mass = rho0 * fvol[j];
            V2<T> const gw = nb.kernel_gradients[i * MAXN + k];
            T const coeff = pi + por(j);
            a += (-mass * coeff) * gw;
        }
        acc[i] = a;
    }
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        int const sc = snb.counts[i];
        if (!sc) { continue; }
        T const coeff = por(i);
        V2<T> a{T(0), T(0)};
        for (int k = 0; k < sc; ++k) {
            uint32_t const j = snb.indices[i * MAXN + k];
            T const mass = rho0 * svol[j];
            V2<T> const gw = snb.kernel_gradients[i * MAXN + k];
            a += (-mass * coeff) * gw;
        }
        acc[i] += a;
    }
}

// iisph_ap_ops.cpp:108-194 iisph_ap_iterate_pressures_in_place (pseudo =
// false: A p scaled by dt^2) and iisph_pseudo_ap_ops.cpp:227-305 (pseudo =
// true: a zero diagonal zeroes the pressure). Jacobi in place: each particle
// reads only its own pressure.
template <class T>
std::pair<T, T> iisph_ap_iterate_pressures_in_place(size_t n, bool pseudo, T dt, T rho0,
                                                    T omega, T* p, T const* rho_star,
                                                    V2<T> const* acc, T const* diag,
                                                    T const* fvol, Neighborhood<T> const& nb,
                                                    T const* svol, Neighborhood<T> const& snb) {
    if (!n) { return {T(0), T(0)}; }
    uint64_t isum = 0;
    T emax = T(0);
#pragma omp parallel for schedule(static, 512) reduction(+ : isum) reduction(max : emax)
    for (size_t i = 0; i < n; ++i) {
        int const fc = nb.counts[i];
        int const sc = snb.counts[i];
        if (!fc && !sc) {
            p[i] = T(0);
            continue;
        }
        if (pseudo && diag[i] == T(0)) {
            p[i] = T(0);
            continue;
        }
        T api = T(0);
        V2<T> const ai = acc[i];
        for (int k = 0; k < fc; ++k) {
            uint32_t const j = nb.indices[i * MAXN + k];
            T const mass = rho0 * fvol[j];
            V2<T> const gw = nb.kernel_gradients[i * MAXN + k];
            api += mass * ((ai - acc[j]).dot(gw));
        }
        for (int k = 0; k < sc; ++k) {
            uint32_t const j = snb.indices[i * MAXN + k];
            T const mass = rho0 * svol[j];
            V2<T> const gw = snb.kernel_gradients[i * MAXN + k];
            api += mass * (ai.dot(gw));
        }
        if (!pseudo) { api *= sqr(dt); }
        T const source = rho0 - rho_star[i];
        T const q = api - source;
        T const err = std::max(T(0), is_safe_divide(q, rho0) ? q / rho0 : T(0));
        emax = std::max(emax, err);
        isum += static_cast<uint64_t>(err * T(1000));
        T const numer = omega * (source - api);
        T const step = is_safe_divide(numer, diag[i]) ? numer / diag[i] : T(0);
        p[i] = std::max(T(0), p[i] + step);
    }
    double const num = static_cast<double>(isum);
    double const den = static_cast<double>(1000 * n);
    return {static_cast<T>(num / den), emax};
}

// iisph_ap.cpp:44-138 (pseudo = false) and iisph_pseudo_ap.cpp:63-155
// (pseudo = true; its integration, iisph_pseudo_ap_ops.cpp:307-341, moves by
// the displacement and adds it over dt to the velocity).
template <class T>
void iisph_ap_sub_step(bool pseudo, T t, T dt, Config<T> const& cfg, std::vector<V2<T>>& pos,
                       std::vector<V2<T>>& vel, Solids<T> const& s, Temp<T>& tp, Stats* st) {
    size_t const n = pos.size();
    T const rho0 = cfg.target_density;
    compute_all_neighborhoods(cfg, pos, s, tp);
    tp.external_forces.resize(n);
    external_forces(n, t, dt, cfg, pos.data(), vel.data(), tp.nb, tp.external_forces.data());
    tp.densities.resize(n);
    compute_densities(n, cfg.support, rho0, tp.densities.data(), tp.fluid_volumes.data(),
                      tp.nb, s.volumes.data(), tp.snb);
    integrate_velocities_forces(n, dt, rho0, vel.data(), tp.fluid_volumes.data(),
                                tp.external_forces.data());
    tp.density_stars.resize(n);
    tp.aiis.resize(n);
    iisph_ap_density_stars_and_diagonals(n, pseudo, dt, rho0, tp.density_stars.data(),
                                         tp.aiis.data(), tp.densities.data(),
                                         tp.fluid_volumes.data(), vel.data(), tp.nb,
                                         s.volumes.data(), s.vel.data(), tp.snb);
    tp.pressures.assign(n, T(0));
    tp.pressure_accelerations.resize(n);
    int iter = 0;
    std::pair<T, T> err{T(0), T(0)};
    for (; iter < cfg.iisph_max_iterations; ++iter) {
        compute_pressure_accelerations(n, rho0, tp.pressure_accelerations.data(),
                                       tp.pressures.data(), tp.densities.data(),
                                       tp.fluid_volumes.data(), tp.nb, s.volumes.data(), tp.snb);
        err = iisph_ap_iterate_pressures_in_place(
          n, pseudo, dt, rho0, cfg.iisph_omega, tp.pressures.data(), tp.density_stars.data(),
          tp.pressure_accelerations.data(), tp.aiis.data(), tp.fluid_volumes.data(), tp.nb,
          s.volumes.data(), tp.snb);
        if (iter > 1 && err.first <= cfg.iisph_error_average_threshold &&
            err.second <= cfg.iisph_error_max_threshold) {
            break;
        }
    }
    if (st) {
        st->iterations = (iter < cfg.iisph_max_iterations) ? iter + 1 : iter;
        st->error_average = static_cast<float>(err.first);
        st->error_max = static_cast<float>(err.second);
    }
    compute_pressure_accelerations(n, rho0, tp.pressure_accelerations.data(),
                                   tp.pressures.data(), tp.densities.data(),
                                   tp.fluid_volumes.data(), tp.nb, s.volumes.data(), tp.snb);
    V2<T> const* a = tp.pressure_accelerations.data();
#pragma omp parallel for schedule(static, 512)
    for (size_t i = 0; i < n; ++i) {
        if (pseudo) {
            V2<T> const v = vel[i];
            pos[i] = pos[i] + (dt * v) + a[i];
            if (is_safe_divide(v, dt)) { vel[i] = v + (a[i] / dt); }
        } else {
            // dynamics.cpp:53-62 integrate_velocities_and_positions_in_place
            vel[i] += dt * a[i];
            pos[i] += dt * vel[i];
        }
    }
}

// dfsph_p.cpp:265-374 (active #else branch :325-373)
template <class T>
void dfsph_p_sub_step(T t, T dt, Config<T> const& cfg, std::vector<V2<T>>& pos,
                      std::vector<V2<T>>& vel, Solids<T> const& s, Temp<T>& tp, Stats* st) {
    size_t const n = pos.size();
    T const rho0 = cfg.target_density;
    compute_all_neighborhoods(cfg, pos, s, tp);
    tp.densities.resize(n);
    compute_densities(n, cfg.support, rho0, tp.densities.data(), tp.fluid_volumes.data(),
                      tp.nb, s.volumes.data(), tp.snb);
    tp.alphas.resize(n);
    dfsph_p_compute_shared_coeffs(n, rho0, tp.alphas.data(), tp.densities.data(),
                                  tp.fluid_volumes.data(), tp.nb, s.volumes.data(), tp.snb);
    tp.density_stars.resize(n);
    dfsph_p_correct_divergence_error(dt, cfg, vel, s, tp, st);
    tp.external_forces.resize(n);
    // dfsph_p.cpp:350: the forces at t + dt.
    external_forces(n, t + dt, dt, cfg, pos.data(), vel.data(), tp.nb,
                    tp.external_forces.data());
    integrate_velocities_forces(n, dt, rho0, vel.data(), tp.fluid_volumes.data(),
                                tp.external_forces.data());
    dfsph_p_correct_density_error(dt, cfg, vel, s, tp, st);
    accumulate(n, dt, pos.data(), vel.data());
}

//------------------------------------------------------------------------------
// Scene = one simulation: config, solids, fluid state and scratch.
enum Method { PCISPH = 0, IISPH = 1, IISPH_AP = 2, IISPH_PSEUDO_AP = 3, DFSPH_P = 4 };

template <class T>
struct Scene {
    Config<T> cfg;
    Solids<T> solids;
    std::vector<V2<T>> pos, vel;
    std::vector<T> colors;  // State::colors (state.h:13-17), RGBA per particle
    Temp<T> tmp;
    Stats last{};
};

// colors.cpp:9-20 target_density_colors (the default User_colors_function,
// colors.cpp:42-55): c = clamp(1 - rho / (1.1 rho0), 0, 1), RGBA (c, c, 1, 1),
// from the densities of the last sub-step.
template <class T>
void compute_colors(Scene<T>& sc) {
    size_t const n = sc.pos.size();
    sc.colors.assign(4 * n, T(0));
    if (sc.tmp.densities.size() < n) { return; }
    T const denom = T(1.1f) * sc.cfg.target_density;
    for (size_t i = 0; i < n; ++i) {
        T const norm_d = sc.tmp.densities[i] / denom;
        T const c = std::clamp(T(1) - norm_d, T(0), T(1));
        sc.colors[4 * i] = c;
        sc.colors[4 * i + 1] = c;
        sc.colors[4 * i + 2] = T(1);
        sc.colors[4 * i + 3] = T(1);
    }
}

// colors.cpp:57-70 convert_colors: each channel to uint8 as
// static_cast<uint8_t>(255 * clamp(f, 0, 1)).
inline void convert_colors(size_t n, uint8_t* out, float const* rgba) {
    for (size_t k = 0; k < 4 * n; ++k) {
        out[k] = static_cast<uint8_t>(255.0f * std::clamp(rgba[k], 0.0f, 1.0f));
    }
}

// Per-frame temp resize/init: iisph.cpp:153-166, dfsph_p.cpp:214-229 and the
// constant volumes of adaptive_time_step.h:56-61.
template <class T>
void frame_init(Scene<T>& sc, int method) {
    size_t const n = sc.pos.size();
    sc.tmp.fluid_volumes.resize(n);
    compute_constant_volumes(n, sc.cfg.target_density, sc.cfg.mass_per_particle,
                             sc.tmp.fluid_volumes.data());
    if (method == DFSPH_P) {
        sc.tmp.divergence_kappas.assign(n, T(0));
        sc.tmp.density_kappas.assign(n, T(0));
    }
}

template <class T>
void sub_step(Scene<T>& sc, int method, T t, T dt) {
    Stats st{};
    switch (method) {
    case IISPH: iisph_sub_step(t, dt, sc.cfg, sc.pos, sc.vel, sc.solids, sc.tmp, &st); break;
    case PCISPH: pcisph_sub_step(t, dt, sc.cfg, sc.pos, sc.vel, sc.solids, sc.tmp, &st); break;
    case DFSPH_P: dfsph_p_sub_step(t, dt, sc.cfg, sc.pos, sc.vel, sc.solids, sc.tmp, &st); break;
    case IISPH_AP:
        iisph_ap_sub_step(false, t, dt, sc.cfg, sc.pos, sc.vel, sc.solids, sc.tmp, &st);
        break;
    case IISPH_PSEUDO_AP:
        iisph_ap_sub_step(true, t, dt, sc.cfg, sc.pos, sc.vel, sc.solids, sc.tmp, &st);
        break;
    default: break;
    }
    sc.last = st;
}

// adaptive_time_step.h:35-122 (IISPH 60/6 CFL 0.2, iisph.cpp:168-188;
// DFSPH 60/4 CFL 0.4, dfsph_p.cpp:376-396), and pcisph.cpp:166-204.
template <class T>
int frame_step(Scene<T>& sc, int method, int64_t global_time) {
    size_t const n = sc.pos.size();
    frame_init(sc, method);
    int sub_steps = 0;
    if (method == PCISPH) {
        int64_t const per = sc.cfg.time_per_step / sc.cfg.sub_steps;
        int64_t time = global_time;
        for (int k = 0; k < sc.cfg.sub_steps; ++k) {
            sub_step(sc, method, static_cast<T>(to_seconds(time)), sc.cfg.seconds_per_sub_step);
            time += per;
            ++sub_steps;
        }
        compute_colors(sc);  // pcisph.cpp:186-192
        return sub_steps;
    }
    int const min_den = 60;
    int const max_den = (method == IISPH) ? 6 : 4;
    T const cfl = (method == IISPH) ? T(0.2f) : T(0.4f);
    int64_t const per = sc.cfg.time_per_step / min_den;
    int const divisor = min_den / max_den;
    int remaining = min_den;
    int64_t time = global_time;
    while (remaining > 0) {
        int64_t const cs = cfl_maximum_time_step(n, sc.cfg.support, cfl, sc.vel.data());
        int ns = static_cast<int>(std::ceil(to_seconds(cs) / to_seconds(per)));
        ns = std::clamp(ns, 1, divisor);
        ns = std::min(ns, remaining);
        int64_t const sub = ns * per;
        sub_step(sc, method, static_cast<T>(to_seconds(time)), static_cast<T>(to_seconds(sub)));
        remaining -= ns;
        ++sub_steps;
        time += sub;
    }
    compute_colors(sc);  // adaptive_time_step.h:105-110
    return sub_steps;
}

//------------------------------------------------------------------------------
// Scene inputs. solids.cpp:13-50 emit_solid_box / estimate count.
inline size_t estimate_solid_box_emission_count(float minx, float miny, float maxx,
                                                float maxy, float R) {
    float const sx = maxx - minx;
    float const sy = maxy - miny;
    auto const nx = static_cast<size_t>(std::ceil(sx / (2.0f * R)));
    auto const ny = static_cast<size_t>(std::ceil(sy / (std::sqrt(3.0f) * R)));
    return nx * ny;
}
inline size_t emit_solid_box(float minx, float miny, float maxx, float maxy, float R,
                             size_t max_count, float* out_xy) {
    bool push_x = false;
    float const dx = 2.0f * R;
    float const dy = std::sqrt(3.0f) * R;
    float const sx0 = minx + R;
    float const sy0 = miny + R;
    float const stopx = maxx - R;
    float const stopy = maxy - R;
    float px = sx0;
    float py = sy0;
    size_t count = 0;
    for (; py <= stopy; py += dy) {
        float const offset_x = R;
        px = push_x ? sx0 : sx0 + R;
        px += offset_x;
        for (; px <= stopx; px += dx) {
            out_xy[2 * count] = px;
            out_xy[2 * count + 1] = py;
            ++count;
            if (count == max_count) { return count; }
        }
        push_x = !push_x;
    }
    return count;
}

// config.cpp:16-132 Simulation_config calibration.
inline void simulation_config(float support, float rho0, int64_t time_per_step,
                              int sub_steps, float* mass, float* seconds_per_sub_step,
                              float* draw_radius, float* pcisph_denom) {
    float const H = support;
    float const R = 0.99f * (H * 0.5f);
    *draw_radius = R;
    float const lo = -15.0f * R;
    float const hi = 15.0f * R;
    size_t const emax = estimate_solid_box_emission_count(lo, lo, hi, hi, R);
    std::vector<float> xy(2 * emax);
    size_t const emitted = emit_solid_box(lo, lo, hi, hi, R, emax, xy.data());
    V2<float> best{};
    float best_r2 = std::numeric_limits<float>::max();
    for (size_t i = 0; i < emitted; ++i) {
        V2<float> const p{xy[2 * i], xy[2 * i + 1]};
        float const r2 = p.length2();
        if (r2 < best_r2) {
            best = p;
            best_r2 = r2;
        }
    }
    V2<float> gsum{};
    float gdot = 0.0f;
    float much = 0.0f;
    for (size_t i = 0; i < emitted; ++i) {
        V2<float> const p{xy[2 * i], xy[2 * i + 1]};
        V2<float> const d = best - p;
        float const len = d.length();
        float const w = K<float>::W(len, H);
        if (w > 0.0f) { much += w; }
        V2<float> const g = K<float>::GradW(d, H);
        gsum += g;
        gdot += g.dot(g);
    }
    *mass = rho0 / much;
    int64_t const per = time_per_step / sub_steps;
    *seconds_per_sub_step = static_cast<float>(to_seconds(per));
    float const beta = 2.0f * sqr(*seconds_per_sub_step / much);
    *pcisph_denom = beta * (gsum.dot(gsum) + gdot);
}

// solids.cpp:139-206 world_walls_initial_solid_state, the positions: four
// walls 3H thick, then (random_obstacles) the 16 "tiny_block" boxes whose
// offsets come from std::mt19937_64{1919} through
// std::uniform_real_distribution<float>{-0.06f, 0.06f}. The reference passes
// three draws as the arguments of one call (solids.cpp:172); its clang build
// evaluates them left to right (x offset, y offset, half-size offset), which is
// written out here because g++ would evaluate the call's arguments in another
// order. Boxes are emitted in order into one array of 100 + the estimated
// counts (solids.cpp:183-199).
inline std::vector<float> world_walls_positions(float support, float length,
                                                bool random_obstacles) {
    float const H = support;
    float const R = 0.99f * (H * 0.5f);
    float const L = length;
    float const border = H * 3.0f;
    struct Box {
        float x0, y0, x1, y1;
    };
    std::vector<Box> boxes;
    boxes.push_back({-border, -border, 0.0f, L + border});
    boxes.push_back({L, -border, L + border, L + border});
    boxes.push_back({-0.5f * border, -border, L + 0.5f * border, 0.0f});
    boxes.push_back({-0.5f * border, L, L + 0.5f * border, L + border});
    if (random_obstacles) {
        std::mt19937_64 gen{1919};
        std::uniform_real_distribution<float> dist{-0.06f, 0.06f};
        int const N = 4;
        for (int by = 0; by < N; ++by) {
            float const byf = float(by + 1) / (N + 1);
            for (int bx = 0; bx < N; ++bx) {
                float const bxf = float(bx + 1) / (N + 1);
                V2<float> unit;
                unit.x = bxf + dist(gen);
                unit.y = byf + dist(gen);
                float const roff = dist(gen);
                // solids.cpp:160-163 tiny_block
                float const r = 0.04f + std::abs(roff);
                V2<float> const lo = L * (unit - V2<float>{r, r});
                V2<float> const hi = L * (unit + V2<float>{r, r});
                boxes.push_back({lo.x, lo.y, hi.x, hi.y});
            }
        }
    }
    size_t total = 100;
    for (auto const& b : boxes) {
        total += estimate_solid_box_emission_count(b.x0, b.y0, b.x1, b.y1, R);
    }
    std::vector<float> xy(2 * total);
    size_t remaining = total;
    size_t at = 0;
    for (auto const& b : boxes) {
        size_t const k = emit_solid_box(b.x0, b.y0, b.x1, b.y1, R, remaining, xy.data() + 2 * at);
        remaining -= k;
        at += k;
    }
    xy.resize(2 * at);
    return xy;
}

// initial_state.cpp:14-97 dam_break_initial_state, the positions: a hex
// block [0, L/2] x [0, 3.25 L / 4] emitted with radius R = 0.99 H / 2, then
// every particle with a solid closer than 2R removed. The solids are looked up
// through their own spatial index (cell 2H, the 3x3 cells around the
// particle's cell, each cell's sorted slots in order) as the reference does
// with the solid state's block map (initial_state.cpp:36-72).
inline std::vector<float> dam_break_positions(float support, float length, size_t ns,
                                              float const* solid_xy) {
    float const H = support;
    float const R = 0.99f * (H * 0.5f);
    float const L = length;
    float const x1 = L / 2.0f;
    float const y1 = 3.25f * L / 4.0f;
    size_t const max_count = 100 + estimate_solid_box_emission_count(0.0f, 0.0f, x1, y1, R);
    std::vector<float> xy(2 * max_count);
    size_t const n = emit_solid_box(0.0f, 0.0f, x1, y1, R, max_count, xy.data());
    std::vector<V2<float>> spos(ns);
    for (size_t b = 0; b < ns; ++b) { spos[b] = V2<float>{solid_xy[2 * b], solid_xy[2 * b + 1]}; }
    float const cell = 2.0f * H;
    SpatialIndex<float> si;
    build_spatial_index(ns, spos.data(), cell, si);
    std::vector<uint8_t> kill(n, 0);
#pragma omp parallel for schedule(static, 512)
    for (size_t pi = 0; pi < n; ++pi) {
        V2<float> const pos{xy[2 * pi], xy[2 * pi + 1]};
        int32_t g[2];
        grid_coord(pos, cell, g);
        bool k = false;
        for (int32_t j = g[1] - 1; !k && j <= g[1] + 1; ++j) {
            for (int32_t i = g[0] - 1; !k && i <= g[0] + 1; ++i) {
                auto const it = si.block_map.find(
                  z_index(static_cast<uint32_t>(i), static_cast<uint32_t>(j)));
                if (it == si.block_map.end()) { continue; }
                for (uint32_t sip = it->second.first; sip != it->second.second; ++sip) {
                    uint32_t const oi = si.pairs[sip].second;
                    float const distance = (spos[oi] - pos).length();
                    if (distance < 2 * R) {
                        k = true;
                        break;
                    }
                }
            }
        }
        kill[pi] = k ? 1 : 0;
    }
    std::vector<float> out;
    out.reserve(2 * n);
    for (size_t i = 0; i < n; ++i) {
        if (!kill[i]) {
            out.push_back(xy[2 * i]);
            out.push_back(xy[2 * i + 1]);
        }
    }
    return out;
}

}  // namespace orc

//==============================================================================
// C ABI for ctypes (tests and the CPU baseline only).
using namespace orc;

extern "C" {

uint64_t orc_z_index(uint32_t x, uint32_t y) { return z_index(x, y); }
uint64_t orc_single_axis_index_2(uint32_t i) { return single_axis_index_2(i); }

float orc_length(float x, float y) { return V2<float>{x, y}.length(); }
float orc_anti_coupling_angle(uint64_t i) { return anti_coupling_angle(i); }
uint64_t orc_lehmer_first_draw(uint64_t seed) { Lehmer64 g{seed}; return g(); }
float orc_W(float r, float h) { return K<float>::W(r, h); }
float orc_dW(float r, float h) { return K<float>::dW(r, h); }
void orc_GradW(float x, float y, float h, float* out) {
    V2<float> const g = K<float>::GradW(V2<float>{x, y}, h);
    out[0] = g.x;
    out[1] = g.y;
}
int orc_is_safe_divide_f(float n, float d) { return is_safe_divide(n, d) ? 1 : 0; }
int orc_is_safe_divide_v2f(float nx, float ny, float d) {
    return is_safe_divide(V2<float>{nx, ny}, d) ? 1 : 0;
}
int orc_is_safe_divide_d(double n, double d) { return is_safe_divide(n, d) ? 1 : 0; }

void orc_grid_coords(size_t n, float cell, float const* xy, int32_t* out) {
    for (size_t i = 0; i < n; ++i) {
        grid_coord(V2<float>{xy[2 * i], xy[2 * i + 1]}, cell, &out[2 * i]);
    }
}

// Spatial index of n points: keys, sorted pairs, block indices, blocks.
// Returns the block count.
size_t orc_spatial_index(size_t n, float cell, float const* xy, int32_t* grid_out,
                         uint64_t* keys_out, uint64_t* sorted_keys_out,
                         uint32_t* sorted_idx_out, uint32_t* block_indices_out,
                         uint32_t* block_begin_out, uint32_t* block_end_out) {
    SpatialIndex<float> si;
    build_spatial_index(n, reinterpret_cast<V2<float> const*>(xy), cell, si);
    for (size_t i = 0; i < n; ++i) {
        if (grid_out) {
            grid_out[2 * i] = si.grid[2 * i];
            grid_out[2 * i + 1] = si.grid[2 * i + 1];
        }
        if (keys_out) { keys_out[i] = si.keys[i]; }
        if (sorted_keys_out) { sorted_keys_out[i] = si.pairs[i].first; }
        if (sorted_idx_out) { sorted_idx_out[i] = si.pairs[i].second; }
        if (block_indices_out) { block_indices_out[i] = si.block_indices[i]; }
    }
    for (size_t b = 0; b < si.blocks.size(); ++b) {
        if (block_begin_out) { block_begin_out[b] = si.blocks[b].first; }
        if (block_end_out) { block_end_out[b] = si.blocks[b].second; }
    }
    return si.blocks.size();
}

// Neighbourhoods of n_self points against n_other points (same_array=1 when
// they are the same array). Outputs are [n_self * 64].
void orc_neighborhoods(size_t n_self, float const* self_xy, size_t n_other,
                       float const* other_xy, int same_array, float max_distance,
                       uint8_t* counts, uint32_t* indices, float* distances, float* vectors) {
    SpatialIndex<float> self_si, other_si;
    auto const* sp = reinterpret_cast<V2<float> const*>(self_xy);
    auto const* op = reinterpret_cast<V2<float> const*>(other_xy);
    build_spatial_index(n_other, op, max_distance, other_si);
    std::vector<int32_t> grid(2 * n_self);
    for (size_t i = 0; i < n_self; ++i) { grid_coord(sp[i], max_distance, &grid[2 * i]); }
    Neighborhood<float> nb;
    nb.resize(n_self);
    create_neighborhoods(n_self, max_distance, nb, sp, grid.data(), op, other_si,
                         same_array != 0);
    std::memcpy(counts, nb.counts.data(), n_self);
    std::memcpy(indices, nb.indices.data(), n_self * MAXN * sizeof(uint32_t));
    if (distances) { std::memcpy(distances, nb.distances.data(), n_self * MAXN * sizeof(float)); }
    if (vectors) { std::memcpy(vectors, nb.vectors_to.data(), n_self * MAXN * 2 * sizeof(float)); }
}

// Wall seconds of one neighbour search of n points, the region the
// reference times in sph_common/tests/test_neighborhood.cpp:178-231: grid
// coords, keys, sorted pairs, block indices, blocks and block map, the list
// arrays' allocation and create_regular_neighborhoods. *pairs receives the sum
// of the neighbour counts (a check that the work was done).
double orc_search_seconds(size_t n, float const* xy, float cell, uint64_t* pairs) {
    auto const t0 = std::chrono::steady_clock::now();
    auto const* p = reinterpret_cast<V2<float> const*>(xy);
    SpatialIndex<float> si;
    build_spatial_index(n, p, cell, si);
    Neighborhood<float> nb;
    nb.resize_lists(n);
    create_neighborhoods(n, cell, nb, p, si.grid.data(), p, si, true);
    auto const t1 = std::chrono::steady_clock::now();
    uint64_t total = 0;
    for (size_t i = 0; i < n; ++i) { total += nb.counts[i]; }
    if (pairs) { *pairs = total; }
    return std::chrono::duration<double>(t1 - t0).count();
}

void orc_simulation_config(float support, float rho0, int64_t time_per_step, int sub_steps,
                           float* out4) {
    simulation_config(support, rho0, time_per_step, sub_steps, &out4[0], &out4[1], &out4[2],
                      &out4[3]);
}
size_t orc_estimate_solid_box_emission_count(float minx, float miny, float maxx, float maxy,
                                             float R) {
    return estimate_solid_box_emission_count(minx, miny, maxx, maxy, R);
}
size_t orc_emit_solid_box(float minx, float miny, float maxx, float maxy, float R,
                          size_t max_count, float* out_xy) {
    return emit_solid_box(minx, miny, maxx, maxy, R, max_count, out_xy);
}
// Counts are returned; out_xy (may be NULL) receives up to capacity points.
size_t orc_world_walls_positions(float support, float length, int random_obstacles,
                                 float* out_xy, size_t capacity) {
    std::vector<float> const xy = world_walls_positions(support, length, random_obstacles != 0);
    size_t const n = xy.size() / 2;
    if (out_xy) { std::memcpy(out_xy, xy.data(), 2 * std::min(n, capacity) * sizeof(float)); }
    return n;
}
size_t orc_dam_break_positions(float support, float length, size_t ns, float const* solid_xy,
                               float* out_xy, size_t capacity) {
    std::vector<float> const xy = dam_break_positions(support, length, ns, solid_xy);
    size_t const n = xy.size() / 2;
    if (out_xy) { std::memcpy(out_xy, xy.data(), 2 * std::min(n, capacity) * sizeof(float)); }
    return n;
}
// The float draws of std::uniform_real_distribution<float>{a, b} over
// std::mt19937_64{seed} (the random-obstacle offsets' source).
void orc_mt19937_64_uniform_float(uint64_t seed, float a, float b, size_t n, float* out) {
    std::mt19937_64 gen{seed};
    std::uniform_real_distribution<float> dist{a, b};
    for (size_t i = 0; i < n; ++i) { out[i] = dist(gen); }
}

// Scene configuration block passed from Python (all floats stored as double so
// one struct serves both precisions).
struct OrcConfig {
    double support, gravity, viscosity, target_density;
    int64_t time_per_step;
    int32_t sub_steps;
    int32_t iisph_max_iterations;
    double iisph_error_average_threshold, iisph_error_max_threshold, iisph_omega;
    int32_t dfsph_max_iterations;
    int32_t pad;
    double dfsph_density_avg, dfsph_density_max, dfsph_divergence_avg, dfsph_divergence_max;
    double seconds_per_sub_step, mass_per_particle, pressure_correction_denom;
};

struct OrcScene {
    int fp64;
    Scene<float> f;
    Scene<double> d;
};

extern "C++" template <class T>
void load_cfg(Config<T>& c, OrcConfig const* o) {
    // Thresholds and constants are float in the reference; the FP64 build
    // widens the same float values.
    c.support = static_cast<T>(static_cast<float>(o->support));
    c.gravity = static_cast<T>(static_cast<float>(o->gravity));
    c.viscosity = static_cast<T>(static_cast<float>(o->viscosity));
    c.target_density = static_cast<T>(static_cast<float>(o->target_density));
    c.time_per_step = o->time_per_step;
    c.sub_steps = o->sub_steps;
    c.iisph_max_iterations = o->iisph_max_iterations;
    c.iisph_error_average_threshold = static_cast<T>(static_cast<float>(o->iisph_error_average_threshold));
    c.iisph_error_max_threshold = static_cast<T>(static_cast<float>(o->iisph_error_max_threshold));
    c.iisph_omega = static_cast<T>(static_cast<float>(o->iisph_omega));
    c.dfsph_max_iterations = o->dfsph_max_iterations;
    c.dfsph_density_avg = static_cast<T>(static_cast<float>(o->dfsph_density_avg));
    c.dfsph_density_max = static_cast<T>(static_cast<float>(o->dfsph_density_max));
    c.dfsph_divergence_avg = static_cast<T>(static_cast<float>(o->dfsph_divergence_avg));
    c.dfsph_divergence_max = static_cast<T>(static_cast<float>(o->dfsph_divergence_max));
    c.seconds_per_sub_step = static_cast<T>(static_cast<float>(o->seconds_per_sub_step));
    c.mass_per_particle = static_cast<T>(static_cast<float>(o->mass_per_particle));
    c.pressure_correction_denom = static_cast<T>(static_cast<float>(o->pressure_correction_denom));
}

extern "C++" template <class T>
void load_scene(Scene<T>& sc, OrcConfig const* c, size_t ns, float const* sxy,
                       size_t nf, float const* fxy, float const* fv) {
    load_cfg(sc.cfg, c);
    sc.solids.pos.resize(ns);
    for (size_t i = 0; i < ns; ++i) {
        sc.solids.pos[i] = V2<T>{static_cast<T>(sxy[2 * i]), static_cast<T>(sxy[2 * i + 1])};
    }
    init_solids(sc.cfg, sc.solids);
    sc.pos.resize(nf);
    sc.vel.resize(nf);
    for (size_t i = 0; i < nf; ++i) {
        sc.pos[i] = V2<T>{static_cast<T>(fxy[2 * i]), static_cast<T>(fxy[2 * i + 1])};
        sc.vel[i] = fv ? V2<T>{static_cast<T>(fv[2 * i]), static_cast<T>(fv[2 * i + 1])}
                       : V2<T>{};
    }
}

void* orc_scene_create(OrcConfig const* cfg, size_t n_solid, float const* solid_xy,
                       size_t n_fluid, float const* fluid_xy, float const* fluid_vel,
                       int fp64) {
    auto* s = new OrcScene();
    s->fp64 = fp64;
    if (fp64) {
        load_scene(s->d, cfg, n_solid, solid_xy, n_fluid, fluid_xy, fluid_vel);
    } else {
        load_scene(s->f, cfg, n_solid, solid_xy, n_fluid, fluid_xy, fluid_vel);
    }
    return s;
}
void orc_scene_destroy(void* p) { delete static_cast<OrcScene*>(p); }

// External force model of a scene: flags (1 gravity, 2 pole, 4 drag, 8
// anti-coupling), v = {gravity, pole magnitude, pole x, pole y, drag
// magnitude, drag diameter, anti-coupling max distance, magnitude, draw
// radius}, and an optional user callback (float scenes only).
extern "C++" template <class T>
void set_forces(Config<T>& c, uint32_t flags, double const* v, UserForces fn, void* user) {
    c.force_flags = flags;
    c.gravity = static_cast<T>(static_cast<float>(v[0]));
    c.pole_magnitude = static_cast<T>(static_cast<float>(v[1]));
    c.pole_x = static_cast<T>(static_cast<float>(v[2]));
    c.pole_y = static_cast<T>(static_cast<float>(v[3]));
    c.drag_magnitude = static_cast<T>(static_cast<float>(v[4]));
    c.drag_diameter = static_cast<T>(static_cast<float>(v[5]));
    c.anti_max_distance = static_cast<T>(static_cast<float>(v[6]));
    c.anti_magnitude = static_cast<T>(static_cast<float>(v[7]));
    c.draw_radius = static_cast<T>(static_cast<float>(v[8]));
    c.user_forces = fn;
    c.user_data = user;
}
int orc_scene_set_forces(void* p, uint32_t flags, double const* v, UserForces fn, void* user) {
    auto* s = static_cast<OrcScene*>(p);
    if (s->fp64) {
        if (fn) { return 1; }
        set_forces(s->d.cfg, flags, v, fn, user);
    } else {
        set_forces(s->f.cfg, flags, v, fn, user);
    }
    return 0;
}
void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads() { return omp_get_max_threads(); }

void orc_scene_frame_init(void* p, int method) {
    auto* s = static_cast<OrcScene*>(p);
    if (s->fp64) { frame_init(s->d, method); } else { frame_init(s->f, method); }
}
void orc_scene_sub_step(void* p, int method, double t, double dt, Stats* st) {
    auto* s = static_cast<OrcScene*>(p);
    if (s->fp64) {
        sub_step(s->d, method, t, dt);
        if (st) { *st = s->d.last; }
    } else {
        sub_step(s->f, method, static_cast<float>(t), static_cast<float>(dt));
        if (st) { *st = s->f.last; }
    }
}
int orc_scene_frame_step(void* p, int method, int64_t global_time, Stats* st) {
    auto* s = static_cast<OrcScene*>(p);
    int r;
    if (s->fp64) {
        r = frame_step(s->d, method, global_time);
        if (st) { *st = s->d.last; }
    } else {
        r = frame_step(s->f, method, global_time);
        if (st) { *st = s->f.last; }
    }
    return r;
}

size_t orc_scene_solid_count(void* p) {
    auto* s = static_cast<OrcScene*>(p);
    return s->fp64 ? s->d.solids.pos.size() : s->f.solids.pos.size();
}
void orc_scene_solid_volumes(void* p, double* out) {
    auto* s = static_cast<OrcScene*>(p);
    size_t const n = orc_scene_solid_count(p);
    for (size_t i = 0; i < n; ++i) {
        out[i] = s->fp64 ? s->d.solids.volumes[i] : s->f.solids.volumes[i];
    }
}

// Copy a named per-particle field (original particle order) into out (double).
// Names: pos, vel (2 comps), density, density_star, pressure, aii, dii (2),
// sum_dij_pj (2), alpha, kappa_density, kappa_divergence, pressure_force (2),
// pressure_acceleration (2; IISPH_AP accelerations, IISPH_PSEUDO_AP
// displacements).
extern "C++" template <class T>
int get_field(Scene<T>& sc, char const* name, double* out) {
    size_t const n = sc.pos.size();
    auto v2 = [&](std::vector<V2<T>> const& v) {
        if (v.size() < n) { return -1; }
        for (size_t i = 0; i < n; ++i) {
            out[2 * i] = v[i].x;
            out[2 * i + 1] = v[i].y;
        }
        return 2;
    };
    auto sc1 = [&](std::vector<T> const& v) {
        if (v.size() < n) { return -1; }
        for (size_t i = 0; i < n; ++i) { out[i] = v[i]; }
        return 1;
    };
    std::string const s(name);
    if (s == "pos") { return v2(sc.pos); }
    if (s == "vel") { return v2(sc.vel); }
    if (s == "density") { return sc1(sc.tmp.densities); }
    if (s == "density_star") { return sc1(sc.tmp.density_stars); }
    if (s == "pressure") { return sc1(sc.tmp.pressures); }
    if (s == "aii") { return sc1(sc.tmp.aiis); }
    if (s == "dii") { return v2(sc.tmp.diis); }
    if (s == "sum_dij_pj") { return v2(sc.tmp.sum_dij_pjs); }
    if (s == "alpha") { return sc1(sc.tmp.alphas); }
    if (s == "kappa_density") { return sc1(sc.tmp.density_kappas); }
    if (s == "kappa_divergence") { return sc1(sc.tmp.divergence_kappas); }
    if (s == "pressure_force") { return v2(sc.tmp.pressure_forces); }
    if (s == "pressure_acceleration") { return v2(sc.tmp.pressure_accelerations); }
    if (s == "external_force") { return v2(sc.tmp.external_forces); }
    return -1;
}
int orc_scene_get_field(void* p, char const* name, double* out) {
    auto* s = static_cast<OrcScene*>(p);
    return s->fp64 ? get_field(s->d, name, out) : get_field(s->f, name, out);
}
// Raw float copy of positions/velocities (float precision only).
// State::colors after the last frame step, RGBA floats and converted bytes
// (colors.cpp:57-70); returns 0 when no frame has run.
int orc_scene_get_colors(void* p, float* rgba, uint8_t* rgba8) {
    auto* s = static_cast<OrcScene*>(p);
    size_t const n = s->fp64 ? s->d.pos.size() : s->f.pos.size();
    size_t const m = s->fp64 ? s->d.colors.size() : s->f.colors.size();
    if (m != 4 * n) { return 0; }
    for (size_t k = 0; k < 4 * n; ++k) {
        rgba[k] = s->fp64 ? static_cast<float>(s->d.colors[k]) : s->f.colors[k];
    }
    if (rgba8) { convert_colors(n, rgba8, rgba); }
    return 1;
}
void orc_convert_colors(size_t n, float const* rgba, uint8_t* out) { convert_colors(n, out, rgba); }

void orc_scene_get_state_f64(void* p, double* pos, double* vel) {
    auto* s = static_cast<OrcScene*>(p);
    size_t const n = s->fp64 ? s->d.pos.size() : s->f.pos.size();
    for (size_t i = 0; i < n; ++i) {
        if (s->fp64) {
            pos[2 * i] = s->d.pos[i].x;
            pos[2 * i + 1] = s->d.pos[i].y;
            vel[2 * i] = s->d.vel[i].x;
            vel[2 * i + 1] = s->d.vel[i].y;
        } else {
            pos[2 * i] = s->f.pos[i].x;
            pos[2 * i + 1] = s->f.pos[i].y;
            vel[2 * i] = s->f.vel[i].x;
            vel[2 * i + 1] = s->f.vel[i].y;
        }
    }
}
void orc_scene_get_state_f32(void* p, float* pos, float* vel) {
    auto* s = static_cast<OrcScene*>(p);
    size_t const n = s->fp64 ? s->d.pos.size() : s->f.pos.size();
    for (size_t i = 0; i < n; ++i) {
        if (s->fp64) {
            pos[2 * i] = static_cast<float>(s->d.pos[i].x);
            pos[2 * i + 1] = static_cast<float>(s->d.pos[i].y);
            vel[2 * i] = static_cast<float>(s->d.vel[i].x);
            vel[2 * i + 1] = static_cast<float>(s->d.vel[i].y);
        } else {
            pos[2 * i] = s->f.pos[i].x;
            pos[2 * i + 1] = s->f.pos[i].y;
            vel[2 * i] = s->f.vel[i].x;
            vel[2 * i + 1] = s->f.vel[i].y;
        }
    }
}
// Neighbourhood of the last sub-step (original order), fluid (which=0) or solid (1).
void orc_scene_get_neighborhood(void* p, int which, uint8_t* counts, uint32_t* indices) {
    auto* s = static_cast<OrcScene*>(p);
    if (s->fp64) {
        auto const& nb = which ? s->d.tmp.snb : s->d.tmp.nb;
        std::memcpy(counts, nb.counts.data(), nb.counts.size());
        std::memcpy(indices, nb.indices.data(), nb.indices.size() * 4);
    } else {
        auto const& nb = which ? s->f.tmp.snb : s->f.tmp.nb;
        std::memcpy(counts, nb.counts.data(), nb.counts.size());
        std::memcpy(indices, nb.indices.data(), nb.indices.size() * 4);
    }
}

// The reference neighbourhood test fixture (sph_common/tests/test_neighborhood.cpp:18-31):
// mt19937_64{17272}, uniform_real_distribution<float>{-100, 100}, x then y.
void orc_fixture_points_mt19937_64(uint64_t seed, size_t n, float lo, float hi, float* out_xy) {
    std::mt19937_64 gen{seed};
    std::uniform_real_distribution<float> d{lo, hi};
    for (size_t i = 0; i < n; ++i) {
        out_xy[2 * i] = d(gen);
        out_xy[2 * i + 1] = d(gen);
    }
}

}  // extern "C"
