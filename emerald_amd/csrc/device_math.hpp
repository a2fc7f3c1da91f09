// Device arithmetic for the 2D SPH hot path on gfx950.
//
// Every helper here evaluates the same IEEE-754 single-precision operations, in
// the same order, as the reference CPU code it cites, so that the GPU results
// are bit-identical to the reference's x86-64 SSE build. This relies on the
// build flags in emerald_amd/build.py:
//   -ffp-contract=off   (no a*b+c fusion into FMA)
//   no fast-math; HIP's default correctly-rounded f32 divide and sqrt
//   f32 denormals preserved (no flush-to-zero)
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace emsph {

// Imath Vec2<float> stand-in: component-wise operators, dot = x*x' + y*y'.
struct __align__(8) V2 {
    float x;
    float y;
};

__device__ __forceinline__ V2 v2(float x, float y) { return V2{x, y}; }
__device__ __forceinline__ V2 add(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ V2 sub(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ V2 neg(V2 a) { return V2{-a.x, -a.y}; }
// scalar * vector and vector * scalar are both (s*x, s*y) in IEEE arithmetic.
__device__ __forceinline__ V2 scale(float s, V2 a) { return V2{s * a.x, s * a.y}; }
__device__ __forceinline__ V2 divs(V2 a, float s) { return V2{a.x / s, a.y / s}; }
__device__ __forceinline__ float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }

// Imath Vec2<T>::length with its lengthTiny branch for |v|^2 < 2*FLT_MIN.
__device__ __forceinline__ float imath_length(V2 v) {
    float const l2 = v.x * v.x + v.y * v.y;
    if (l2 < 2.0f * FLT_MIN) {
        float ax = (v.x >= 0.0f) ? v.x : -v.x;
        float ay = (v.y >= 0.0f) ? v.y : -v.y;
        float mx = ax;
        if (mx < ay) { mx = ay; }
        if (mx == 0.0f) { return 0.0f; }
        ax /= mx;
        ay /= mx;
        return mx * __builtin_sqrtf(ax * ax + ay * ay);
    }
    return __builtin_sqrtf(l2);
}

// util/safe_divide.h:14-99 is_safe_divide for float and Vec2<float>.
__device__ __forceinline__ bool is_safe_div(float n, float d) {
    if (d == 0.0f) { return false; }
    float const ad = __builtin_fabsf(d);
    if (ad >= 1.0f) { return true; }
    float const mr = ad / FLT_MIN;  // exact: FLT_MIN is a power of two
    return mr > __builtin_fabsf(n);
}
__device__ __forceinline__ bool is_safe_div(V2 n, float d) {
    if (d == 0.0f) { return false; }
    float const ad = __builtin_fabsf(d);
    if (ad >= 1.0f) { return true; }
    float const mr = ad / FLT_MIN;
    return (mr > __builtin_fabsf(n.x)) && (mr > __builtin_fabsf(n.y));
}

// Correctly rounded single-precision a / b for a divisor b whose reciprocal
// inv = RN64(1 / b) is known. When |a / b| >= 2^-126 (a normal float result,
// or an overflow), RN32(RN64(a * inv)) == RN32(a / b). Proof sketch: in the
// binade [2^E, 2^(E+1)) the float midpoints are M 2^(E-24) with M odd and
// above 2^24, so a / b is never one (M b would need 25 significant bits) and
// a - M 2^(E-24) b is a nonzero multiple of 2^(E-24+e_b): a / b lies more
// than 2^-49 (relative) from every midpoint, while RN64(a * inv) is within
// 2^-52 (relative) of a / b, on the same side of every midpoint. No double
// involved is subnormal. Subnormal and zero results (where a / b can be an
// exact midpoint, e.g. 3 2^-149 / 6) and NaN take the IEEE divide.
__device__ __forceinline__ float div_by_recip(float a, float b, double inv) {
    double const q = static_cast<double>(a) * inv;
    if (!(__builtin_fabs(q) >= 0x1p-126)) { return a / b; }
    return static_cast<float>(q);
}

// Constants of the cubic-spline kernel for one support h, evaluated on the
// host in the same float operation order as sph_common/kernels.h:25-55, with
// the double reciprocals of the divisors (for div_by_recip).
struct KernelConsts {
    float h;       // support (the reference's params.support)
    float k;       // h * h * h * pi
    float hh;      // h * h
    float hhh;     // h * h * h
    float four_h;  // 4 * h
    float w0;      // W(0, h)
    double inv_h, inv_k, inv_hh, inv_hhh, inv_four_h;  // RN64(1 / x)
    // Upper end of the distances r in [2^-48, r_fast_hi] for which every
    // quotient of kernel_W_GradW_r is normal or zero (so div_by_recip's fast
    // path is exact without per-quotient tests): 4h when h lies in
    // [2^-20, 2^10], else -1 (no fast range). See kernel_W_GradW_r.
    float r_fast_hi;
};

// sph_common/kernels.h:16-35 W(r, h). The two pieces' final divides by k are
// one divide of the selected numerator: the same operations per lane, with
// no divergent branches across a wave (lanes at q <= 1 and q > 1 are common).
__device__ __forceinline__ float kernel_W(float r, KernelConsts const& c) {
    float const q = div_by_recip(r, c.h, c.inv_h);  // r / h
    float const wa = 1.0f + q * q * (q * 0.75f - 1.5f);
    float const t = 2.0f - q;
    float const wb = (t * t * t) / 4.0f;
    float const w = div_by_recip(q <= 1.0f ? wa : wb, c.k, c.inv_k);  // / k
    return (q <= 2.0f) ? w : 0.0f;
}

// sph_common/kernels.h:37-55 dW(r, h), branch-free in the same way: the first
// divide of either piece is one divide of selected operands
// (-3 r / h^2 or -3 (2 - r/h)^2 / 4h), and the final divide by k is shared.
__device__ __forceinline__ float kernel_dW(float r, KernelConsts const& c) {
    float const q = div_by_recip(r, c.h, c.inv_h);  // r / h
    bool const inner = q <= 1.0f;
    float const t = 2.0f - q;
    float const num = inner ? -3.0f * r : -3.0f * (t * t);
    float x = inner ? div_by_recip(num, c.hh, c.inv_hh)           // / h^2
                    : div_by_recip(num, c.four_h, c.inv_four_h);  // / 4h
    if (inner) { x = x + div_by_recip(2.25f * r * r, c.hhh, c.inv_hhh); }  // + 2.25 r^2 / h^3
    float const dw = div_by_recip(x, c.k, c.inv_k);  // / k
    return (q <= 2.0f) ? dw : 0.0f;
}

// sph_common/kernels.h:57-64 GradW(dP, h)
__device__ __forceinline__ V2 kernel_GradW(V2 dP, KernelConsts const& c) {
    float const r = imath_length(dP);
    if (is_safe_div(dP, r)) {
        float const dw = kernel_dW(r, c);
        V2 const u = divs(dP, r);
        return V2{u.x * dw, u.y * dw};
    }
    return V2{0.0f, 0.0f};
}

// Same, with r = |dP| already known (r == imath_length(dP) bit for bit).
__device__ __forceinline__ V2 kernel_GradW_r(V2 dP, float r, KernelConsts const& c) {
    if (is_safe_div(dP, r)) {
        float const dw = kernel_dW(r, c);
        V2 const u = divs(dP, r);
        return V2{u.x * dw, u.y * dw};
    }
    return V2{0.0f, 0.0f};
}

// W(r) and GradW(dP) of one neighbour, r = |dP| (r == imath_length(dP) bit
// for bit): the values of kernel_W(r) and kernel_GradW_r(dP, r), bit for bit,
// with q = r / h computed once and the pieces' operands selected instead of
// branched on. For r in [2^-48, r_fast_hi] (h in [2^-20, 2^10]) every
// quotient is normal or exactly zero, so div_by_recip's product is exact
// without its per-quotient test:
//   r / h >= 2^-48 / 2^10;  wa / k >= 2^-2 / (4 * 2^30) (wa >= 1/4 for q <= 1);
//   wb / k >= 2^-71 / 2^32 or 0 (t = 2 - q is exact, a multiple of 2^-23 for
//   q in [1, 4], so t^3 / 4 >= 2^-71 unless t = 0);
//   3 r / h^2 >= 3 * 2^-48 / 2^20;  3 t^2 / 4h >= 2^-56 or 0;
//   2.25 r^2 / h^3 >= 2.25 * 2^-96 / 2^30 > 2^-126;
//   x / k with |x| >= 0.74 r / h^2 for q <= 1 (-3 + 2.25 q <= -0.75), and
//   x = -3 t^2 / 4h otherwise: >= 2^-100 or 0;
// and r <= 4h keeps every quotient below 2^84. Other distances take
// kernel_W and kernel_GradW_r, whose divides are tested one by one.
__device__ __forceinline__ void kernel_W_GradW_r(V2 dP, float r, KernelConsts const& c,
                                                 float& w, V2& gw) {
    if (!(r >= 0x1p-48f && r <= c.r_fast_hi)) {
        w = kernel_W(r, c);
        gw = kernel_GradW_r(dP, r, c);
        return;
    }
    float const q = static_cast<float>(static_cast<double>(r) * c.inv_h);  // r / h
    bool const inner = q <= 1.0f;
    float const t = 2.0f - q;
    float const wa = 1.0f + q * q * (q * 0.75f - 1.5f);
    float const wb = (t * t * t) / 4.0f;
    float const wk = static_cast<float>(static_cast<double>(inner ? wa : wb) * c.inv_k);  // / k
    float const num = inner ? -3.0f * r : -3.0f * (t * t);
    float x = static_cast<float>(static_cast<double>(num) *
                                 (inner ? c.inv_hh : c.inv_four_h));  // / h^2 or / 4h
    float const x2 = static_cast<float>(static_cast<double>(2.25f * r * r) * c.inv_hhh);  // 2.25 r^2 / h^3
    x = inner ? x + x2 : x;
    float const dwk = static_cast<float>(static_cast<double>(x) * c.inv_k);  // / k
    bool const in_support = q <= 2.0f;
    w = in_support ? wk : 0.0f;
    float const dw = in_support ? dwk : 0.0f;
    if (is_safe_div(dP, r)) {
        V2 const u = divs(dP, r);
        gw = V2{u.x * dw, u.y * dw};
    } else {
        gw = V2{0.0f, 0.0f};
    }
}

// sph_common/common.cpp:43-67 grid coordinate: floor(p / cell) per axis with
// the bounds minimum at the origin.
__device__ __forceinline__ int2 grid_coord(V2 p, float cell) {
    float const qx = (p.x - 0.0f) / cell;
    float const qy = (p.y - 0.0f) / cell;
    return make_int2(static_cast<int32_t>(__builtin_floorf(qx)),
                     static_cast<int32_t>(__builtin_floorf(qy)));
}

// z_index/z_index.cpp:17-48: full 64-bit interleave of two 32-bit axes (x on
// even bits, y on odd bits), written as bit spreading instead of a LUT.
__device__ __host__ __forceinline__ uint64_t spread_bits_2(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}
// Inverse of spread_bits_2: the even bits of x packed into 32 bits.
__device__ __host__ __forceinline__ uint32_t compact_bits_2(uint64_t x) {
    x &= 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
    return static_cast<uint32_t>(x);
}

__device__ __host__ __forceinline__ uint64_t morton2(int32_t gx, int32_t gy) {
    return spread_bits_2(static_cast<uint32_t>(gx)) |
           (spread_bits_2(static_cast<uint32_t>(gy)) << 1);
}

// Hash of a (cell key, scene) pair for the static solid cell set (64-bit
// finalizer; the same on host and device).
__device__ __host__ __forceinline__ uint32_t cell_set_hash(uint64_t key, uint32_t tag) {
    uint64_t h = key ^ (static_cast<uint64_t>(tag) * 0x9E3779B97F4A7C15ull);
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return static_cast<uint32_t>(h);
}

// Workgroup -> tile remap so that each XCD (workgroups are dealt round-robin
// over the 8 XCDs) walks one contiguous run of Morton-sorted tiles and keeps
// its neighbour gathers in its own L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t XCDS = 8;
    uint32_t const xcd = bid % XCDS;
    uint32_t const idx = bid / XCDS;
    // Blocks with bid >= nblocks - (nblocks % XCDS) would map past the end when
    // nblocks is not a multiple of 8; fall back to identity for the tail.
    uint32_t const full = nblocks - (nblocks % XCDS);
    if (bid >= full) { return bid; }
    return xcd * (full / XCDS) + idx;
}

}  // namespace emsph
