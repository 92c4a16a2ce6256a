// Core of the correctly rounded binary32 exp (expf_cr in exp_cr.hpp):
// argument reduction, table-driven evaluation as an unevaluated sum of three
// floats, and the exact rounding decision.  Shared by the device code and the
// generator of its tables (gen_exp_cr.cpp, built for the CPU).
#pragma once

#include <stdint.h>

namespace cx {

struct ExpCrTab {
    float h[256], m[256], l[256]; // 2^(j/256) = h + m + l
};

namespace ecr {

__device__ __forceinline__ uint32_t bits(float f) { return __builtin_bit_cast(uint32_t, f); }
__device__ __forceinline__ float from_bits(uint32_t b) { return __builtin_bit_cast(float, b); }

__device__ __forceinline__ float two_sum(float a, float b, float *e)
{
    const float s = a + b;
    const float bb = s - a;
    *e = (a - (s - bb)) + (b - bb);
    return s;
}

__device__ __forceinline__ float two_prod(float a, float b, float *e)
{
    const float p = a * b;
    *e = __builtin_fmaf(a, b, -p);
    return p;
}

// (ah + al) + (bh + bl) as a double-float, relative error within 3u^2 of the
// result (accurate double-word addition, Joldes, Muller and Popescu 2017)
__device__ __forceinline__ float dw_add(float ah, float al, float bh, float bl, float *lo)
{
    float e, f;
    const float s = two_sum(ah, bh, &e);
    const float t = two_sum(al, bl, &f);
    float c = e + t;
    float v, w;
    const float u = s + c;
    v = c - (u - s);
    w = f + v;
    const float h = u + w;
    *lo = w - (h - u);
    return h;
}

// ln2/256 = kC1 + kC2 + kC3 + kC4 (kC1 with 8 significant bits)
constexpr float kInvC = 0x1.715476p+8f; // 256/ln2
constexpr float kC1 = 0x1.62p-9f;
constexpr float kC2 = 0x1.c85fe0p-18f;
constexpr float kC3 = -0x1.718432p-43f;
constexpr float kC4 = -0x1.4361c4p-68f;

// exp(x) ~ (s + t + c) 2^k for x in [-110, 89]: s = RN(y), t its exact
// TwoSum remainder, c the small remainder
struct Core {
    float s, t, c;
    int32_t k;
};

__device__ __forceinline__ Core core(float x, const ExpCrTab &tab)
{
    const float nf = __builtin_rintf(x * kInvC);
    const int32_t n = (int32_t)nf;
    // r = x - nf ln2/256 as a double-float
    float ul, pl, rl;
    const float uh = two_sum(x, -(nf * kC1), &ul); // nf kC1 exact
    float ph = two_prod(nf, kC2, &pl);
    float rh = dw_add(uh, ul, -ph, -pl, &rl);
    ph = two_prod(nf, kC3, &pl);
    rh = dw_add(rh, rl, -ph, -pl, &rl);
    rh = dw_add(rh, rl, -(nf * kC4), 0.0f, &rl);

    const int32_t j = n & 255;
    const float th = tab.h[j], tm = tab.m[j], tl = tab.l[j];
    float b1, qe, b2;
    const float a1 = two_prod(th, rh, &b1); // Th rh
    const float q = two_prod(rh, rh, &qe);  // rh^2
    const float a2 = two_prod(th, 0.5f * q, &b2);
    // r^3 (1/6 + r/24 + r^2/120) from rh
    const float p3 =
        (rh * q) * __builtin_fmaf(rh, __builtin_fmaf(rh, 1.0f / 120.0f, 1.0f / 24.0f),
                                  0.16666667f);
    // the rest of Th (r + r^2/2) beyond a1 + a2: Th (rl + qe/2 + rh rl), the
    // product remainders; plus Tm (r + r^2/2), Th r^3 (...), Tl
    const float small =
        __builtin_fmaf(th, rl + __builtin_fmaf(rh, rl, 0.5f * qe) + p3,
                       tm * __builtin_fmaf(0.5f, q, rh)) +
        (b1 + b2) + tl;
    float e1, e2;
    const float c1 = two_sum(a1, a2, &e1);
    const float c2 = two_sum(c1, tm, &e2);
    const float clo = (e1 + e2) + small;
    Core r;
    r.s = two_sum(th, c2, &r.t);
    r.c = clo;
    r.k = n >> 8;
    return r;
}

// Rounds (s + t + c) 2^k to binary32; *gate: the value lies within
// gate_ulps result ulps of a rounding midpoint
__device__ __forceinline__ float round_core(const Core &y, float gate_ulps, bool *gate)
{
    const uint32_t sb = bits(y.s);
    const int32_t es = (int32_t)(sb >> 23) - 127; // s in [0.99, 2.02]
    const int32_t e = es + y.k;
    *gate = false;
    if (e > 127)
        return from_bits(0x7f800000u);
    if (e >= -126) {
        // the neighbour of s on t's side and the distance of s + t + c from
        // the midpoint between them (t - half gap is exact where it matters)
        const bool up = y.t >= 0.0f;
        const float nb = from_bits(up ? sb + 1u : sb - 1u);
        float gap = up ? nb - y.s : y.s - nb;
        // at 2^-126 the grid below s is the subnormal one, as wide as above
        if (!up && e == -126 && (sb & 0x7fffffu) == 0u)
            gap = from_bits(sb + 1u) - y.s;
        const float half = 0.5f * gap;
        const float v = up ? (y.t - half) + y.c : (y.t + half) + y.c;
        // v > 0 (up) / v < 0 (down): beyond the midpoint
        const bool cross = up ? v > 0.0f : v < 0.0f;
        const bool tie = v == 0.0f;
        *gate = __builtin_fabsf(v) < gate_ulps * gap;
        uint32_t rb = sb;
        if (cross || (tie && (sb & 1u)))
            rb = up ? sb + 1u : sb - 1u;
        if (!up && e == -126 && (sb & 0x7fffffu) == 0u)
            rb = sb; // cannot cross: |t| is at most a quarter of the gap
        // scale by 2^k through the exponent field (stays normal, or carries
        // into the infinity encoding)
        const int32_t er = (int32_t)(rb >> 23) - 127 + y.k;
        if (er > 127)
            return from_bits(0x7f800000u);
        if (er < -126)
            return 0.0f; // not reached: rounding down stops at 2^-126
        return from_bits(rb + ((uint32_t)y.k << 23));
    }
    // subnormal result: significand M of s (24 bits) on the grid 2^sh
    const int32_t sh = -126 - e;
    if (sh > 24)
        return 0.0f;
    const uint32_t m = (sb & 0x7fffffu) | 0x800000u;
    const uint32_t q = m >> sh, rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1);
    const float v = y.t + y.c; // its sign is exact
    bool up = rem > half;
    if (rem == half) {
        up = v > 0.0f || (v == 0.0f && (q & 1u));
        const float ulp_s = from_bits((uint32_t)(es + 104) << 23); // 2^(es - 23)
        *gate = __builtin_fabsf(v) < gate_ulps * ulp_s * (float)(1u << sh);
    }
    return from_bits(q + (up ? 1u : 0u));
}

} // namespace ecr

} // namespace cx
