// Compensated FP32 ("double-float") arithmetic for the exact rescore of the
// mixed-precision engine: an unevaluated sum hi + lo of two floats carries
// about 48 significand bits.  Error-free transforms: TwoSum (Knuth),
// FastTwoSum (Dekker), TwoProduct through fmaf.
#pragma once

#include <hip/hip_runtime.h>

namespace cx {

struct df {
    float hi, lo;
};

__device__ __forceinline__ df two_sum(float a, float b)
{
    float s = a + b;
    float bb = s - a;
    float e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__device__ __forceinline__ df fast_two_sum(float a, float b)
{
    float s = a + b;
    float e = b - (s - a);
    return {s, e};
}

__device__ __forceinline__ df two_prod(float a, float b)
{
    float p = a * b;
    float e = fmaf(a, b, -p);
    return {p, e};
}

__device__ __forceinline__ df df_neg(df a) { return {-a.hi, -a.lo}; }

// accurate double-float addition (both lows folded with TwoSum)
__device__ __forceinline__ df df_add(df a, df b)
{
    df s = two_sum(a.hi, b.hi);
    df t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ df df_sub(df a, df b) { return df_add(a, df_neg(b)); }

// addition for accumulating a long sum: error ~ 2^-47 of the running magnitude
__device__ __forceinline__ df df_acc(df s, df p)
{
    df t = two_sum(s.hi, p.hi);
    t.lo += s.lo + p.lo;
    return fast_two_sum(t.hi, t.lo);
}

__device__ __forceinline__ df df_mul(df a, df b)
{
    df p = two_prod(a.hi, b.hi);
    p.lo = fmaf(a.hi, b.lo, fmaf(a.lo, b.hi, p.lo));
    return fast_two_sum(p.hi, p.lo);
}

__device__ __forceinline__ df df_mul_f(df a, float b)
{
    df p = two_prod(a.hi, b);
    p.lo = fmaf(a.lo, b, p.lo);
    return fast_two_sum(p.hi, p.lo);
}

// a / b: three quotient terms, each the remainder (exact up to the
// double-float subtraction) times the hardware reciprocal of b.hi (about 1 ulp,
// v_rcp_f32); the approximation only enters through the third term's size,
// so the result stays within 2^-44 relative.
__device__ __forceinline__ df df_div(df a, df b)
{
    const float rb = __builtin_amdgcn_rcpf(b.hi);
    float q1 = a.hi * rb;
    df r = df_sub(a, df_mul_f(b, q1));
    float q2 = r.hi * rb;
    df r2 = df_sub(r, df_mul_f(b, q2));
    float q3 = r2.hi * rb;
    df q = fast_two_sum(q1, q2);
    return df_add(q, df{q3, 0.0f});
}

// sqrt(a): the hardware root of a.hi (about 1 ulp, v_sqrt_f32) corrected by
// one Newton step on the double-float remainder, within 2^-43 relative.
__device__ __forceinline__ df df_sqrt(df a)
{
    if (!(a.hi > 0.0f))
        return {0.0f, 0.0f};
    const float y = __builtin_amdgcn_sqrtf(a.hi);
    const df r = df_sub(a, two_prod(y, y));
    const float c = r.hi * (0.5f * __builtin_amdgcn_rcpf(y));
    return fast_two_sum(y, c);
}

__device__ __forceinline__ bool df_lt(df a, df b)
{
    return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}

// 1/i! split into (hi, lo) at compile time
struct DfConst {
    float hi, lo;
};

constexpr double fact_inv(int i)
{
    double f = 1.0;
    for (int j = 2; j <= i; j++)
        f *= (double)j;
    return 1.0 / f;
}

constexpr DfConst split_const(double v)
{
    return {(float)v, (float)(v - (double)(float)v)};
}

// 2^(k/64), k < 64, as double-float pairs: the double value from its Taylor
// series at compile time (exp(k ln2 / 64), 30 terms), split into two floats.
struct Exp2Tab {
    DfConst v[64];
};

constexpr Exp2Tab make_exp2_tab()
{
    Exp2Tab t{};
    constexpr double ln2 = 0.69314718055994530942;
    for (int k = 0; k < 64; k++) {
        const double x = k * ln2 / 64.0;
        double term = 1.0, sum = 1.0;
        for (int i = 1; i < 30; i++) {
            term *= x / i;
            sum += term;
        }
        t.v[k] = split_const(sum);
    }
    return t;
}

__constant__ constexpr Exp2Tab kExp2Tab = make_exp2_tab();

// exp(x) for x <= ~0 in double-float: x = (64 h + k) ln2/64 + r with N =
// 64 h + k the integer nearest x 64/ln2 and ln2/64 as three floats (the first
// with 10 significant bits, so N L1 is exact for |N| < 2^14, i.e. x > -104),
// |r| <= ln2/128; exp(r) = 1 + r + r^2/2 + r^2 g(r) with the double-float
// terms exact up to their additions and g(r) = r/6 + r^2/24 + r^3/120 in FP32
// (r^2 g(r) < 2.6e-8, so its FP32 rounding and the dropped r^6/720 < 4e-17
// stay below 2^-48); times 2^(k/64) from the table and scaled by 2^h.  Within
// 2^-43 relative.
__device__ __forceinline__ df df_exp(df x)
{
    if (x.hi < -104.0f)
        return {0.0f, 0.0f};
    if (x.hi > 88.0f)
        return {__builtin_huge_valf(), 0.0f};
    constexpr double l64 = 0.69314718055994530942 / 64.0;
    constexpr float L1 = 709.0f / 65536.0f;            // l64 cut to 10 significant bits
    constexpr float L2 = (float)(l64 - (double)L1);
    constexpr float L3 = (float)(l64 - (double)L1 - (double)L2);
    const float nf = __builtin_rintf(x.hi * 92.332482616893656f); // 64 / ln2
    df r = {x.hi - nf * L1, 0.0f};                       // exact (Sterbenz)
    r = df_add(r, df{x.lo, 0.0f});
    r = df_sub(r, two_prod(nf, L2));
    r = df_sub(r, df{nf * L3, 0.0f});
    const float t = r.hi;
    const float g = t * __builtin_fmaf(t, __builtin_fmaf(t, 1.0f / 120.0f, 1.0f / 24.0f),
                                       1.0f / 6.0f);
    const df r2 = df_mul(r, r);
    df e = df_add(df{1.0f, 0.0f}, r);
    e = df_add(e, df{0.5f * r2.hi, 0.5f * r2.lo});
    e = df_add(e, df{r2.hi * g, 0.0f});
    const int n = (int)nf;
    const DfConst c = kExp2Tab.v[n & 63];
    e = df_mul(e, df{c.hi, c.lo});
    const int h = n >> 6;
    return {__builtin_ldexpf(e.hi, h), __builtin_ldexpf(e.lo, h)};
}

} // namespace cx
