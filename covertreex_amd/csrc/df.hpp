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

__device__ __forceinline__ df df_div(df a, df b)
{
    float q1 = a.hi / b.hi;
    df r = df_sub(a, df_mul_f(b, q1));
    float q2 = r.hi / b.hi;
    df r2 = df_sub(r, df_mul_f(b, q2));
    float q3 = r2.hi / b.hi;
    df q = fast_two_sum(q1, q2);
    return df_add(q, df{q3, 0.0f});
}

__device__ __forceinline__ df df_sqrt(df a)
{
    if (!(a.hi > 0.0f))
        return {0.0f, 0.0f};
    float y = __builtin_sqrtf(a.hi);
    df r = df_sub(a, two_prod(y, y));
    float c = r.hi / (2.0f * y);
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

// exp(x) for x <= ~0 in double-float: x = n ln2 + r with ln2 as three floats
// (the first with 15 significant bits so n*L1 is exact for |n| < 256), then
// the degree-14 Taylor polynomial of exp(r), |r| <= 0.35, in Horner form
// (truncation < 1e-17 relative), scaled by 2^n.
__device__ __forceinline__ df df_exp(df x)
{
    if (x.hi < -104.0f)
        return {0.0f, 0.0f};
    if (x.hi > 88.0f)
        return {__builtin_huge_valf(), 0.0f};
    constexpr float L1 = 0.693145751953125f;          // 0x3f317200
    constexpr double ln2 = 0.69314718055994530942;
    constexpr float L2 = (float)(ln2 - (double)L1);
    constexpr float L3 = (float)(ln2 - (double)L1 - (double)L2);
    const float n = __builtin_rintf(x.hi * 1.44269504f);
    df r = {x.hi - n * L1, 0.0f};                      // exact (Sterbenz)
    r = df_add(r, df{x.lo, 0.0f});
    r = df_sub(r, two_prod(n, L2));
    r = df_sub(r, df{n * L3, 0.0f});
    constexpr DfConst c[15] = {
        split_const(fact_inv(0)),  split_const(fact_inv(1)),  split_const(fact_inv(2)),
        split_const(fact_inv(3)),  split_const(fact_inv(4)),  split_const(fact_inv(5)),
        split_const(fact_inv(6)),  split_const(fact_inv(7)),  split_const(fact_inv(8)),
        split_const(fact_inv(9)),  split_const(fact_inv(10)), split_const(fact_inv(11)),
        split_const(fact_inv(12)), split_const(fact_inv(13)), split_const(fact_inv(14))};
    df p = {c[14].hi, c[14].lo};
#pragma unroll
    for (int i = 13; i >= 0; i--)
        p = df_add(df_mul(p, r), df{c[i].hi, c[i].lo});
    const int e = (int)n;
    return {__builtin_ldexpf(p.hi, e), __builtin_ldexpf(p.lo, e)};
}

} // namespace cx
