// Device helpers shared by the build and query kernels (gfx950).
//
// The "reference arithmetic" functions below evaluate the residual metric in
// the reference's f32 order (src/metric.rs:376-563) so that the device tree
// and the reference-form query reproduce the reference's comparisons bit for
// bit.  Every file including this header compiles with contraction off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "exp_cr.hpp"

#pragma clang fp contract(off)

namespace cx {

constexpr int32_t kLevelUnset = INT32_MIN;

// Read-only metric data in tree order.
struct MetricView {
    const float *__restrict__ scaled; // n x d (coords * inv_ls)
    const float *__restrict__ norms;  // n
    const float *__restrict__ v;      // n x m
    const float *__restrict__ p;      // n
    int32_t d, m;
    float var;
    int32_t kernel; // 0 RBF, 1 Matern 5/2
};

// wide::f32x8::reduce_add on the SSE2 baseline: two in-order f32x4 sums.
__device__ __forceinline__ float reduce8(const float *x)
{
    float lo = ((x[0] + x[1]) + x[2]) + x[3];
    float hi = ((x[4] + x[5]) + x[6]) + x[7];
    return lo + hi;
}

// dot_product_simd_slice (src/metric.rs:693-733).
__device__ __forceinline__ float dot_slice(const float *__restrict__ a,
                                           const float *__restrict__ b, int32_t len)
{
    float acc = 0.0f;
    const int32_t tail = (len / 8) * 8;
    for (int32_t i = 0; i < tail; i += 8) {
        float pr[8];
#pragma unroll
        for (int l = 0; l < 8; l++)
            pr[l] = a[i + l] * b[i + l];
        acc += reduce8(pr);
    }
    for (int32_t j = tail; j < len; j++)
        acc += a[j] * b[j];
    return acc;
}

// dot_slice with 16-byte loads, for rows that are 16-byte aligned with
// len % 4 == 0: the same products and additions in the same order.
__device__ __forceinline__ float dot_slice_v4(const float *__restrict__ a,
                                              const float *__restrict__ b, int32_t len)
{
    float acc = 0.0f;
    const int32_t tail = (len / 8) * 8;
    for (int32_t i = 0; i < tail; i += 8) {
        const float4 a0 = *(const float4 *)(a + i), a1 = *(const float4 *)(a + i + 4);
        const float4 b0 = *(const float4 *)(b + i), b1 = *(const float4 *)(b + i + 4);
        float pr[8] = {a0.x * b0.x, a0.y * b0.y, a0.z * b0.z, a0.w * b0.w,
                       a1.x * b1.x, a1.y * b1.y, a1.z * b1.z, a1.w * b1.w};
        acc += reduce8(pr);
    }
    for (int32_t j = tail; j < len; j++)
        acc += a[j] * b[j];
    return acc;
}

// dot_f32_simd_inline (src/metric.rs:601-643).
__device__ __forceinline__ float dot_inline(const float *__restrict__ a,
                                            const float *__restrict__ b, int32_t len)
{
    const int32_t chunks = len / 8, tail = chunks * 8;
    float acc0[8], acc1[8];
#pragma unroll
    for (int l = 0; l < 8; l++) {
        acc0[l] = 0.0f;
        acc1[l] = 0.0f;
    }
    int32_t i = 0;
    for (int32_t dc = 0; dc < chunks / 2; dc++) {
#pragma unroll
        for (int l = 0; l < 8; l++) {
            float p0 = a[i + l] * b[i + l];
            float p1 = a[i + 8 + l] * b[i + 8 + l];
            acc0[l] = p0 + acc0[l];
            acc1[l] = p1 + acc1[l];
        }
        i += 16;
    }
    if (chunks & 1) {
#pragma unroll
        for (int l = 0; l < 8; l++) {
            float p0 = a[i + l] * b[i + l];
            acc0[l] = p0 + acc0[l];
        }
    }
    float comb[8];
#pragma unroll
    for (int l = 0; l < 8; l++)
        comb[l] = acc0[l] + acc1[l];
    float result = reduce8(comb);
    for (int32_t j = tail; j < len; j++)
        result += a[j] * b[j];
    return result;
}

__device__ __forceinline__ float clamp_rho(float rho)
{
    if (rho != rho) // f32::max returns the non-NaN operand
        return -1.0f;
    rho = rho < -1.0f ? -1.0f : rho;
    return rho > 1.0f ? 1.0f : rho;
}

// dist_sq_ref (below) with the V-row dot already evaluated (the same
// dot_slice value)
__device__ __forceinline__ float dist_sq_ref_dot(const MetricView &mv, int64_t i, int64_t j,
                                                 float dot)
{
    if (i == j)
        return 0.0f;
    const int32_t d = mv.d;
    float dot_scaled = dot_slice(mv.scaled + i * d, mv.scaled + j * d, d);
    float d2 = mv.norms[i] + mv.norms[j] - 2.0f * dot_scaled;
    if (d2 < 0.0f)
        d2 = 0.0f;
    float kval;
    if (mv.kernel == 0) {
        kval = mv.var * expf_cr(-0.5f * d2);
    } else {
        float r = __builtin_sqrtf(d2);
        float a = 2.2360679775f * r;
        float poly = 1.0f + a + (a * a) / 3.0f;
        kval = mv.var * poly * expf_cr(-a);
    }
    float den2 = mv.p[i] * mv.p[j];
    const float eps = 1e-9f;
    const float floor_sq = eps * eps;
    if (!(den2 >= floor_sq))
        den2 = floor_sq;
    float denom = __builtin_sqrtf(den2);
    float rho = denom > 0.0f ? (kval - dot) / denom : 0.0f;
    rho = clamp_rho(rho);
    float ds = 1.0f - __builtin_fabsf(rho);
    return ds > 0.0f ? ds : 0.0f;
}

// distance_sq_idx_f32 (src/metric.rs:489-563): 1 - |rho|, no sqrt.  Used by
// the build, the si_cache pass, root seeding and predecessor-failed children.
__device__ __forceinline__ float dist_sq_ref(const MetricView &mv, int64_t i, int64_t j)
{
    if (i == j)
        return 0.0f;
    const int32_t m = mv.m;
    // V rows start 16-byte aligned when m % 4 == 0 (allocation base + i*m floats)
    const float dot = (m & 3) == 0 ? dot_slice_v4(mv.v + i * m, mv.v + j * m, m)
                                   : dot_slice(mv.v + i * m, mv.v + j * m, m);
    return dist_sq_ref_dot(mv, i, j, dot);
}

// distances_f32_batch_idx_into_with_kth (src/metric.rs:376-487): the
// residual distance sqrt(1 - |rho|) used for query tiles.
__device__ __forceinline__ float dist_batch_ref(const MetricView &mv, int64_t q, int64_t j)
{
    if (q == j)
        return 0.0f;
    const int32_t d = mv.d, m = mv.m;
    const float *qc = mv.scaled + q * d, *pc = mv.scaled + j * d;
    float dot_c = 0.0f;
    if (d == 3) {
        dot_c = qc[0] * pc[0] + qc[1] * pc[1] + qc[2] * pc[2];
    } else {
        for (int32_t c = 0; c < d; c++)
            dot_c += qc[c] * pc[c];
    }
    float d2 = mv.norms[q] + mv.norms[j] - 2.0f * dot_c;
    if (d2 < 0.0f)
        d2 = 0.0f;
    float kval;
    if (mv.kernel == 0) {
        kval = mv.var * expf_cr(-0.5f * d2);
    } else {
        float r = __builtin_sqrtf(d2);
        float a = 2.2360679775f * r;
        float poly = 1.0f + a + (a * a) * 0.3333333333f;
        kval = mv.var * poly * expf_cr(-a);
    }
    float prod = mv.p[q] * mv.p[j];
    const float eps = 1e-9f;
    const float floor_sq = eps * eps;
    if (!(prod >= floor_sq))
        prod = floor_sq;
    float denom = __builtin_sqrtf(prod);
    float dot_v = dot_inline(mv.v + q * m, mv.v + j * m, m);
    float rho = clamp_rho((kval - dot_v) / denom);
    float ds = 1.0f - __builtin_fabsf(rho);
    return ds > 0.0f ? __builtin_sqrtf(ds) : 0.0f;
}

// Relaxed workgroup-scope loads/stores for tree link arrays that a kernel
// both reads and writes (keeps them out of the scalar cache; ordering comes
// from __syncthreads()).
template <typename T> __device__ __forceinline__ T ld_rx(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T> __device__ __forceinline__ void st_rx(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

} // namespace cx
