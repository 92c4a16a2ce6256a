/* Residual-correlation and Euclidean metrics -- oracle restatement.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Build with -ffp-contract=off. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Double-evaluated expf: Cody-Waite reduction x = k ln2 + r (fdlibm's ln2
 * hi/lo split, |k| < 2^21 keeps k*ln2_hi exact), exp(r) by the degree-13
 * Taylor polynomial in Horner form (truncation < 5e-18 for |r| <= ln2/2),
 * scale by 2^k, round once to float.  Every operation is an IEEE double op in
 * a fixed order, so any IEEE machine without contraction gets the same bits. */
static float expf_dd(float x)
{
    if (isnan(x))
        return x;
    if (x > 89.0f)
        return INFINITY;
    if (x < -110.0f)
        return 0.0f;
    const double inv_ln2 = 1.4426950408889634074;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double xd = (double)x;
    double k = nearbyint(xd * inv_ln2);
    double r = (xd - k * ln2_hi) - k * ln2_lo;
    double p = 1.0 / 6227020800.0; /* 1/13! */
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return (float)ldexp(p, (int)k);
}

float ox_expf(int32_t mode, float x)
{
    if (mode == OX_EXP_CR)
        return expf_dd(x);
    return expf(x);
}

void ox_expf_range(int32_t mode, uint32_t first_bits, int64_t count, float *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; i++) {
        const uint32_t b = first_bits + (uint32_t)i;
        float x;
        memcpy(&x, &b, sizeof x);
        out[i] = ox_expf(mode, x);
    }
}

void ox_expf_array(int32_t mode, const float *x, int64_t n, float *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = ox_expf(mode, x[i]);
}

/* ResidualMetric::new, src/metric.rs:84-177 (f32 instantiation). */
ox_metric *ox_metric_residual(int64_t n, int32_t m, int32_t d, const float *coords,
                              const float *v, const float *p, float var,
                              const float *ls, int32_t n_ls, int32_t kernel_type,
                              int32_t exp_mode)
{
    ox_metric *mt = (ox_metric *)calloc(1, sizeof(*mt));
    mt->n = n;
    mt->m = m;
    mt->d = d;
    mt->kind = OX_METRIC_RESIDUAL;
    mt->kernel_type = kernel_type == 1 ? OX_KERNEL_MATERN52 : OX_KERNEL_RBF; /* metric.rs:50-56 */
    mt->exp_mode = exp_mode;
    mt->var = var;
    mt->scaled = (float *)malloc(sizeof(float) * (size_t)(n * d + 1));
    mt->norms = (float *)malloc(sizeof(float) * (size_t)(n + 1));
    mt->v = (float *)malloc(sizeof(float) * (size_t)(n * m + 1));
    mt->p = (float *)malloc(sizeof(float) * (size_t)(n + 1));
    float inv_ls[64];
    const float eps = 1e-6f;
    float fallback = n_ls > 0 ? ls[0] : 1.0f; /* metric.rs:95 */
    for (int32_t c = 0; c < d && c < 64; c++) {
        float raw = c < n_ls ? ls[c] : fallback;
        float safe = fabsf(raw) < eps ? eps : raw; /* metric.rs:98-99 */
        inv_ls[c] = 1.0f / safe;
    }
    for (int64_t r = 0; r < n; r++) {
        float acc = 0.0f; /* metric.rs:110-114: iterator sum, left to right */
        for (int32_t c = 0; c < d; c++) {
            float s = coords[r * d + c] * inv_ls[c];
            mt->scaled[r * d + c] = s;
            acc = acc + s * s;
        }
        mt->norms[r] = acc;
    }
    memcpy(mt->v, v, sizeof(float) * (size_t)(n * m));
    memcpy(mt->p, p, sizeof(float) * (size_t)n);
    return mt;
}

/* Euclidean, src/metric.rs:22-42: payload rows are the points themselves. */
ox_metric *ox_metric_euclidean(int64_t n, int32_t d, const float *points)
{
    ox_metric *mt = (ox_metric *)calloc(1, sizeof(*mt));
    mt->n = n;
    mt->d = d;
    mt->kind = OX_METRIC_EUCLIDEAN;
    mt->scaled = (float *)malloc(sizeof(float) * (size_t)(n * d + 1));
    memcpy(mt->scaled, points, sizeof(float) * (size_t)(n * d));
    return mt;
}

void ox_metric_free(ox_metric *mt)
{
    if (!mt)
        return;
    free(mt->scaled);
    free(mt->norms);
    free(mt->v);
    free(mt->p);
    free(mt);
}

/* wide::f32x8::reduce_add on the SSE2 baseline: two f32x4 halves, each an
 * in-order iterator sum, then the sum of halves (association parity unpinned). */
static inline float reduce8(const float *x)
{
    float lo = ((x[0] + x[1]) + x[2]) + x[3];
    float hi = ((x[4] + x[5]) + x[6]) + x[7];
    return lo + hi;
}

/* dot_product_simd_slice, src/metric.rs:693-733 (f32 branch). */
static float dot_simd_slice(const float *a, const float *b, int32_t len)
{
    float acc = 0.0f;
    int32_t chunks = len / 8, tail = chunks * 8;
    float prod[8];
    for (int32_t i = 0; i < tail; i += 8) {
        for (int l = 0; l < 8; l++)
            prod[l] = a[i + l] * b[i + l];
        acc += reduce8(prod);
    }
    for (int32_t j = tail; j < len; j++)
        acc += a[j] * b[j];
    return acc;
}

/* dot_f32_simd_inline, src/metric.rs:601-643. */
static float dot_simd_inline(const float *a, const float *b, int32_t len)
{
    int32_t chunks = len / 8, tail = chunks * 8;
    float acc0[8] = {0}, acc1[8] = {0};
    int32_t i = 0;
    for (int32_t dc = 0; dc < chunks / 2; dc++) {
        for (int l = 0; l < 8; l++) {
            float p0 = a[i + l] * b[i + l];
            float p1 = a[i + 8 + l] * b[i + 8 + l];
            acc0[l] = p0 + acc0[l];
            acc1[l] = p1 + acc1[l];
        }
        i += 16;
    }
    if (chunks % 2 == 1) {
        for (int l = 0; l < 8; l++) {
            float p0 = a[i + l] * b[i + l];
            acc0[l] = p0 + acc0[l];
        }
    }
    float comb[8];
    for (int l = 0; l < 8; l++)
        comb[l] = acc0[l] + acc1[l];
    float result = reduce8(comb);
    for (int32_t j = tail; j < len; j++)
        result += a[j] * b[j];
    return result;
}

static float euclid_sq(const ox_metric *mt, int64_t i, int64_t j)
{
    /* Euclidean::distance_sq, src/metric.rs:33-41: fold of squared diffs */
    const float *x = mt->scaled + i * mt->d, *y = mt->scaled + j * mt->d;
    float acc = 0.0f;
    for (int32_t c = 0; c < mt->d; c++) {
        float diff = x[c] - y[c];
        acc = acc + diff * diff;
    }
    return acc;
}

/* distance_sq_idx_f32, src/metric.rs:489-563: returns 1-|rho| (no sqrt).
 * For the Euclidean metric, the squared distance. */
float ox_dist_sq(const ox_metric *mt, int64_t i, int64_t j)
{
    if (mt->kind == OX_METRIC_EUCLIDEAN)
        return euclid_sq(mt, i, j);
    if (i == j)
        return 0.0f;
    const int32_t d = mt->d, m = mt->m;
    float dot_scaled = dot_simd_slice(mt->scaled + i * d, mt->scaled + j * d, d);
    float d2 = mt->norms[i] + mt->norms[j] - 2.0f * dot_scaled;
    if (d2 < 0.0f)
        d2 = 0.0f;
    float kval;
    if (mt->kernel_type == OX_KERNEL_RBF) {
        kval = mt->var * ox_expf(mt->exp_mode, -0.5f * d2);
    } else {
        float r = sqrtf(d2);
        const float sqrt5 = 2.2360679775f;
        float a = sqrt5 * r;
        float poly = 1.0f + a + (a * a) / 3.0f;
        kval = mt->var * poly * ox_expf(mt->exp_mode, -a);
    }
    const float eps = 1e-9f;
    float denom_sq = mt->p[i] * mt->p[j];
    float floor_sq = eps * eps;
    if (!(denom_sq >= floor_sq))
        denom_sq = floor_sq;
    float denom = sqrtf(denom_sq);
    float dot = dot_simd_slice(mt->v + i * m, mt->v + j * m, m);
    float rho = denom > 0.0f ? (kval - dot) / denom : 0.0f;
    if (isnan(rho)) /* f32::max returns the non-NaN operand */
        rho = -1.0f;
    if (rho < -1.0f)
        rho = -1.0f;
    if (rho > 1.0f)
        rho = 1.0f;
    float ds = 1.0f - fabsf(rho);
    return ds > 0.0f ? ds : 0.0f;
}

/* distances_f32_batch_idx_into_with_kth, src/metric.rs:376-487: the residual
 * distance sqrt(1-|rho|) used for query tiles.  Euclidean: sqrt of the sum. */
float ox_dist_batch(const ox_metric *mt, int64_t q, int64_t j)
{
    if (mt->kind == OX_METRIC_EUCLIDEAN)
        return sqrtf(euclid_sq(mt, q, j));
    if (q == j)
        return 0.0f;
    const int32_t d = mt->d, m = mt->m;
    const float *qc = mt->scaled + q * d, *pc = mt->scaled + j * d;
    float dot_c = 0.0f;
    if (d == 3) {
        dot_c = qc[0] * pc[0] + qc[1] * pc[1] + qc[2] * pc[2];
    } else {
        for (int32_t c = 0; c < d; c++)
            dot_c += qc[c] * pc[c];
    }
    float d2 = mt->norms[q] + mt->norms[j] - 2.0f * dot_c;
    if (d2 < 0.0f)
        d2 = 0.0f;
    float kval;
    if (mt->kernel_type == OX_KERNEL_RBF) {
        kval = mt->var * ox_expf(mt->exp_mode, -0.5f * d2);
    } else {
        float r = sqrtf(d2);
        const float sqrt5 = 2.2360679775f, inv_three = 0.3333333333f;
        float a = sqrt5 * r;
        float poly = 1.0f + a + (a * a) * inv_three;
        kval = mt->var * poly * ox_expf(mt->exp_mode, -a);
    }
    const float eps = 1e-9f;
    float prod = mt->p[q] * mt->p[j];
    float floor_sq = eps * eps;
    if (!(prod >= floor_sq))
        prod = floor_sq;
    float denom = sqrtf(prod);
    float dot_v = dot_simd_inline(mt->v + q * m, mt->v + j * m, m);
    float rho = (kval - dot_v) / denom;
    if (isnan(rho))
        rho = -1.0f;
    if (rho < -1.0f)
        rho = -1.0f;
    if (rho > 1.0f)
        rho = 1.0f;
    float ds = 1.0f - fabsf(rho);
    return ds > 0.0f ? sqrtf(ds) : 0.0f;
}

void ox_dist_sq_pairs(const ox_metric *mt, const int64_t *a, const int64_t *b, int64_t n,
                      float *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = ox_dist_sq(mt, a[i], b[i]);
}

void ox_dist_batch_pairs(const ox_metric *mt, const int64_t *a, const int64_t *b, int64_t n,
                         float *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = ox_dist_batch(mt, a[i], b[i]);
}
