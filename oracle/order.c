/* Build orderings -- oracle restatement.  TEST INFRASTRUCTURE ONLY. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

/* part1by2, src/pcct.rs:10-18: spread 21 bits to every third bit. */
static uint64_t part1by2(uint64_t n)
{
    n &= 0x1fffffULL;
    n = (n | (n << 32)) & 0x1f00000000ffffULL;
    n = (n | (n << 16)) & 0x1f0000ff0000ffULL;
    n = (n | (n << 8)) & 0x100f00f00f00f00fULL;
    n = (n | (n << 4)) & 0x10c30c30c30c30c3ULL;
    n = (n | (n << 2)) & 0x1249249249249249ULL;
    return n;
}

typedef struct {
    uint64_t code;
    int64_t idx;
} code_idx;

static int cmp_code_idx(const void *a, const void *b)
{
    const code_idx *x = (const code_idx *)a, *y = (const code_idx *)b;
    if (x->code != y->code)
        return x->code < y->code ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* hilbert_like_order, src/pcct.rs:27-67: 21-bit Morton code on D=3, stable
 * sort by code; identity for any other D. */
void ox_morton_order(int64_t n, int32_t d, const float *coords, int64_t *order_out,
                     uint64_t *codes_out)
{
    if (n <= 0)
        return;
    if (d != 3) {
        for (int64_t i = 0; i < n; i++) {
            order_out[i] = i;
            if (codes_out)
                codes_out[i] = 0;
        }
        return;
    }
    float mins[3] = {INFINITY, INFINITY, INFINITY};
    float maxs[3] = {-INFINITY, -INFINITY, -INFINITY};
    /* f32::MAX / f32::MIN starting values compare the same for finite data */
    for (int c = 0; c < 3; c++) {
        mins[c] = 3.40282347e38f;
        maxs[c] = -3.40282347e38f;
    }
    for (int64_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) {
            float v = coords[i * 3 + c];
            if (v < mins[c])
                mins[c] = v;
            if (v > maxs[c])
                maxs[c] = v;
        }
    const float eps = 1e-9f;
    code_idx *ci = (code_idx *)malloc(sizeof(code_idx) * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        uint64_t u[3];
        for (int c = 0; c < 3; c++) {
            float denom = fabsf(maxs[c] - mins[c]);
            if (!(denom >= eps))
                denom = eps;
            float norm = (coords[i * 3 + c] - mins[c]) / denom;
            if (norm < 0.0f)
                norm = 0.0f;
            if (norm > 1.0f)
                norm = 1.0f;
            float s = norm * 2097151.0f;
            /* `as u32`: truncation toward zero, NaN -> 0 */
            u[c] = isnan(s) ? 0u : (uint64_t)(uint32_t)s;
        }
        ci[i].code = part1by2(u[0]) | (part1by2(u[1]) << 1) | (part1by2(u[2]) << 2);
        ci[i].idx = i;
    }
    qsort(ci, (size_t)n, sizeof(code_idx), cmp_code_idx);
    for (int64_t i = 0; i < n; i++) {
        order_out[i] = ci[i].idx;
        if (codes_out)
            codes_out[i] = ci[i].code;
    }
    free(ci);
}

/* _hilbert_permutation code step, covertreex/algo/order/strategy.py:38-99:
 * bits = max(1, min(16, 64 // D)); quantise with round-half-even after
 * (p - lo) / max(hi - lo, 1e-9) in float64; Skilling AxesToTranspose then
 * interleave most-significant bit first.  Codes only; the caller sorts by
 * (code, index). */
void ox_hilbert_codes(int64_t n, int32_t d, const double *points, uint64_t *codes_out)
{
    if (n <= 0 || d <= 0)
        return;
    int bits = 64 / d;
    if (bits > 16)
        bits = 16;
    if (bits < 1)
        bits = 1;
    double *lo = (double *)malloc(sizeof(double) * d), *hi = (double *)malloc(sizeof(double) * d);
    for (int c = 0; c < d; c++) {
        lo[c] = INFINITY;
        hi[c] = -INFINITY;
    }
    for (int64_t i = 0; i < n; i++)
        for (int c = 0; c < d; c++) {
            double v = points[i * d + c];
            lo[c] = v < lo[c] ? v : lo[c];
            hi[c] = v > hi[c] ? v : hi[c];
        }
    const uint64_t max_val = (1ULL << bits) - 1;
    uint64_t *x = (uint64_t *)malloc(sizeof(uint64_t) * d);
    for (int64_t i = 0; i < n; i++) {
        for (int c = 0; c < d; c++) {
            double span = hi[c] - lo[c];
            if (!(span >= 1e-9))
                span = 1e-9;
            double s = (points[i * d + c] - lo[c]) / span;
            double r = nearbyint(s * (double)max_val); /* np.round: half to even */
            if (r < 0.0)
                r = 0.0;
            if (r > (double)max_val)
                r = (double)max_val;
            x[c] = (uint64_t)r;
        }
        /* _axes_to_transpose, strategy.py:63-88 */
        uint64_t M = 1ULL << (bits - 1);
        for (uint64_t Q = M; Q > 1; Q >>= 1) {
            uint64_t P = Q - 1;
            for (int c = 0; c < d; c++) {
                if (x[c] & Q) {
                    x[0] ^= P;
                } else {
                    uint64_t t = (x[0] ^ x[c]) & P;
                    x[0] ^= t;
                    x[c] ^= t;
                }
            }
        }
        for (int c = 1; c < d; c++)
            x[c] ^= x[c - 1];
        uint64_t t = 0;
        for (uint64_t Q = M; Q > 1; Q >>= 1)
            if (x[d - 1] & Q)
                t ^= Q - 1;
        for (int c = 0; c < d; c++)
            x[c] ^= t;
        /* _transpose_to_hilbert, strategy.py:91-99 */
        uint64_t index = 0;
        for (int b = bits - 1; b >= 0; b--) {
            uint64_t word = 0;
            for (int c = 0; c < d; c++)
                word = (word << 1) | ((x[c] >> b) & 1ULL);
            index = (index << d) | word;
        }
        codes_out[i] = index;
    }
    free(x);
    free(lo);
    free(hi);
}
