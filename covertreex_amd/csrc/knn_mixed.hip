// Mixed-precision batched residual kNN (CX_KNN_MIXED): exact neighbours under
// d = sqrt(1 - |rho|) evaluated on the stored inputs, with a certified screen.
//
// Work is laid out as query tiles x candidate tiles in dataset order, so the
// Vecchia mask j < i (src/algo.rs:918-925 / src/lib.rs:444-448) becomes a
// triangle: a query tile only visits candidates below its largest index.
//
//   prep      V -> dataset-order BF16 hi/lo split (V = hi + lo + r, |r| <=
//             2^-18 |V|), row norms |V_i| rounded up.
//   screen    one workgroup per 128-query tile x 2048-candidate segment, four
//             wavefronts of 64 x 64; V.V on MFMA (v_mfma_f32_32x32x16_bf16)
//             as hi.hi + (hi.lo + lo.hi) in two FP32 accumulators, kernel
//             value in FP32, and a certified interval [lb2, ub2] containing
//             the exact squared distance 1 - |rho|.
//             pass 1: per-query histogram of ub2 -> tau_q, an upper bound on
//                     the k-th exact squared distance;
//             pass 2: every candidate with lb2 <= tau_q is appended to the
//                     query's list (a superset of the exact k nearest).
//   rescore   one wavefront per query: compensated FP32 (double-float)
//             distance for each listed candidate, exact top-k, ascending
//             output with a per-result error bound.  A query whose list
//             overflowed is rescored over all its valid candidates.
#include <hipcub/hipcub.hpp>

#include "cx_internal.hpp"
#include "df.hpp"

namespace cx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kQT = 128;     // queries per workgroup tile
constexpr int kCT = 128;     // candidates per candidate tile
constexpr int kKC = 32;      // K chunk staged in LDS
constexpr int kKP = kKC + 8; // padded LDS row (bf16)
constexpr int kBins = 128;   // histogram bins of the squared upper bound over [0, 1)
constexpr float kHistCap = 0.99f; // squared bounds at or above this are not counted
constexpr int kCap = 2048;   // listed candidates per query before overflow
constexpr int kMaxD = 16;
constexpr int kMixMaxK = 256;
constexpr int64_t kSeg = 16 * kCT; // candidates per workgroup (grid y)

// ---- prep ------------------------------------------------------------------

__device__ __forceinline__ uint16_t bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// V (tree order) -> dataset-order hi/lo, zero padded to mp columns.
__global__ void split_kernel(const float *__restrict__ v, const int64_t *__restrict__ inv_order,
                             int64_t n, int32_t m, int32_t mp, uint16_t *__restrict__ vhi,
                             uint16_t *__restrict__ vlo, float *__restrict__ vnorm)
{
    const int64_t row = blockIdx.x; // dataset index
    if (row >= n)
        return;
    const float *src = v + inv_order[row] * (int64_t)m;
    double acc = 0.0;
    for (int32_t c = threadIdx.x; c < mp; c += blockDim.x) {
        float x = c < m ? src[c] : 0.0f;
        uint16_t h = bf16_rne(x);
        uint16_t l = bf16_rne(x - bf16_to_f(h));
        vhi[row * (int64_t)mp + c] = h;
        vlo[row * (int64_t)mp + c] = l;
        acc += (double)x * (double)x;
    }
    __shared__ double s_acc[64];
    for (int o = 32; o > 0; o >>= 1)
        acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0)
        s_acc[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++)
            t += s_acc[w];
        // round the norm up so products of norms bound |V_i . V_j| from above
        vnorm[row] = (float)__builtin_sqrt(t) * (1.0f + 1.2e-7f) + 1e-30f;
    }
}

__global__ void gather_meta_kernel(const float *__restrict__ raw, const float *__restrict__ p,
                                   const int64_t *__restrict__ inv_order, int64_t n, int32_t d,
                                   float *__restrict__ x_ds, float *__restrict__ isp_ds)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t t = inv_order[i];
    for (int32_t c = 0; c < d; c++)
        x_ds[i * d + c] = raw[t * d + c];
    // correctly rounded sqrt and division: relative error <= 2^-23
    isp_ds[i] = 1.0f / __builtin_sqrtf(p[t]);
}

// ---- screen ------------------------------------------------------------------

struct ScreenArgs {
    const uint16_t *vhi, *vlo;
    const float *x;     // n x d raw coords, dataset order
    const float *isp;   // n: 1/sqrt(p), relative error <= 2^-23
    const float *vnorm; // n
    const int64_t *qsorted;  // sorted query dataset indices (nq)
    const int32_t *qrow;     // dataset row used for each sorted query
    int64_t n, nq;
    int32_t d, mp, kernel, pred, k;
    float var;
    float inv_ls[kMaxD];
    float cD, eK;
    float *tau;      // nq
    int32_t *hist;   // nq x kBins
    int32_t *cnt;    // nq
    int32_t *list;   // nq x kCap
    int32_t n_tiles;
};

__device__ __forceinline__ float fast_kernel(float d2, float var, int32_t kernel)
{
    if (kernel == 0)
        return var * __expf(-0.5f * d2);
    float a = __builtin_sqrtf(5.0f * d2);
    return var * (1.0f + a + a * a * 0.333333343f) * __expf(-a);
}

template <int PASS>
__global__ void __launch_bounds__(256) screen_kernel(ScreenArgs A)
{
    __shared__ __attribute__((aligned(16))) uint16_t sAh[kQT * kKP];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[kQT * kKP];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[kCT * kKP];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[kCT * kKP];
    __shared__ float sQx[kQT * kMaxD], sQp[kQT], sQn[kQT], sQtau[kQT];
    __shared__ int64_t sQid[kQT];
    __shared__ int32_t sQlive[kQT];
    __shared__ int32_t sQrow[kQT];
    __shared__ float sCx[kCT * kMaxD], sCp[kCT], sCn[kCT];
    // pass-1 histogram: two 16-bit bins per word (at most kSeg counts per bin)
    __shared__ uint32_t sHist[PASS == 1 ? kQT * kBins / 2 : 1];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    // blockIdx.x = query tile, blockIdx.y = candidate segment; heavy tiles first
    const int32_t tile = A.n_tiles - 1 - (int32_t)blockIdx.x;
    const int64_t s0 = (int64_t)tile * kQT;
    {
        // candidate range of this tile: [0, largest query index) under the mask
        int64_t c_lim = A.n;
        if (A.pred) {
            const int64_t last = s0 + kQT - 1 < A.nq - 1 ? s0 + kQT - 1 : A.nq - 1;
            const int64_t qmax = A.qsorted[last];
            c_lim = qmax < 0 ? A.n : (qmax < A.n ? qmax : A.n);
        }
        if ((int64_t)blockIdx.y * kSeg >= c_lim)
            return;
    }
    const int32_t d = A.d, mp = A.mp;

    for (int i = tid; i < kQT; i += 256) {
        const int64_t s = s0 + i;
        const bool live = s < A.nq;
        const int32_t row = live ? A.qrow[s] : 0;
        sQrow[i] = row;
        sQid[i] = live ? A.qsorted[s] : 0;
        sQlive[i] = live ? 1 : 0;
        for (int c = 0; c < d; c++)
            sQx[i * kMaxD + c] = A.x[(int64_t)row * d + c];
        sQp[i] = A.isp[row];
        sQn[i] = A.vnorm[row];
        if (PASS == 2)
            sQtau[i] = live ? A.tau[s] : -1.0f;
    }
    if (PASS == 1)
        for (int i = tid; i < kQT * kBins / 2; i += 256)
            sHist[i] = 0u;
    __syncthreads();

    int64_t c_end = A.n;
    if (A.pred) {
        const int64_t last = s0 + kQT - 1 < A.nq - 1 ? s0 + kQT - 1 : A.nq - 1;
        int64_t qmax = A.qsorted[last];
        c_end = qmax < 0 ? A.n : (qmax < A.n ? qmax : A.n);
    }
    const int64_t c_begin = (int64_t)blockIdx.y * kSeg;
    c_end = c_end < c_begin + kSeg ? c_end : c_begin + kSeg;
    const int wq0 = (wid >> 1) * 64, wc0 = (wid & 1) * 64;

    for (int64_t c0 = c_begin; c0 < c_end; c0 += kCT) {
        __syncthreads(); // the previous tile's epilogue still reads sCx/sCp/sCn
        for (int i = tid; i < kCT; i += 256) {
            const int64_t j = c0 + i;
            const int64_t jr = j < A.n ? j : 0;
            for (int c = 0; c < d; c++)
                sCx[i * kMaxD + c] = A.x[jr * d + c];
            sCp[i] = A.isp[jr];
            sCn[i] = A.vnorm[jr];
        }
        floatx16 acc_hh[2][2], acc_x[2][2];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                acc_hh[a][b] = floatx16{};
                acc_x[a][b] = floatx16{};
            }
        for (int kc = 0; kc < mp; kc += kKC) {
            __syncthreads();
            // stage: 128 rows x 32 bf16 (4 x 16 B segments) for hi and lo
#pragma unroll
            for (int it = 0; it < 4; it++) {
                const int idx = tid + 256 * it; // [0, 1024)
                const int part = idx >> 9, rem = idx & 511;
                const int row = rem >> 2, seg = rem & 3;
                const uint16_t *srcA = (part ? A.vlo : A.vhi) +
                                       (int64_t)sQrow[row] * mp + kc + seg * 8;
                uint16_t *dstA = (part ? sAl : sAh) + row * kKP + seg * 8;
                *(uint4 *)dstA = *(const uint4 *)srcA;
                const int64_t crow = c0 + row < A.n ? c0 + row : 0;
                const uint16_t *srcB = (part ? A.vlo : A.vhi) + crow * mp + kc + seg * 8;
                uint16_t *dstB = (part ? sBl : sBh) + row * kKP + seg * 8;
                *(uint4 *)dstB = *(const uint4 *)srcB;
            }
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < kKC; ks += 16) {
                const int koff = ks + 8 * (lane >> 5);
                bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int ar = wq0 + 32 * t + (lane & 31);
                    const int br = wc0 + 32 * t + (lane & 31);
                    ah[t] = __builtin_bit_cast(bf16x8, *(const uint4 *)(sAh + ar * kKP + koff));
                    al[t] = __builtin_bit_cast(bf16x8, *(const uint4 *)(sAl + ar * kKP + koff));
                    bh[t] = __builtin_bit_cast(bf16x8, *(const uint4 *)(sBh + br * kKP + koff));
                    bl[t] = __builtin_bit_cast(bf16x8, *(const uint4 *)(sBl + br * kKP + koff));
                }
#pragma unroll
                for (int a = 0; a < 2; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        acc_hh[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            ah[a], bh[b], acc_hh[a][b], 0, 0, 0);
                        acc_x[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            ah[a], bl[b], acc_x[a][b], 0, 0, 0);
                        acc_x[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            al[a], bh[b], acc_x[a][b], 0, 0, 0);
                    }
            }
        }
        // epilogue: certified interval per (query, candidate)
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int ql = wq0 + 32 * a + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    const int cl = wc0 + 32 * b + (lane & 31);
                    const int64_t qid = sQid[ql];
                    const int64_t j = c0 + cl;
                    if (!sQlive[ql] || j >= A.n)
                        continue;
                    if (A.pred && qid >= 0 && j >= qid) // negative index: no bound
                        continue;
                    // squared bounds on d^2 = 1 - |rho|
                    float lb2, ub2;
                    if (j == (int64_t)sQrow[ql]) {
                        lb2 = 0.0f;
                        ub2 = 0.0f;
                    } else {
                        float d2 = 0.0f;
                        for (int c = 0; c < d; c++) {
                            float df_ = (sQx[ql * kMaxD + c] - sCx[cl * kMaxD + c]) * A.inv_ls[c];
                            d2 += df_ * df_;
                        }
                        const float kv = fast_kernel(d2, A.var, A.kernel);
                        const float dot = acc_hh[a][b][r] + acc_x[a][b][r];
                        // 1/sqrt(max(p_q p_j, 1e-18)) from per-row 1/sqrt(p)
                        float w = sQp[ql] * sCp[cl];
                        w = w < 1e9f ? w : 1e9f;
                        const float num = kv - dot;
                        const float ar = __builtin_fabsf(num * w);
                        const float erho =
                            (A.eK + A.cD * (sQn[ql] * sCn[cl]) + __builtin_fabsf(num) * 1.2e-7f) *
                                w * 1.000002f +
                            ar * 6e-7f;
                        const float rlo = ar - erho > 0.0f ? ar - erho : 0.0f;
                        const float rhi = ar + erho < 1.0f ? ar + erho : 1.0f;
                        lb2 = (1.0f - rhi) - 6e-8f;
                        ub2 = (1.0f - rlo) + 6e-8f;
                    }
                    if (PASS == 1) {
                        if (ub2 < kHistCap) {
                            const int bin = (int)(ub2 * (float)kBins);
                            atomicAdd(&sHist[(ql * kBins + bin) >> 1], (bin & 1) ? 65536u : 1u);
                        }
                    } else {
                        if (lb2 <= sQtau[ql]) {
                            const int64_t s = s0 + ql;
                            const int32_t pos = atomicAdd(&A.cnt[s], 1);
                            if (pos < kCap)
                                A.list[s * kCap + pos] = (int32_t)j;
                        }
                    }
                }
    }
    if (PASS == 1) {
        __syncthreads();
        for (int i = tid; i < kQT * kBins / 2; i += 256) {
            const int64_t s = s0 + (2 * i) / kBins;
            const uint32_t h = sHist[i];
            if (s < A.nq && h) {
                const int bin = (2 * i) % kBins;
                if (h & 0xffffu)
                    atomicAdd(&A.hist[s * kBins + bin], (int32_t)(h & 0xffffu));
                if (h >> 16)
                    atomicAdd(&A.hist[s * kBins + bin + 1], (int32_t)(h >> 16));
            }
        }
    }
}

// tau_q (squared scale): upper edge of the first bin where the cumulative count
// of squared upper bounds reaches k; 2.0 (keep everything) when fewer than k
// candidates have a squared upper bound below the histogram cap.
__global__ void tau_kernel(const int32_t *__restrict__ hist, int64_t nq, int32_t k, float *tau)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nq)
        return;
    int cum = 0;
    float t = 2.0f;
    for (int b = 0; b < kBins; b++) {
        cum += hist[s * kBins + b];
        if (cum >= k) {
            t = (float)(b + 1) / (float)kBins;
            break;
        }
    }
    tau[s] = t;
}

// ---- rescore -----------------------------------------------------------------

struct RescoreArgs {
    const float *v;       // n x m tree order
    const float *raw;     // n x d tree order
    const float *p;       // n tree order
    const int64_t *inv_order;
    const int64_t *qsorted;
    const int64_t *qperm; // sorted position -> original query position
    const int32_t *qrow;
    const int32_t *cnt;
    const int32_t *list;
    int64_t n, nq;
    int32_t d, m, kernel, pred, k;
    float var;
    float inv_ls_hi[kMaxD], inv_ls_lo[kMaxD];
    int64_t *out_idx;
    float *out_dist, *out_err;
    int32_t *counter;
};

// Exact-arithmetic surrogate of the residual distance between dataset rows q
// and j: every step in double-float.  Returns the distance and a bound on its
// error.
__device__ __forceinline__ df accurate_distance(const RescoreArgs &A, int64_t qt, int64_t jt,
                                                float *err)
{
    const int32_t d = A.d, m = A.m;
    const float *xq = A.raw + qt * d, *xj = A.raw + jt * d;
    df d2 = {0.0f, 0.0f};
    for (int c = 0; c < d; c++) {
        df diff = two_sum(xq[c], -xj[c]);
        df s = df_mul(diff, df{A.inv_ls_hi[c], A.inv_ls_lo[c]});
        d2 = df_add(d2, df_mul(s, s));
    }
    df kv;
    if (A.kernel == 0) {
        kv = df_exp(df_mul_f(d2, -0.5f));
    } else {
        df a = df_sqrt(df_mul_f(d2, 5.0f));
        constexpr DfConst third = split_const(1.0 / 3.0);
        df poly = df_add(df_add(df{1.0f, 0.0f}, a), df_mul(df_mul(a, a), df{third.hi, third.lo}));
        kv = df_mul(poly, df_exp(df_neg(a)));
    }
    kv = df_mul_f(kv, A.var);
    const float *vq = A.v + qt * m, *vj = A.v + jt * m;
    df dot = {0.0f, 0.0f};
    float sabs = 0.0f;
    for (int c = 0; c < m; c++) {
        dot = df_acc(dot, two_prod(vq[c], vj[c]));
        sabs += __builtin_fabsf(vq[c] * vj[c]);
    }
    df num = df_sub(kv, dot);
    df pp = two_prod(A.p[qt], A.p[jt]);
    if (pp.hi < 1e-18f)
        pp = {1e-18f, 0.0f};
    df den = df_sqrt(pp);
    df rho = df_div(num, den);
    if (rho.hi < 0.0f)
        rho = df_neg(rho);
    if (rho.hi > 1.0f || (rho.hi == 1.0f && rho.lo > 0.0f)) // |rho| clamped to 1
        rho = {1.0f, 0.0f};
    df s = df_sub(df{1.0f, 0.0f}, rho);
    df dist = df_sqrt(s);
    // error: dot accumulation (m+8) * 2^-46 * sum|products|, kernel 2^-44,
    // then through the division and the square root; the output rounding to
    // f32 adds half an ulp.
    const float e_num = ((float)(m + 8) * 1.5e-14f * (sabs * 1.01f + __builtin_fabsf(kv.hi)) +
                         __builtin_fabsf(kv.hi) * 6e-14f);
    const float e_rho = e_num / den.hi * 1.001f + rho.hi * 3e-14f + 1e-37f;
    const float sd = dist.hi;
    float e_d = sd > 0.0f ? e_rho / (2.0f * sd) : __builtin_sqrtf(e_rho);
    const float e_sq = __builtin_sqrtf(e_rho);
    e_d = e_d < e_sq ? e_d : e_sq;
    *err = e_d * 1.01f + sd * 6e-8f;
    return dist;
}

__global__ void __launch_bounds__(64) rescore_kernel(RescoreArgs A)
{
    __shared__ float s_dh[kMixMaxK], s_dl[kMixMaxK], s_er[kMixMaxK];
    __shared__ int32_t s_j[kMixMaxK];
    __shared__ int64_t s_qi;
    const int lane = threadIdx.x;
    const int32_t k = A.k;
    for (;;) {
        if (lane == 0)
            s_qi = atomicAdd(A.counter, 1);
        __syncthreads();
        const int64_t s = s_qi;
        __syncthreads();
        if (s >= A.nq)
            break;
        const int64_t qid = A.qsorted[s];
        const int64_t pos_out = A.qperm[s];
        const int32_t qrow = A.qrow[s];
        const int64_t qt = A.inv_order[qrow];
        for (int j = lane; j < k; j += 64) {
            A.out_idx[pos_out * k + j] = -1;
            A.out_dist[pos_out * k + j] = 3.40282347e38f;
            if (A.out_err)
                A.out_err[pos_out * k + j] = 0.0f;
        }
        // Vecchia bound on the original index (negative -> unbounded)
        const int64_t limit = A.pred ? (qid < 0 ? A.n : (qid < A.n ? qid : A.n)) : A.n;
        const int32_t c = A.cnt[s];
        const bool overflow = c > kCap;
        const int64_t total = overflow ? limit : (int64_t)c;
        int count = 0;
        float maxh = -1.0f, maxl = 0.0f;
        int32_t maxj = -1;
        int maxpos = 0;
        for (int64_t base = 0; base < total; base += 64) {
            const int64_t e = base + lane;
            const bool live = e < total;
            int64_t j = 0;
            if (live)
                j = overflow ? e : (int64_t)A.list[s * kCap + e];
            df dist = {0.0f, 0.0f};
            float err = 0.0f;
            if (live) {
                if (j == qrow)
                    dist = {0.0f, 0.0f};
                else
                    dist = accurate_distance(A, qt, A.inv_order[j], &err);
            }
            bool better = live && (count < k || df_lt(dist, df{maxh, maxl}) ||
                                   (dist.hi == maxh && dist.lo == maxl && j < maxj));
            uint64_t mask = __ballot(better);
            while (mask) {
                const int l = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                const float dh = __shfl(dist.hi, l, 64);
                const float dl = __shfl(dist.lo, l, 64);
                const float er = __shfl(err, l, 64);
                const int32_t jj = (int32_t)__shfl((int)j, l, 64);
                int slot;
                if (count < k) {
                    slot = count++;
                } else {
                    const bool lt = dh < maxh || (dh == maxh && (dl < maxl ||
                                                  (dl == maxl && jj < maxj)));
                    if (!lt)
                        continue;
                    slot = maxpos;
                }
                if (lane == 0) {
                    s_dh[slot] = dh;
                    s_dl[slot] = dl;
                    s_er[slot] = er;
                    s_j[slot] = jj;
                }
                __syncthreads();
                // refresh the maximum (largest distance, then largest index)
                float bh = -1.0f, bl = 0.0f;
                int32_t bj = -1;
                int bp = 0;
                for (int t = lane; t < count; t += 64) {
                    const float h = s_dh[t], lo = s_dl[t];
                    const int32_t jv = s_j[t];
                    if (h > bh || (h == bh && (lo > bl || (lo == bl && jv > bj)))) {
                        bh = h;
                        bl = lo;
                        bj = jv;
                        bp = t;
                    }
                }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    const float oh = __shfl_xor(bh, o, 64), ol = __shfl_xor(bl, o, 64);
                    const int32_t oj = __shfl_xor(bj, o, 64);
                    const int op = __shfl_xor(bp, o, 64);
                    if (oh > bh || (oh == bh && (ol > bl || (ol == bl && oj > bj)))) {
                        bh = oh;
                        bl = ol;
                        bj = oj;
                        bp = op;
                    }
                }
                maxh = bh;
                maxl = bl;
                maxj = bj;
                maxpos = bp;
            }
        }
        __syncthreads();
        // ascending output: (distance, dataset index)
        for (int t = lane; t < count; t += 64) {
            const float h = s_dh[t], lo = s_dl[t];
            const int32_t jv = s_j[t];
            int rank = 0;
            for (int u = 0; u < count; u++) {
                const float uh = s_dh[u], ul = s_dl[u];
                rank += uh < h || (uh == h && (ul < lo || (ul == lo && s_j[u] < jv)));
            }
            A.out_idx[pos_out * k + rank] = jv;
            A.out_dist[pos_out * k + rank] = h + lo;
            if (A.out_err)
                A.out_err[pos_out * k + rank] = s_er[t];
        }
        __syncthreads();
    }
}

// ---- driver -----------------------------------------------------------------

__global__ void query_rows_kernel(const int64_t *__restrict__ qsorted, int64_t nq, int64_t n,
                                  const int64_t *__restrict__ order, int32_t *qrow)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nq)
        return;
    const int64_t q = qsorted[s];
    // out-of-range queries read tree node 0, as the reference maps them (src/lib.rs:414-418)
    qrow[s] = (int32_t)((q >= 0 && q < n) ? q : order[0]);
}

__global__ void iota64_kernel(int64_t *out, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = i;
}

static void prepare_mixed(TreeState &t)
{
    if (t.mixed_ready)
        return;
    hipStream_t st = t.stream;
    const int64_t n = t.n;
    t.mp = ((t.m + kKC - 1) / kKC) * kKC;
    const int64_t npad = ((n + kCT - 1) / kCT) * kCT + kCT;
    t.vhi.alloc((size_t)npad * t.mp);
    t.vlo.alloc((size_t)npad * t.mp);
    CX_HIP(hipMemsetAsync(t.vhi.ptr, 0, t.vhi.bytes(), st));
    CX_HIP(hipMemsetAsync(t.vlo.ptr, 0, t.vlo.bytes(), st));
    t.vnorm.alloc(npad);
    CX_HIP(hipMemsetAsync(t.vnorm.ptr, 0, t.vnorm.bytes(), st));
    split_kernel<<<(unsigned)n, 256, 0, st>>>(t.v.ptr, t.inv_order.ptr, n, t.m, t.mp, t.vhi.ptr,
                                              t.vlo.ptr, t.vnorm.ptr);
    t.x_ds.alloc((size_t)npad * t.d);
    t.isp_ds.alloc(npad);
    CX_HIP(hipMemsetAsync(t.x_ds.ptr, 0, t.x_ds.bytes(), st));
    CX_HIP(hipMemsetAsync(t.isp_ds.ptr, 0, t.isp_ds.bytes(), st));
    gather_meta_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
        t.raw.ptr, t.p.ptr, t.inv_order.ptr, n, t.d, t.x_ds.ptr, t.isp_ds.ptr);
    CX_HIP(hipGetLastError());
    CX_HIP(hipStreamSynchronize(st));
    t.mixed_ready = true;
}

void knn_mixed(TreeState &t, const int64_t *d_queries, int64_t nq, int32_t k, int32_t predecessor,
               int64_t *d_idx, float *d_dist, float *d_err)
{
    CX_REQUIRE(k >= 1 && k <= kMixMaxK, "mixed-precision kNN supports 1 <= k <= 256");
    CX_REQUIRE(t.d <= kMaxD, "mixed-precision kNN supports d <= 16");
    if (nq <= 0)
        return;
    prof_mark(t, 0);
    prepare_mixed(t);
    prof_mark(t, 1);
    hipStream_t st = t.stream;
    const int64_t n = t.n;

    // workspace: sorted queries, permutation, rows, tau, counts, lists
    auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_qs = off;
    off += align(sizeof(int64_t) * nq);
    const size_t o_perm = off;
    off += align(sizeof(int64_t) * nq);
    const size_t o_iota = off;
    off += align(sizeof(int64_t) * nq);
    const size_t o_row = off;
    off += align(sizeof(int32_t) * nq);
    const size_t o_tau = off;
    off += align(sizeof(float) * nq);
    const size_t o_cnt = off;
    off += align(sizeof(int32_t) * nq);
    const size_t o_hist = off;
    off += align(sizeof(int32_t) * nq * kBins);
    const size_t o_ctr = off;
    off += 256;
    const size_t o_list = off;
    off += align(sizeof(int32_t) * (size_t)nq * kCap);
    size_t sort_bytes = 0;
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const int64_t *)nullptr,
                                              (int64_t *)nullptr, (const int64_t *)nullptr,
                                              (int64_t *)nullptr, (int)nq, 0, 64, st));
    const size_t o_sort = off;
    off += align(sort_bytes);
    uint8_t *ws = workspace(t, off);
    int64_t *qs = (int64_t *)(ws + o_qs), *perm = (int64_t *)(ws + o_perm);
    int64_t *iota = (int64_t *)(ws + o_iota);
    int32_t *qrow = (int32_t *)(ws + o_row), *cnt = (int32_t *)(ws + o_cnt);
    int32_t *ctr = (int32_t *)(ws + o_ctr), *list = (int32_t *)(ws + o_list);
    float *tau = (float *)(ws + o_tau);
    int32_t *hist = (int32_t *)(ws + o_hist);

    const unsigned qb = (unsigned)((nq + 255) / 256);
    iota64_kernel<<<qb, 256, 0, st>>>(iota, nq);
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(ws + o_sort, sort_bytes, d_queries, qs, iota, perm,
                                              (int)nq, 0, 64, st));
    query_rows_kernel<<<qb, 256, 0, st>>>(qs, nq, n, t.order.ptr, qrow);
    prof_mark(t, 2);
    CX_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * nq, st));
    CX_HIP(hipMemsetAsync(hist, 0, sizeof(int32_t) * nq * kBins, st));
    CX_HIP(hipMemsetAsync(ctr, 0, sizeof(int32_t), st));

    ScreenArgs sa{};
    sa.vhi = t.vhi.ptr;
    sa.vlo = t.vlo.ptr;
    sa.x = t.x_ds.ptr;
    sa.isp = t.isp_ds.ptr;
    sa.vnorm = t.vnorm.ptr;
    sa.qsorted = qs;
    sa.qrow = qrow;
    sa.n = n;
    sa.nq = nq;
    sa.d = t.d;
    sa.mp = t.mp;
    sa.kernel = t.kernel;
    sa.pred = predecessor ? 1 : 0;
    sa.k = k;
    sa.var = t.var;
    for (int c = 0; c < t.d; c++)
        sa.inv_ls[c] = t.inv_ls[c];
    // |computed V.V - exact| <= cD |V_q||V_j|: the hi/lo split drops at most
    // 3.1 * 2^-18 of each product, the hi.hi FP32 accumulation adds gamma_mp,
    // the cross terms (2^-8 smaller) gamma_2mp, and the final sum one rounding.
    const double u = 1.0 / 16777216.0;
    const double g1 = (t.mp + 16) * u / (1.0 - (t.mp + 16) * u);
    const double g2 = (2 * t.mp + 16) * u / (1.0 - (2 * t.mp + 16) * u);
    sa.cD = (float)((3.1 / 262144.0 + g1 * 1.01 + g2 * 0.01 + 2 * u) * 1.05);
    sa.eK = (float)(std::fabs((double)t.var) / 262144.0); // fast exp/kernel in FP32
    sa.tau = tau;
    sa.hist = hist;
    sa.cnt = cnt;
    sa.list = list;
    sa.n_tiles = (int32_t)((nq + kQT - 1) / kQT);
    const dim3 grid((unsigned)sa.n_tiles, (unsigned)((n + kSeg - 1) / kSeg));
    screen_kernel<1><<<grid, 256, 0, st>>>(sa);
    CX_HIP(hipGetLastError());
    prof_mark(t, 3);
    tau_kernel<<<qb, 256, 0, st>>>(hist, nq, k, tau);
    prof_mark(t, 4);
    screen_kernel<2><<<grid, 256, 0, st>>>(sa);
    CX_HIP(hipGetLastError());
    prof_mark(t, 5);

    RescoreArgs ra{};
    ra.v = t.v.ptr;
    ra.raw = t.raw.ptr;
    ra.p = t.p.ptr;
    ra.inv_order = t.inv_order.ptr;
    ra.qsorted = qs;
    ra.qperm = perm;
    ra.qrow = qrow;
    ra.cnt = cnt;
    ra.list = list;
    ra.n = n;
    ra.nq = nq;
    ra.d = t.d;
    ra.m = t.m;
    ra.kernel = t.kernel;
    ra.pred = predecessor ? 1 : 0;
    ra.k = k;
    ra.var = t.var;
    for (int c = 0; c < t.d; c++) {
        const double il = 1.0 / (double)t.ls[c];
        ra.inv_ls_hi[c] = (float)il;
        ra.inv_ls_lo[c] = (float)(il - (double)(float)il);
    }
    ra.out_idx = d_idx;
    ra.out_dist = d_dist;
    ra.out_err = d_err;
    ra.counter = ctr;
    const int64_t waves = nq < 4096 ? nq : 4096;
    rescore_kernel<<<(unsigned)waves, 64, 0, st>>>(ra);
    CX_HIP(hipGetLastError());
    prof_mark(t, 6);
}

} // namespace cx
