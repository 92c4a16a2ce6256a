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
constexpr int kScreenThreads = 512; // eight wavefronts

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
    const uint16_t *vhi, *vlo; // npad x mp, dataset order, zero rows past n
    const float *x;     // n x d raw coords, dataset order
    const float *isp;   // n: 1/sqrt(p), relative error <= 2^-23
    const float *vnorm; // n
    const int64_t *qsorted;  // sorted query dataset indices (nq)
    const int32_t *qrow;     // dataset row used for each sorted query
    int32_t n;
    int64_t nq;
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

// Kernel value in FP32 from the hardware exp2/sqrt (about 1 ulp each); its
// error is far below eK = var * 2^-18.
__device__ __forceinline__ float fast_kernel(float d2, float var, int32_t kernel)
{
    constexpr float kLog2e = 1.44269504f;
    if (kernel == 0)
        return var * __builtin_amdgcn_exp2f(d2 * (-0.5f * kLog2e));
    const float a = __builtin_amdgcn_sqrtf(5.0f * d2);
    const float poly = __builtin_fmaf(a, __builtin_fmaf(a, 0.333333343f, 1.0f), 1.0f);
    return var * poly * __builtin_amdgcn_exp2f(-a * kLog2e);
}

// One workgroup = 128 candidates (MFMA rows) x 128 queries (MFMA columns) per
// candidate tile, looping over a 2048-candidate segment.  Eight wavefronts:
// wavefront w owns candidate rows 32*(w&3).. and query columns 64*(w>>2)..,
// i.e. two 32x32 output blocks, each with a hi.hi and a cross accumulator.
template <int PASS>
__global__ void __launch_bounds__(kScreenThreads, 4) screen_kernel(ScreenArgs A)
{
    __shared__ __attribute__((aligned(16))) uint16_t sAh[kCT * kKP];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[kCT * kKP];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[kQT * kKP];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[kQT * kKP];
    __shared__ __attribute__((aligned(16))) float sCw[kCT];
    __shared__ __attribute__((aligned(16))) float sCn[kCT];
    __shared__ int32_t sQrow[kQT];
    // pass-1 histogram: two 16-bit bins per word (at most kSeg counts per bin)
    __shared__ uint32_t sHist[PASS == 1 ? kQT * kBins / 2 : 1];
    // coordinates, structure of arrays: candidates [d][kCT], then queries [d][kQT]
    extern __shared__ __attribute__((aligned(16))) float sDyn[];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int cb = wid & 3, qh = wid >> 2;
    // blockIdx.x = query tile, blockIdx.y = candidate segment; heavy tiles first
    const int32_t tile = A.n_tiles - 1 - (int32_t)blockIdx.x;
    const int64_t s0 = (int64_t)tile * kQT;
    const int32_t n = A.n, d = A.d, mp = A.mp;
    int32_t c_end = n;
    if (A.pred) {
        // candidate range of this tile: [0, largest query index) under the mask
        const int64_t last = s0 + kQT - 1 < A.nq - 1 ? s0 + kQT - 1 : A.nq - 1;
        const int64_t qmax = A.qsorted[last];
        c_end = qmax < 0 ? n : (qmax < n ? (int32_t)qmax : n);
    }
    const int32_t c_begin = (int32_t)blockIdx.y * (int32_t)kSeg;
    if (c_begin >= c_end)
        return;
    c_end = c_end < c_begin + (int32_t)kSeg ? c_end : c_begin + (int32_t)kSeg;
    float *sCx = sDyn, *sQx = sDyn + d * kCT;

    for (int i = tid; i < kQT; i += kScreenThreads) {
        const int64_t s = s0 + i;
        const int32_t row = s < A.nq ? A.qrow[s] : 0;
        sQrow[i] = row;
        for (int c = 0; c < d; c++)
            sQx[c * kQT + i] = A.x[(int64_t)row * d + c];
    }
    if (PASS == 1)
        for (int i = tid; i < kQT * kBins / 2; i += kScreenThreads)
            sHist[i] = 0u;

    // this lane's two query columns: mask limit, own row, weight, norm term,
    // and (pass 2) the emission threshold on the |rho| upper bound
    int32_t qcol[2], qlim[2], qself[2];
    float qw[2], qcn[2], qthr[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
        qcol[b] = 64 * qh + 32 * b + (lane & 31);
        const int64_t s = s0 + qcol[b];
        const bool live = s < A.nq;
        const int64_t qid = live ? A.qsorted[s] : 0;
        const int32_t row = live ? A.qrow[s] : 0;
        qlim[b] = !live ? 0 : (A.pred ? (qid < 0 ? n : (qid < n ? (int32_t)qid : n)) : n);
        qself[b] = live ? row : -1;
        qw[b] = A.isp[row];
        qcn[b] = A.cD * A.vnorm[row];
        qthr[b] = 2.0f;
        if (PASS == 2 && live) {
            const float t = A.tau[s];
            // |rho| >= 1 - tau is implied by d^2 <= tau; 2^-22 covers rounding
            qthr[b] = t >= 2.0f ? -1.0f : (1.0f - t) - 2.3841858e-7f;
        }
    }

    const int rbase = 32 * cb + 4 * (lane >> 5);
    for (int32_t c0 = c_begin; c0 < c_end; c0 += kCT) {
        __syncthreads(); // the previous tile's epilogue still reads sCx/sCw/sCn
        for (int i = tid; i < kCT; i += kScreenThreads) {
            const int32_t j = c0 + i;
            const int32_t jr = j < n ? j : 0;
            for (int c = 0; c < d; c++)
                sCx[c * kCT + i] = A.x[(int64_t)jr * d + c];
            sCw[i] = A.isp[jr];
            sCn[i] = A.vnorm[jr];
        }
        floatx16 acc_h[2], acc_x[2];
#pragma unroll
        for (int b = 0; b < 2; b++) {
            acc_h[b] = floatx16{};
            acc_x[b] = floatx16{};
        }
        for (int kc = 0; kc < mp; kc += kKC) {
            __syncthreads();
            // stage 128 candidate + 128 query rows x 32 bf16, hi and lo:
            // 2048 segments of 16 B, four per thread
            uint4 val[4];
#pragma unroll
            for (int it = 0; it < 4; it++) {
                const int idx = tid + kScreenThreads * it;
                const int side = idx >> 10, part = (idx >> 9) & 1;
                const int row = (idx >> 2) & 127, seg = idx & 3;
                const int64_t vrow = side ? (int64_t)sQrow[row] : (int64_t)(c0 + row);
                const uint16_t *src = (part ? A.vlo : A.vhi) + vrow * mp + kc + seg * 8;
                val[it] = *(const uint4 *)src;
            }
#pragma unroll
            for (int it = 0; it < 4; it++) {
                const int idx = tid + kScreenThreads * it;
                const int side = idx >> 10, part = (idx >> 9) & 1;
                const int row = (idx >> 2) & 127, seg = idx & 3;
                uint16_t *dst = (side ? (part ? sBl : sBh) : (part ? sAl : sAh)) + row * kKP +
                                seg * 8;
                *(uint4 *)dst = val[it];
            }
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < kKC; ks += 16) {
                const int koff = ks + 8 * (lane >> 5);
                const int ar = 32 * cb + (lane & 31);
                const bf16x8 ah = __builtin_bit_cast(bf16x8, *(const uint4 *)(sAh + ar * kKP + koff));
                const bf16x8 al = __builtin_bit_cast(bf16x8, *(const uint4 *)(sAl + ar * kKP + koff));
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    const int br = qcol[b];
                    const bf16x8 bh =
                        __builtin_bit_cast(bf16x8, *(const uint4 *)(sBh + br * kKP + koff));
                    const bf16x8 bl =
                        __builtin_bit_cast(bf16x8, *(const uint4 *)(sBl + br * kKP + koff));
                    acc_h[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc_h[b], 0, 0, 0);
                    acc_x[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc_x[b], 0, 0, 0);
                    acc_x[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc_x[b], 0, 0, 0);
                }
            }
        }
        // epilogue: output register r of block b is candidate row
        // rbase + 8*(r/4) + r%4, query column qcol[b]
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int rg = rbase + 8 * g;
            const float4 cw = *(const float4 *)(sCw + rg);
            const float4 cn = *(const float4 *)(sCn + rg);
            float d2[2][4];
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int rr = 0; rr < 4; rr++)
                    d2[b][rr] = 0.0f;
            for (int c = 0; c < d; c++) {
                const float4 cx = *(const float4 *)(sCx + c * kCT + rg);
                const float il = A.inv_ls[c];
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    const float qx = sQx[c * kQT + qcol[b]];
                    const float t0 = (qx - cx.x) * il, t1 = (qx - cx.y) * il;
                    const float t2 = (qx - cx.z) * il, t3 = (qx - cx.w) * il;
                    d2[b][0] = __builtin_fmaf(t0, t0, d2[b][0]);
                    d2[b][1] = __builtin_fmaf(t1, t1, d2[b][1]);
                    d2[b][2] = __builtin_fmaf(t2, t2, d2[b][2]);
                    d2[b][3] = __builtin_fmaf(t3, t3, d2[b][3]);
                }
            }
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int r = 4 * g + rr;
                    const int32_t j = c0 + rg + rr;
                    const float cwv = rr == 0 ? cw.x : rr == 1 ? cw.y : rr == 2 ? cw.z : cw.w;
                    const float cnv = rr == 0 ? cn.x : rr == 1 ? cn.y : rr == 2 ? cn.z : cn.w;
                    const float kv = fast_kernel(d2[b][rr], A.var, A.kernel);
                    const float num = kv - (acc_h[b][r] + acc_x[b][r]);
                    const float an = __builtin_fabsf(num);
                    // 1/sqrt(max(p_q p_j, 1e-18)) from per-row 1/sqrt(p)
                    const float w = __builtin_fminf(qw[b] * cwv, 1e9f);
                    const float ar = an * w;
                    // certified |rho| error: kernel, dot (cD |V_q||V_j|), the
                    // numerator's rounding, the weight's and the products'
                    const float e1 = __builtin_fmaf(an, 1.2e-7f, __builtin_fmaf(qcn[b], cnv, A.eK));
                    const float erho = __builtin_fmaf(ar, 6e-7f, e1 * (w * 1.000002f));
                    const bool valid = j < qlim[b];
                    const bool self = j == qself[b];
                    if (PASS == 1) {
                        // squared upper bound 1 - (|rho| lower bound), rounded up
                        float ub2 = (1.0f - (ar - erho)) + 1.5e-7f;
                        ub2 = self ? 0.0f : ub2;
                        if (valid && ub2 < kHistCap) {
                            const int bin = (int)(ub2 * (float)kBins);
                            atomicAdd(&sHist[(qcol[b] * kBins + bin) >> 1],
                                      (bin & 1) ? 65536u : 1u);
                        }
                    } else {
                        if (valid && (ar + erho >= qthr[b] || self)) {
                            const int64_t s = s0 + qcol[b];
                            const int32_t pos = atomicAdd(&A.cnt[s], 1);
                            if (pos < kCap)
                                A.list[s * kCap + pos] = j;
                        }
                    }
                }
        }
    }
    if (PASS == 1) {
        __syncthreads();
        for (int i = tid; i < kQT * kBins / 2; i += kScreenThreads) {
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
        const int64_t pos_out = A.qperm[s];
        const int32_t qrow = A.qrow[s];
        const int64_t qt = A.inv_order[qrow];
        const int32_t c = A.cnt[s];
        if (c > kCap)
            continue; // overflowed list: scored by the hard-query kernels
        for (int j = lane; j < k; j += 64) {
            A.out_idx[pos_out * k + j] = -1;
            A.out_dist[pos_out * k + j] = 3.40282347e38f;
            if (A.out_err)
                A.out_err[pos_out * k + j] = 0.0f;
        }
        const int64_t total = c;
        int count = 0;
        float maxh = -1.0f, maxl = 0.0f;
        int32_t maxj = -1;
        int maxpos = 0;
        for (int64_t base = 0; base < total; base += 64) {
            const int64_t e = base + lane;
            const bool live = e < total;
            int64_t j = 0;
            if (live)
                j = A.list[s * kCap + e];
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

// ---- hard queries ------------------------------------------------------------
// A query whose list overflowed (its screen interval was too wide, e.g. tiny
// p) is scored over all its valid candidates: one workgroup per (query,
// 4096-candidate segment) keeps a running top-k of compensated distances in
// LDS; a second kernel merges the per-segment lists.

constexpr int kHardChunk = 1024;             // candidates per sort round
constexpr int kHardSeg = 4 * kHardChunk;     // candidates per workgroup
constexpr int kHardSort = 2048;              // running top-k + one chunk, padded
constexpr int kHardBatch = 2048;             // hard queries per launch pair

struct HardEntry {
    float dh, dl, err;
    int32_t j;
};

struct HardArgs {
    RescoreArgs r;
    const int32_t *hard; // sorted positions of the hard queries
    int32_t h0, hn;      // this batch: hard[h0 .. h0 + hn)
    int32_t nseg;        // segments per query
    HardEntry *part;     // hn x nseg x k
    int32_t *counter;
};

__device__ __forceinline__ bool entry_lt(float ah, float al, int32_t aj, float bh, float bl,
                                         int32_t bj)
{
    return ah < bh || (ah == bh && (al < bl || (al == bl && aj < bj)));
}

// ascending bitonic sort of kHardSort (distance hi, lo, index) entries in LDS
__device__ void sort_entries(float *h, float *l, float *e, int32_t *jv)
{
    for (int size = 2; size <= kHardSort; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = threadIdx.x; t < kHardSort / 2; t += blockDim.x) {
                const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const bool sw = up ? entry_lt(h[hi], l[hi], jv[hi], h[lo], l[lo], jv[lo])
                                   : entry_lt(h[lo], l[lo], jv[lo], h[hi], l[hi], jv[hi]);
                if (sw) {
                    float a = h[lo]; h[lo] = h[hi]; h[hi] = a;
                    a = l[lo]; l[lo] = l[hi]; l[hi] = a;
                    a = e[lo]; e[lo] = e[hi]; e[hi] = a;
                    const int32_t b = jv[lo]; jv[lo] = jv[hi]; jv[hi] = b;
                }
            }
        }
    __syncthreads();
}

__device__ __forceinline__ void clear_entries(float *h, float *l, float *e, int32_t *jv, int from)
{
    for (int t = from + (int)threadIdx.x; t < kHardSort; t += blockDim.x) {
        h[t] = __builtin_huge_valf();
        l[t] = 0.0f;
        e[t] = 0.0f;
        jv[t] = 0x7fffffff;
    }
}

__device__ __forceinline__ int64_t vecchia_limit(const RescoreArgs &A, int64_t qid)
{
    // Vecchia bound on the original index (negative -> unbounded)
    return A.pred ? (qid < 0 ? A.n : (qid < A.n ? qid : A.n)) : A.n;
}

__global__ void hard_collect_kernel(const int32_t *__restrict__ cnt, int64_t nq, int32_t *hard,
                                    int32_t *n_hard)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nq && cnt[s] > kCap)
        hard[atomicAdd(n_hard, 1)] = (int32_t)s;
}

__global__ void __launch_bounds__(256) hard_score_kernel(HardArgs H)
{
    __shared__ float sh[kHardSort], sl[kHardSort], se[kHardSort];
    __shared__ int32_t sj[kHardSort];
    __shared__ int32_t s_item;
    const RescoreArgs &A = H.r;
    const int32_t k = A.k;
    for (;;) {
        if (threadIdx.x == 0)
            s_item = atomicAdd(H.counter, 1);
        __syncthreads();
        const int32_t item = s_item;
        __syncthreads();
        if (item >= H.hn * H.nseg)
            break;
        const int32_t hl = item / H.nseg, seg = item % H.nseg;
        const int64_t s = H.hard[H.h0 + hl];
        const int32_t qrow = A.qrow[s];
        const int64_t qt = A.inv_order[qrow];
        const int64_t limit = vecchia_limit(A, A.qsorted[s]);
        const int64_t beg = (int64_t)seg * kHardSeg;
        if (beg >= limit)
            continue; // the merge reads only segments below the limit
        const int64_t end = beg + kHardSeg < limit ? beg + kHardSeg : limit;
        clear_entries(sh, sl, se, sj, 0);
        for (int64_t c0 = beg; c0 < end; c0 += kHardChunk) {
            __syncthreads();
            // the running top-k sits in [0, k); this chunk fills [k, k + chunk)
            for (int t = threadIdx.x; t < kHardChunk; t += blockDim.x) {
                const int64_t j = c0 + t;
                float dh = __builtin_huge_valf(), dl = 0.0f, er = 0.0f;
                int32_t jj = 0x7fffffff;
                if (j < end) {
                    jj = (int32_t)j;
                    if (j == qrow) {
                        dh = 0.0f;
                    } else {
                        const df dd = accurate_distance(A, qt, A.inv_order[j], &er);
                        dh = dd.hi;
                        dl = dd.lo;
                    }
                }
                sh[k + t] = dh;
                sl[k + t] = dl;
                se[k + t] = er;
                sj[k + t] = jj;
            }
            sort_entries(sh, sl, se, sj);
        }
        HardEntry *out = H.part + ((int64_t)hl * H.nseg + seg) * k;
        for (int t = threadIdx.x; t < k; t += blockDim.x)
            out[t] = HardEntry{sh[t], sl[t], se[t], sj[t]};
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) hard_merge_kernel(HardArgs H)
{
    __shared__ float sh[kHardSort], sl[kHardSort], se[kHardSort];
    __shared__ int32_t sj[kHardSort];
    const RescoreArgs &A = H.r;
    const int32_t k = A.k;
    for (int32_t hl = blockIdx.x; hl < H.hn; hl += gridDim.x) {
        const int64_t s = H.hard[H.h0 + hl];
        const int64_t limit = vecchia_limit(A, A.qsorted[s]);
        const int32_t nseg = (int32_t)((limit + kHardSeg - 1) / kHardSeg);
        const int32_t per = (kHardSort - k) / k; // lists merged per round
        __syncthreads();
        clear_entries(sh, sl, se, sj, 0);
        for (int32_t g = 0; g < nseg; g += per) {
            const int32_t nl = nseg - g < per ? nseg - g : per;
            __syncthreads();
            // entries left past k + nl*k by an earlier round rank behind the
            // current top-k, so they never re-enter it
            for (int t = threadIdx.x; t < nl * k; t += blockDim.x) {
                const HardEntry e = H.part[((int64_t)hl * H.nseg + g + t / k) * k + t % k];
                sh[k + t] = e.dh;
                sl[k + t] = e.dl;
                se[k + t] = e.err;
                sj[k + t] = e.j;
            }
            sort_entries(sh, sl, se, sj);
        }
        const int64_t pos_out = A.qperm[s];
        for (int t = threadIdx.x; t < k; t += blockDim.x) {
            const bool ok = sh[t] < __builtin_huge_valf();
            A.out_idx[pos_out * k + t] = ok ? (int64_t)sj[t] : -1;
            A.out_dist[pos_out * k + t] = ok ? sh[t] + sl[t] : 3.40282347e38f;
            if (A.out_err)
                A.out_err[pos_out * k + t] = ok ? se[t] : 0.0f;
        }
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
    CX_REQUIRE(t.n < ((int64_t)1 << 31) - 4 * kSeg,
               "mixed-precision kNN supports fewer than 2^31 - 8192 points");
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
    const size_t o_hard = off;
    off += align(sizeof(int32_t) * nq);
    const size_t o_part = off;
    const int64_t hard_cap = nq < kHardBatch ? nq : kHardBatch;
    off += align(sizeof(HardEntry) * (size_t)hard_cap * ((n + kHardSeg - 1) / kHardSeg) * k);
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
    int32_t *hard = (int32_t *)(ws + o_hard);
    HardEntry *part = (HardEntry *)(ws + o_part);
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
    CX_HIP(hipMemsetAsync(ctr, 0, 256, st));

    ScreenArgs sa{};
    sa.vhi = t.vhi.ptr;
    sa.vlo = t.vlo.ptr;
    sa.x = t.x_ds.ptr;
    sa.isp = t.isp_ds.ptr;
    sa.vnorm = t.vnorm.ptr;
    sa.qsorted = qs;
    sa.qrow = qrow;
    sa.n = (int32_t)n;
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
    const size_t dyn_lds = sizeof(float) * 2 * (size_t)t.d * kCT;
    screen_kernel<1><<<grid, kScreenThreads, dyn_lds, st>>>(sa);
    CX_HIP(hipGetLastError());
    prof_mark(t, 3);
    tau_kernel<<<qb, 256, 0, st>>>(hist, nq, k, tau);
    prof_mark(t, 4);
    screen_kernel<2><<<grid, kScreenThreads, dyn_lds, st>>>(sa);
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
    // queries whose list overflowed go to the hard-query kernels
    hard_collect_kernel<<<qb, 256, 0, st>>>(cnt, nq, hard, ctr + 1);
    if (!t.host_word)
        CX_HIP(hipHostMalloc((void **)&t.host_word, 64, hipHostMallocDefault));
    CX_HIP(hipMemcpyAsync(t.host_word, ctr + 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    const int64_t waves = nq < 4096 ? nq : 4096;
    rescore_kernel<<<(unsigned)waves, 64, 0, st>>>(ra);
    CX_HIP(hipGetLastError());
    prof_mark(t, 6);
    CX_HIP(hipStreamSynchronize(st));
    const int32_t n_hard = *t.host_word;
    HardArgs ha{};
    ha.r = ra;
    ha.hard = hard;
    ha.nseg = (int32_t)((n + kHardSeg - 1) / kHardSeg);
    ha.part = part;
    for (int32_t h0 = 0; h0 < n_hard; h0 += kHardBatch) {
        ha.h0 = h0;
        ha.hn = n_hard - h0 < kHardBatch ? n_hard - h0 : kHardBatch;
        ha.counter = ctr + 2;
        if (h0 > 0)
            CX_HIP(hipMemsetAsync(ha.counter, 0, sizeof(int32_t), st));
        hard_score_kernel<<<1024, 256, 0, st>>>(ha);
        CX_HIP(hipGetLastError());
        const int64_t blocks = ha.hn < 2048 ? ha.hn : 2048;
        hard_merge_kernel<<<(unsigned)blocks, 256, 0, st>>>(ha);
        CX_HIP(hipGetLastError());
    }
    prof_mark(t, 7);
    t.last_hard = n_hard;
}

} // namespace cx
