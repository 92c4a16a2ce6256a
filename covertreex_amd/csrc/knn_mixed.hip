// Mixed-precision batched residual kNN (CX_KNN_MIXED): exact neighbours under
// d = sqrt(1 - |rho|) evaluated on the stored inputs, with a certified screen.
//
// Work is laid out as query tiles x candidate tiles in dataset order, so the
// Vecchia mask j < i (src/algo.rs:918-925 / src/lib.rs:444-448) becomes a
// triangle: a query tile only visits candidates below its largest index.
//
//   prep      V -> dataset-order BF16 hi/lo split (V = hi + lo + r, |r| <=
//             2^-18 |V|), row norms |V_i| rounded up.
//   screen    one workgroup per 128-query tile x 2048-candidate segment, eight
//             wavefronts of 32 candidates x 64 queries; V.V on MFMA
//             (v_mfma_f32_32x32x16_bf16) as hi.hi + (hi.lo + lo.hi) in two
//             FP32 accumulators, kernel value in FP32, and a certified
//             interval [lb2, ub2] containing the exact squared distance
//             1 - |rho|.
//             pass 1: per-query histogram of ub2 -> tau_q, an upper bound on
//                     the k-th exact squared distance;
//             pass 2: every candidate with lb2 <= tau_q is appended to the
//                     query's list (a superset of the exact k nearest).
//   rescore   one workgroup per query: compensated FP32 (double-float)
//             distance for each listed candidate, exact top-k, ascending
//             output with a per-result error bound.  A query whose list
//             overflowed its kCap slots is listed again, in full, into a pool
//             sized by its pass-2 count, and rescored from there.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

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

struct MetaScale {
    float il[kMaxD];
};

// dataset-order coordinates scaled by 1/lengthscale, and 1/sqrt(p)
__global__ void gather_meta_kernel(const float *__restrict__ raw, const float *__restrict__ p,
                                   const int64_t *__restrict__ inv_order, int64_t n, int32_t d,
                                   MetaScale sc, float *__restrict__ x_ds,
                                   float *__restrict__ isp_ds)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t t = inv_order[i];
    for (int32_t c = 0; c < d; c++)
        x_ds[i * d + c] = raw[t * d + c] * sc.il[c];
    // correctly rounded sqrt and division: relative error <= 2^-23
    isp_ds[i] = 1.0f / __builtin_sqrtf(p[t]);
}

// ---- screen ------------------------------------------------------------------

struct ScreenArgs {
    const uint16_t *vhi, *vlo; // npad x mp, dataset order, zero rows past n
    const float *x;     // n x d coords / lengthscale, dataset order
    const float *isp;   // n: 1/sqrt(p), relative error <= 2^-23
    const float *vnorm; // n
    const int64_t *qsorted;  // sorted query dataset indices (nq)
    const int32_t *qrow;     // dataset row used for each sorted query
    int32_t n;
    int64_t nq;
    int32_t d, mp, kernel, pred, k;
    float var;
    float cD, eK;
    float kA, kB, kC;  // folded kernel constants (fast_kernel)
    float *tau;      // nq
    int32_t *hist;   // nq x kBins
    int32_t *cnt;    // nq
    int32_t *list;   // nq x kCap
    // this launch covers query tiles [tile0, tile0 + n_tiles) x all segments;
    // codes hold one byte per (candidate, query) pair of those tiles
    int32_t tile0, n_tiles, n_seg;
    uint32_t *codes;
};

// Kernel value in FP32 from the hardware exp2/sqrt (about 1 ulp each) with
// the constants folded: RBF var * 2^(kA d2), kA = -log2(e)/2; Matern-5/2 with
// a' = sqrt(kA d2) = log2(e) sqrt(5 d2): (var + kB a' + kC a'^2) 2^-a'.  The
// error stays far below eK.
template <int KT> __device__ __forceinline__ float fast_kernel(float d2, const ScreenArgs &A)
{
    if (KT == 0)
        return A.var * __builtin_amdgcn_exp2f(d2 * A.kA);
    const float a = __builtin_amdgcn_sqrtf(d2 * A.kA);
    const float poly = __builtin_fmaf(a, __builtin_fmaf(a, A.kC, A.kB), A.var);
    return poly * __builtin_amdgcn_exp2f(-a);
}

// First word of the packed codes of (tile block, candidate tile, wavefront,
// lane): eight consecutive words per lane, word 2g + b holding candidate rows
// rbase + 8g .. + 3 (one byte each) against query column qcol[b], so a
// wavefront stores and reloads 2 KiB contiguous per candidate tile.
__device__ __forceinline__ int64_t code_word(int64_t blk, int ct, int wid, int lane)
{
    return (((blk * (kSeg / kCT) + ct) * 8 + wid) * 64 + lane) * 8;
}

// Per-lane query columns of a 128-query tile (shared by the screen and the
// filter): column, Vecchia limit on the candidate index, own dataset row.
__device__ __forceinline__ void tile_queries(const ScreenArgs &A, int64_t s0, int qh, int lane,
                                             int32_t *qcol, int32_t *qlim, int32_t *qself,
                                             int32_t *qrowv)
{
#pragma unroll
    for (int b = 0; b < 2; b++) {
        qcol[b] = 64 * qh + 32 * b + (lane & 31);
        const int64_t s = s0 + qcol[b];
        const bool live = s < A.nq;
        const int64_t qid = live ? A.qsorted[s] : 0;
        const int32_t row = live ? A.qrow[s] : 0;
        qlim[b] = !live ? 0 : (A.pred ? (qid < 0 ? A.n : (qid < A.n ? (int32_t)qid : A.n)) : A.n);
        qself[b] = live ? row : -1;
        qrowv[b] = row;
    }
}

// candidate range [c_begin, c_end) of this (tile, segment); empty if none
__device__ __forceinline__ void tile_range(const ScreenArgs &A, int64_t s0, int32_t *c_begin,
                                           int32_t *c_end)
{
    int32_t e = A.n;
    if (A.pred) {
        // [0, largest query index) under the mask
        const int64_t last = s0 + kQT - 1 < A.nq - 1 ? s0 + kQT - 1 : A.nq - 1;
        const int64_t qmax = A.qsorted[last];
        e = qmax < 0 ? A.n : (qmax < A.n ? (int32_t)qmax : A.n);
    }
    const int32_t b = (int32_t)blockIdx.y * (int32_t)kSeg;
    *c_begin = b;
    *c_end = e < b + (int32_t)kSeg ? e : b + (int32_t)kSeg;
}

// Pass 1.  One workgroup = 128 candidates (MFMA rows) x 128 queries (MFMA
// columns) per candidate tile, looping over a 2048-candidate segment.  Eight
// wavefronts: wavefront w owns candidate rows 32*(w&3).. and query columns
// 64*(w>>2).., i.e. two 32x32 output blocks, each with a hi.hi and a cross
// accumulator.  Per pair: the certified interval of |rho|; its upper end
// feeds the query's histogram, its lower end, as a squared-distance lower
// bound, is stored as a one-byte code for the filter.
// DD > 0 fixes the coordinate dimension (query coordinates then live in
// registers); DD == 0 reads them from LDS for any d <= kMaxD.  KT is the
// kernel (0 RBF, 1 Matern-5/2).  Coordinates arrive scaled by 1/lengthscale.
template <int DD, int KT>
__global__ void __launch_bounds__(kScreenThreads, 4) screen_kernel(ScreenArgs A)
{
    __shared__ __attribute__((aligned(16))) uint16_t sAh[kCT * kKP];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[kCT * kKP];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[kQT * kKP];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[kQT * kKP];
    __shared__ __attribute__((aligned(16))) float sCw[kCT];
    __shared__ __attribute__((aligned(16))) float sCn[kCT];
    __shared__ int32_t sQrow[kQT];
    // histogram: two 16-bit bins per word (at most kSeg counts per bin)
    __shared__ uint32_t sHist[kQT * kBins / 2];
    // coordinates, structure of arrays: candidates [d][kCT], then queries [d][kQT]
    extern __shared__ __attribute__((aligned(16))) float sDyn[];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int cb = wid & 3, qh = wid >> 2;
    // blockIdx.x = query tile within the launch, heavy tiles first
    const int32_t tl = A.n_tiles - 1 - (int32_t)blockIdx.x;
    const int64_t s0 = (int64_t)(A.tile0 + tl) * kQT;
    const int64_t blk = (int64_t)tl * A.n_seg + blockIdx.y;
    const int32_t mp = A.mp;
    const int32_t d = DD > 0 ? DD : A.d;
    int32_t c_begin, c_end;
    tile_range(A, s0, &c_begin, &c_end);
    if (c_begin >= c_end)
        return;
    float *sCx = sDyn, *sQx = sDyn + d * kCT;

    for (int i = tid; i < kQT; i += kScreenThreads) {
        const int64_t s = s0 + i;
        const int32_t row = s < A.nq ? A.qrow[s] : 0;
        sQrow[i] = row;
        for (int c = 0; c < d; c++)
            sQx[c * kQT + i] = A.x[(int64_t)row * d + c];
    }
    for (int i = tid; i < kQT * kBins / 2; i += kScreenThreads)
        sHist[i] = 0u;

    // this lane's two query columns: mask limit, weight 1/sqrt(p), norm term
    int32_t qcol[2], qlim[2], qself[2], qrowv[2];
    tile_queries(A, s0, qh, lane, qcol, qlim, qself, qrowv);
    float qw[2], qcn[2];
    float qx[2][DD > 0 ? DD : 1];
#pragma unroll
    for (int b = 0; b < 2; b++) {
        qw[b] = A.isp[qrowv[b]];
        qcn[b] = A.cD * 1.000002f * A.vnorm[qrowv[b]];
        if (DD > 0)
#pragma unroll
            for (int c = 0; c < (DD > 0 ? DD : 1); c++)
                qx[b][c] = A.x[(int64_t)qrowv[b] * d + c];
    }
    const float eKp = A.eK * 1.000002f;
    // a pair counts when 1 - rlo + 1.5e-7 < kHistCap; screen first on rlo
    // itself with a slightly lower threshold
    const float thr1 = 1.0f + 1.5e-7f - kHistCap - 1e-6f;

    const int rbase = 32 * cb + 4 * (lane >> 5);
    for (int32_t c0 = c_begin; c0 < c_end; c0 += kCT) {
        __syncthreads(); // the previous tile's epilogue still reads sCx/sCw/sCn
        for (int i = tid; i < kCT; i += kScreenThreads) {
            const int32_t j = c0 + i;
            const int32_t jr = j < A.n ? j : 0;
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
        const int ct = (c0 - c_begin) / kCT;
        uint32_t words[8];
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
            if (DD > 0) {
#pragma unroll
                for (int c = 0; c < (DD > 0 ? DD : 1); c++) {
                    const float4 cx = *(const float4 *)(sCx + c * kCT + rg);
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        const float t0 = qx[b][c] - cx.x, t1 = qx[b][c] - cx.y;
                        const float t2 = qx[b][c] - cx.z, t3 = qx[b][c] - cx.w;
                        d2[b][0] = __builtin_fmaf(t0, t0, d2[b][0]);
                        d2[b][1] = __builtin_fmaf(t1, t1, d2[b][1]);
                        d2[b][2] = __builtin_fmaf(t2, t2, d2[b][2]);
                        d2[b][3] = __builtin_fmaf(t3, t3, d2[b][3]);
                    }
                }
            } else {
                for (int c = 0; c < d; c++) {
                    const float4 cx = *(const float4 *)(sCx + c * kCT + rg);
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        const float q = sQx[c * kQT + qcol[b]];
                        const float t0 = q - cx.x, t1 = q - cx.y;
                        const float t2 = q - cx.z, t3 = q - cx.w;
                        d2[b][0] = __builtin_fmaf(t0, t0, d2[b][0]);
                        d2[b][1] = __builtin_fmaf(t1, t1, d2[b][1]);
                        d2[b][2] = __builtin_fmaf(t2, t2, d2[b][2]);
                        d2[b][3] = __builtin_fmaf(t3, t3, d2[b][3]);
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < 2; b++) {
                uint32_t word = 0u;
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int r = 4 * g + rr;
                    const int32_t j = c0 + rg + rr;
                    const float cwv = rr == 0 ? cw.x : rr == 1 ? cw.y : rr == 2 ? cw.z : cw.w;
                    const float cnv = rr == 0 ? cn.x : rr == 1 ? cn.y : rr == 2 ? cn.z : cn.w;
                    const float kv = fast_kernel<KT>(d2[b][rr], A);
                    const float num = kv - (acc_h[b][r] + acc_x[b][r]);
                    // 1/sqrt(max(p_q p_j, 1e-18)) from per-row 1/sqrt(p)
                    const float w = __builtin_fminf(qw[b] * cwv, 1e9f);
                    const float ar = __builtin_fabsf(num) * w;
                    // certified |rho| error: w (eK + cD |V_q||V_j|) (1 + 2e-6)
                    // for the kernel, the dot and the final sum, plus 7.2e-7 ar
                    // for the roundings of num, w and ar
                    const float wE = w * __builtin_fmaf(qcn[b], cnv, eKp);
                    const bool valid = j < qlim[b];
                    // histogram of the squared upper bound 1 - (|rho| lower
                    // bound); the pair's own row may be counted at its
                    // computed bound or not at all, tau stays an upper bound
                    const float rlo = __builtin_fmaf(ar, 1.0f - 7.2e-7f, -wE);
                    if (valid && rlo > thr1) {
                        const float ub2 = (1.0f - rlo) + 1.5e-7f;
                        if (ub2 < kHistCap) {
                            const int bin = (int)(ub2 * (float)kBins);
                            atomicAdd(&sHist[(qcol[b] * kBins + bin) >> 1],
                                      (bin & 1) ? 65536u : 1u);
                        }
                    }
                    // code = floor(256 * squared lower bound), 255 if invalid
                    const float rhi = __builtin_fmaf(ar, 1.0f + 7.2e-7f, wE);
                    const float lb2 = (1.0f - rhi) - 1.5e-7f;
                    const float cf = __builtin_fminf(__builtin_fmaxf(lb2 * 256.0f, 0.0f), 255.0f);
                    const uint32_t code = valid ? (uint32_t)cf : 255u;
                    word |= code << (8 * rr);
                }
                words[2 * g + b] = word;
            }
        }
        uint4 *cdst = (uint4 *)(A.codes + code_word(blk, ct, wid, lane));
        cdst[0] = make_uint4(words[0], words[1], words[2], words[3]);
        cdst[1] = make_uint4(words[4], words[5], words[6], words[7]);
    }
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

// Pass 2: the same (tile, segment) grid and lane layout reads the codes back
// and lists every valid candidate whose code is at most floor(256 tau_q)
// (every pair with squared lower bound <= tau_q among them), plus the query's
// own row.  One atomic per query column and wavefront: lanes l and l + 32
// share a column and split the reserved range.
__global__ void __launch_bounds__(kScreenThreads) filter_kernel(ScreenArgs A)
{
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int cb = wid & 3, qh = wid >> 2;
    const int32_t tl = A.n_tiles - 1 - (int32_t)blockIdx.x;
    const int64_t s0 = (int64_t)(A.tile0 + tl) * kQT;
    const int64_t blk = (int64_t)tl * A.n_seg + blockIdx.y;
    int32_t c_begin, c_end;
    tile_range(A, s0, &c_begin, &c_end);
    if (c_begin >= c_end)
        return;
    int32_t qcol[2], qlim[2], qself[2], qrowv[2];
    tile_queries(A, s0, qh, lane, qcol, qlim, qself, qrowv);
    uint32_t qt[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
        const int64_t s = s0 + qcol[b];
        const float t = s < A.nq ? A.tau[s] : 0.0f;
        qt[b] = t >= 2.0f ? 255u : (uint32_t)(t * 256.0f);
    }
    const int rbase = 32 * cb + 4 * (lane >> 5);
    for (int32_t c0 = c_begin; c0 < c_end; c0 += kCT) {
        const int ct = (c0 - c_begin) / kCT;
        uint32_t emit[2] = {0u, 0u};
        const uint4 *csrc = (const uint4 *)(A.codes + code_word(blk, ct, wid, lane));
        const uint4 w0 = csrc[0], w1 = csrc[1];
        const uint32_t words[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const uint32_t word = words[2 * g + b];
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const int32_t j = c0 + rbase + 8 * g + rr;
                    const uint32_t code = (word >> (8 * rr)) & 255u;
                    if (j < qlim[b] && code <= qt[b])
                        emit[b] |= 1u << (4 * g + rr);
                }
            }
#pragma unroll
        for (int b = 0; b < 2; b++) {
            // the query's own row (distance 0) is always listed
            const int32_t dr = qself[b] - c0 - rbase;
            if (dr >= 0 && dr < 28 && (dr & 7) < 4 && qself[b] < qlim[b])
                emit[b] |= 1u << (4 * (dr >> 3) + (dr & 3));
            const int c_me = __builtin_popcount(emit[b]);
            const int c_other = __shfl_xor(c_me, 32, 64);
            const int64_t s = s0 + qcol[b];
            int base = 0;
            if (lane < 32 && c_me + c_other > 0)
                base = atomicAdd(&A.cnt[s], c_me + c_other);
            base = __shfl(base, lane & 31, 64);
            int p = base + (lane < 32 ? 0 : c_other);
            uint32_t mk = emit[b];
            int32_t *dst = A.list + s * kCap;
            while (mk) {
                const int r = __builtin_ctz(mk);
                mk &= mk - 1;
                if (p < kCap)
                    dst[p] = c0 + rbase + 8 * (r >> 2) + (r & 3);
                p++;
            }
        }
    }
}

template <int KT> static void launch_screen_kt(const ScreenArgs &A, size_t lds, hipStream_t st)
{
    const dim3 grid((unsigned)A.n_tiles, (unsigned)A.n_seg);
    switch (A.d) {
    case 2:
        screen_kernel<2, KT><<<grid, kScreenThreads, lds, st>>>(A);
        break;
    case 3:
        screen_kernel<3, KT><<<grid, kScreenThreads, lds, st>>>(A);
        break;
    case 5:
        screen_kernel<5, KT><<<grid, kScreenThreads, lds, st>>>(A);
        break;
    default:
        screen_kernel<0, KT><<<grid, kScreenThreads, lds, st>>>(A);
    }
    CX_HIP(hipGetLastError());
}

static void launch_screen(const ScreenArgs &A, size_t lds, hipStream_t st)
{
    if (A.kernel == 0)
        launch_screen_kt<0>(A, lds, st);
    else
        launch_screen_kt<1>(A, lds, st);
}

// tau_q (squared scale): upper edge of the first bin where the cumulative count
// of squared upper bounds reaches k; 2.0 (keep everything) when fewer than k
// candidates have a squared upper bound below the histogram cap.
__global__ void tau_kernel(const int32_t *__restrict__ hist, int64_t s_begin, int64_t s_end,
                           int32_t k, float *tau)
{
    const int64_t s = s_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= s_end)
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

constexpr int kRsThreads = 256;                 // one workgroup per query
constexpr int kRsSort = 2048;                   // running top-k + one round, padded
constexpr int kRsChunk = kRsSort - kMixMaxK;    // listed candidates per sort round
constexpr int kHMaxM = 512;                     // query rows staged in LDS up to this rank

struct HardEntry {
    float dh, dl, err;
    int32_t j;
};

struct RescoreArgs {
    const float *v;       // n x m tree order
    const float *raw;     // n x d tree order
    const float *p;       // n tree order
    const int64_t *inv_order;
    const int64_t *qsorted;
    const int64_t *qperm; // sorted position -> original query position
    const int32_t *qrow;
    const int32_t *cnt;   // listed candidates per sorted query
    const int32_t *list;  // kCap slots per sorted query
    int64_t n, nq;
    int32_t d, m, kernel, pred, k;
    float var;
    float inv_ls_hi[kMaxD], inv_ls_lo[kMaxD];
    int64_t *out_idx;
    float *out_dist, *out_err;
    int64_t *counter;
};

// The residual distance from a compensated dot V_q.V_j (double-float) and
// sum|V_q[c] V_j[c]|: kernel value, numerator, division and square root in
// double-float.  Returns the distance and a bound on its error.
__device__ __forceinline__ df finish_distance(const RescoreArgs &A, int64_t qt, int64_t jt, df dot,
                                              float sabs, float *err)
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

__device__ __forceinline__ bool entry_lt(float ah, float al, int32_t aj, float bh, float bl,
                                         int32_t bj)
{
    return ah < bh || (ah == bh && (al < bl || (al == bl && aj < bj)));
}

// ascending bitonic sort of the first `size` (a power of two) entries
// (distance hi, lo, index; error rides along) in LDS
__device__ void sort_entries(float *h, float *l, float *e, int32_t *jv, int size_total)
{
    for (int size = 2; size <= size_total; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = threadIdx.x; t < size_total / 2; t += blockDim.x) {
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

// Compensated dot of V rows q and j over this lane's columns (Dot2 within a
// lane: TwoProduct + TwoSum per term), eight lanes per row pair.  Rows are
// float4-aligned when m % 4 == 0; the lanes of a group then read 128
// contiguous bytes per load.
__device__ __forceinline__ void dot2_step(float a, float b, float &s, float &c, float &sabs)
{
    const float ph = a * b;
    const float pl = __builtin_fmaf(a, b, -ph);
    const float t = s + ph;
    const float bb = t - s;
    c += ((s - (t - bb)) + (ph - bb)) + pl;
    s = t;
    sabs += __builtin_fabsf(ph);
}

__device__ __forceinline__ void group_dot(const float *__restrict__ vq,
                                          const float *__restrict__ vj, int32_t m, int sub,
                                          df &dot, float &sabs)
{
    float s = 0.0f, c = 0.0f;
    sabs = 0.0f;
    if ((m & 3) == 0) {
#pragma unroll 4
        for (int32_t i = 4 * sub; i < m; i += 32) {
            const float4 a = *(const float4 *)(vq + i);
            const float4 b = *(const float4 *)(vj + i);
            dot2_step(a.x, b.x, s, c, sabs);
            dot2_step(a.y, b.y, s, c, sabs);
            dot2_step(a.z, b.z, s, c, sabs);
            dot2_step(a.w, b.w, s, c, sabs);
        }
    } else {
#pragma unroll 4
        for (int32_t i = sub; i < m; i += 8)
            dot2_step(vq[i], vj[i], s, c, sabs);
    }
    dot = two_sum(s, c);
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        const df other = {__shfl_xor(dot.hi, o, 64), __shfl_xor(dot.lo, o, 64)};
        dot = df_add(dot, other);
        sabs += __shfl_xor(sabs, o, 64);
    }
}

// Score one list slice into LDS slots [k, k + n_in) and merge it into the
// running top-k at [0, k): eight lanes per candidate for the dot (results
// parked in the slots), then one lane per candidate for the distance.
__device__ void score_slice(const RescoreArgs &A, const int32_t *__restrict__ lst, int32_t n_in,
                            int32_t qrow, int64_t qt, const float *vq, float *sh, float *sl,
                            float *se, int32_t *sj)
{
    const int tid = threadIdx.x, grp = tid >> 3, sub = tid & 7;
    const int32_t k = A.k, m = A.m;
    __syncthreads();
    for (int32_t e = grp; e < n_in; e += kRsThreads / 8) {
        const int32_t j = lst[e];
        df dot;
        float sabs;
        group_dot(vq, A.v + A.inv_order[j] * m, m, sub, dot, sabs);
        if (sub == 0) {
            sh[k + e] = dot.hi;
            sl[k + e] = dot.lo;
            se[k + e] = sabs;
            sj[k + e] = j;
        }
    }
    __syncthreads();
    for (int32_t e = tid; e < n_in; e += kRsThreads) {
        const int32_t j = sj[k + e];
        float dh = 0.0f, dl = 0.0f, er = 0.0f;
        if (j != qrow) {
            const df dd = finish_distance(A, qt, A.inv_order[j], df{sh[k + e], sl[k + e]},
                                          se[k + e], &er);
            dh = dd.hi;
            dl = dd.lo;
        }
        sh[k + e] = dh;
        sl[k + e] = dl;
        se[k + e] = er;
    }
    int p2 = 2;
    while (p2 < k + n_in)
        p2 <<= 1;
    // entries past k left by an earlier round rank behind the current top-k,
    // so they never re-enter it
    sort_entries(sh, sl, se, sj, p2);
}

__device__ __forceinline__ void clear_entries(float *sh, float *sl, float *se, int32_t *sj)
{
    for (int t = threadIdx.x; t < kRsSort; t += kRsThreads) {
        sh[t] = __builtin_huge_valf();
        sl[t] = 0.0f;
        se[t] = 0.0f;
        sj[t] = 0x7fffffff;
    }
}

__device__ __forceinline__ void write_row(const RescoreArgs &A, int64_t pos_out, const float *sh,
                                          const float *sl, const float *se, const int32_t *sj)
{
    const int32_t k = A.k;
    for (int t = threadIdx.x; t < k; t += kRsThreads) {
        const bool ok = sh[t] < __builtin_huge_valf();
        A.out_idx[pos_out * k + t] = ok ? (int64_t)sj[t] : -1;
        A.out_dist[pos_out * k + t] = ok ? sh[t] + sl[t] : 3.40282347e38f;
        if (A.out_err)
            A.out_err[pos_out * k + t] = ok ? se[t] : 0.0f;
    }
}

// One workgroup per query, its kCap-slot list scored in rounds of kRsChunk,
// output row written.  Queries whose list overflowed are left to the hard
// kernels.
__global__ void __launch_bounds__(kRsThreads) rescore_kernel(RescoreArgs A)
{
    __shared__ float sh[kRsSort], sl[kRsSort], se[kRsSort];
    __shared__ int32_t sj[kRsSort];
    __shared__ int64_t s_item;
    __shared__ __attribute__((aligned(16))) float s_vq[kHMaxM];
    const int32_t m = A.m;
    for (;;) {
        if (threadIdx.x == 0)
            s_item = (int64_t)atomicAdd((unsigned long long *)A.counter, 1ull);
        __syncthreads();
        const int64_t s = s_item;
        __syncthreads();
        if (s >= A.nq)
            break;
        const int32_t count = A.cnt[s];
        if (count > kCap)
            continue; // scored over all its candidates by the hard kernels
        const int32_t qrow = A.qrow[s];
        const int64_t qt = A.inv_order[qrow];
        // the query row in LDS when it fits (score_slice syncs before use)
        const float *vq = A.v + qt * m;
        if (m <= kHMaxM) {
            for (int c = threadIdx.x; c < m; c += kRsThreads)
                s_vq[c] = vq[c];
            vq = s_vq;
        }
        clear_entries(sh, sl, se, sj);
        for (int32_t c0 = 0; c0 < count; c0 += kRsChunk) {
            const int32_t n_in = count - c0 < kRsChunk ? count - c0 : kRsChunk;
            score_slice(A, A.list + s * kCap + c0, n_in, qrow, qt, vq, sh, sl, se, sj);
        }
        write_row(A, A.qperm[s], sh, sl, se, sj);
        __syncthreads();
    }
}

// ---- hard queries ------------------------------------------------------------
// A query whose list overflowed kCap (a wide screen interval, e.g. tiny p) is
// scored over all its valid candidates.  One workgroup = kHQ consecutive hard
// queries x a kHSeg-candidate segment: every candidate row is read once for
// all kHQ queries (eight lanes per candidate, kHQ compensated dots per lane
// group), distances are finished one lane per pair, and each query keeps a
// running top-k in LDS.  A merge kernel combines the segments.

constexpr int kHQ = 4;        // hard queries per workgroup
constexpr int kHSeg = 1024;   // candidates per workgroup
constexpr int kHRound = 128;  // candidates per sort round
constexpr int kHSort = 512;   // running top-k (k <= 256) + one round, padded

struct HardArgs {
    RescoreArgs r;
    const int32_t *hard;  // sorted positions of the hard queries
    int32_t n_hard;
    int32_t g0, gn;       // this batch: query groups [g0, g0 + gn)
    int32_t nseg;         // segments per group (ceil(n / kHSeg))
    HardEntry *part;      // gn x nseg x kHQ x k segment top-k lists
    int64_t *counter;
};

// ascending bitonic sort of the first `size` entries of each of `na` arrays
// laid out [a][kHSort]
__device__ void sort_arrays(float *h, float *l, float *e, int32_t *jv, int na, int size_total)
{
    const int half = size_total / 2;
    for (int size = 2; size <= size_total; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = threadIdx.x; t < na * half; t += blockDim.x) {
                const int a = t / half, tt = t - a * half;
                const int lo = a * kHSort + 2 * tt - (tt & (stride - 1)), hi = lo + stride;
                const bool up = ((lo - a * kHSort) & size) == 0;
                const bool sw = up ? entry_lt(h[hi], l[hi], jv[hi], h[lo], l[lo], jv[lo])
                                   : entry_lt(h[lo], l[lo], jv[lo], h[hi], l[hi], jv[hi]);
                if (sw) {
                    float x = h[lo]; h[lo] = h[hi]; h[hi] = x;
                    x = l[lo]; l[lo] = l[hi]; l[hi] = x;
                    x = e[lo]; e[lo] = e[hi]; e[hi] = x;
                    const int32_t y = jv[lo]; jv[lo] = jv[hi]; jv[hi] = y;
                }
            }
        }
    __syncthreads();
}

__device__ __forceinline__ int64_t hard_limit(const RescoreArgs &A, int64_t s)
{
    const int64_t qid = A.qsorted[s];
    // Vecchia bound on the original index (negative -> unbounded)
    return A.pred ? (qid < 0 ? A.n : (qid < A.n ? qid : A.n)) : A.n;
}

__global__ void __launch_bounds__(kRsThreads) hard_tile_kernel(HardArgs H)
{
    __shared__ float sh[kHQ * kHSort], sl[kHQ * kHSort], se[kHQ * kHSort];
    __shared__ int32_t sj[kHQ * kHSort];
    __shared__ int64_t s_item;
    __shared__ int64_t s_lim[kHQ], s_qt[kHQ];
    __shared__ int32_t s_qrow[kHQ];
    __shared__ __attribute__((aligned(16))) float s_vq[kHQ * kHMaxM];
    const RescoreArgs &A = H.r;
    const int tid = threadIdx.x, grp = tid >> 3, sub = tid & 7;
    const int32_t k = A.k, m = A.m;
    // float4-aligned rows that fit LDS are staged
    const bool staged = (m & 3) == 0 && m <= kHMaxM;
    for (;;) {
        if (tid == 0)
            s_item = (int64_t)atomicAdd((unsigned long long *)H.counter, 1ull);
        __syncthreads();
        const int64_t item = s_item;
        __syncthreads();
        if (item >= (int64_t)H.gn * H.nseg)
            break;
        const int32_t gl = (int32_t)(item / H.nseg), seg = (int32_t)(item % H.nseg);
        const int32_t h0 = (H.g0 + gl) * kHQ;
        const int nq_g = H.n_hard - h0 < kHQ ? H.n_hard - h0 : kHQ;
        if (tid < kHQ) {
            int64_t lim = 0, qtv = 0;
            int32_t row = -1;
            if (tid < nq_g) {
                const int64_t s = H.hard[h0 + tid];
                lim = hard_limit(A, s);
                row = A.qrow[s];
                qtv = A.inv_order[row];
            }
            s_lim[tid] = lim;
            s_qt[tid] = qtv;
            s_qrow[tid] = row;
        }
        __syncthreads();
        int64_t qt[kHQ];
        int64_t lim_max = 0;
#pragma unroll
        for (int i = 0; i < kHQ; i++) {
            qt[i] = s_qt[i];
            lim_max = s_lim[i] > lim_max ? s_lim[i] : lim_max;
        }
        if (staged)
            for (int t = tid; t < kHQ * m; t += kRsThreads) {
                const int i = t / m, col = t - i * m;
                s_vq[i * kHMaxM + col] = A.v[s_qt[i] * m + col];
            }
        const int64_t beg = (int64_t)seg * kHSeg;
        if (beg >= lim_max)
            continue; // no query of the group has candidates here
        const int64_t end = beg + kHSeg < lim_max ? beg + kHSeg : lim_max;
        for (int t = tid; t < kHQ * kHSort; t += kRsThreads) {
            sh[t] = __builtin_huge_valf();
            sl[t] = 0.0f;
            se[t] = 0.0f;
            sj[t] = 0x7fffffff;
        }
        for (int64_t c0 = beg; c0 < end; c0 += kHRound) {
            const int n_in = end - c0 < kHRound ? (int)(end - c0) : kHRound;
            __syncthreads();
            // phase A: eight lanes per candidate, kHQ dots each
            for (int e = grp; e < n_in; e += kRsThreads / 8) {
                const int64_t j = c0 + e;
                const float *vj = A.v + A.inv_order[j] * m;
                float s_[kHQ], c_[kHQ], sa[kHQ];
#pragma unroll
                for (int i = 0; i < kHQ; i++) {
                    s_[i] = 0.0f;
                    c_[i] = 0.0f;
                    sa[i] = 0.0f;
                }
                if (staged) {
                    // query rows in LDS, candidate row in float4 pieces
#pragma unroll 2
                    for (int32_t col = 4 * sub; col < m; col += 32) {
                        const float4 b = *(const float4 *)(vj + col);
#pragma unroll
                        for (int i = 0; i < kHQ; i++) {
                            const float4 a = *(const float4 *)(s_vq + i * kHMaxM + col);
                            dot2_step(a.x, b.x, s_[i], c_[i], sa[i]);
                            dot2_step(a.y, b.y, s_[i], c_[i], sa[i]);
                            dot2_step(a.z, b.z, s_[i], c_[i], sa[i]);
                            dot2_step(a.w, b.w, s_[i], c_[i], sa[i]);
                        }
                    }
                } else {
                    for (int32_t col = sub; col < m; col += 8) {
                        const float b = vj[col];
#pragma unroll
                        for (int i = 0; i < kHQ; i++)
                            dot2_step(A.v[qt[i] * m + col], b, s_[i], c_[i], sa[i]);
                    }
                }
#pragma unroll
                for (int i = 0; i < kHQ; i++) {
                    df dot = two_sum(s_[i], c_[i]);
                    float sab = sa[i];
#pragma unroll
                    for (int o = 1; o < 8; o <<= 1) {
                        const df other = {__shfl_xor(dot.hi, o, 64), __shfl_xor(dot.lo, o, 64)};
                        dot = df_add(dot, other);
                        sab += __shfl_xor(sab, o, 64);
                    }
                    if (sub == 0) {
                        const int slot = i * kHSort + k + e;
                        sh[slot] = dot.hi;
                        sl[slot] = dot.lo;
                        se[slot] = sab;
                        sj[slot] = (int32_t)j;
                    }
                }
            }
            __syncthreads();
            // phase B: one lane per (query, candidate) pair
            for (int t = tid; t < kHQ * n_in; t += kRsThreads) {
                const int i = t / n_in, e = t - i * n_in;
                const int slot = i * kHSort + k + e;
                const int32_t j = sj[slot];
                float dh = __builtin_huge_valf(), dl = 0.0f, er = 0.0f;
                int32_t jj = 0x7fffffff;
                if (i < nq_g && (int64_t)j < s_lim[i]) {
                    jj = j;
                    if (j == s_qrow[i]) {
                        dh = 0.0f;
                    } else {
                        const df dd = finish_distance(A, s_qt[i], A.inv_order[j],
                                                      df{sh[slot], sl[slot]}, se[slot], &er);
                        dh = dd.hi;
                        dl = dd.lo;
                    }
                }
                sh[slot] = dh;
                sl[slot] = dl;
                se[slot] = er;
                sj[slot] = jj;
            }
            int p2 = 2;
            while (p2 < k + n_in)
                p2 <<= 1;
            // entries past k left by an earlier round rank behind the current
            // top-k, so they never re-enter it
            sort_arrays(sh, sl, se, sj, kHQ, p2);
        }
        HardEntry *out = H.part + ((int64_t)gl * H.nseg + seg) * kHQ * k;
        for (int t = tid; t < kHQ * k; t += kRsThreads) {
            const int i = t / k, r = t - i * k;
            const int slot = i * kHSort + r;
            out[t] = HardEntry{sh[slot], sl[slot], se[slot], sj[slot]};
        }
        __syncthreads();
    }
}

// Merge the segment lists of each hard query of the batch; write its row.
__global__ void __launch_bounds__(kRsThreads) hard_merge_kernel(HardArgs H)
{
    __shared__ float sh[kRsSort], sl[kRsSort], se[kRsSort];
    __shared__ int32_t sj[kRsSort];
    const RescoreArgs &A = H.r;
    const int32_t k = A.k;
    const int32_t per = (kRsSort - k) / k; // lists merged per round
    const int32_t h_begin = H.g0 * kHQ;
    const int32_t h_end = (H.g0 + H.gn) * kHQ < H.n_hard ? (H.g0 + H.gn) * kHQ : H.n_hard;
    for (int32_t h = h_begin + blockIdx.x; h < h_end; h += gridDim.x) {
        const int64_t s = H.hard[h];
        const int64_t limit = hard_limit(A, s);
        const int32_t nl_all = (int32_t)((limit + kHSeg - 1) / kHSeg);
        const int32_t gl = h / kHQ - H.g0, i = h % kHQ;
        __syncthreads();
        clear_entries(sh, sl, se, sj);
        for (int32_t g = 0; g < nl_all; g += per) {
            const int32_t nl = nl_all - g < per ? nl_all - g : per;
            __syncthreads();
            for (int t = threadIdx.x; t < nl * k; t += kRsThreads) {
                const int32_t seg = g + t / k;
                const HardEntry e =
                    H.part[(((int64_t)gl * H.nseg + seg) * kHQ + i) * k + t % k];
                sh[k + t] = e.dh;
                sl[k + t] = e.dl;
                se[k + t] = e.err;
                sj[k + t] = e.j;
            }
            int p2 = 2;
            while (p2 < k + nl * k)
                p2 <<= 1;
            sort_entries(sh, sl, se, sj, p2);
        }
        write_row(A, A.qperm[s], sh, sl, se, sj);
    }
}

__global__ void hard_flag_kernel(const int32_t *__restrict__ cnt, int64_t nq, int32_t *flag)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nq)
        flag[s] = cnt[s] > kCap ? 1 : 0;
}

// ordered compaction: hard[pos[s]] = s; the count goes to n_hard
__global__ void hard_scatter_kernel(const int32_t *__restrict__ flag,
                                    const int32_t *__restrict__ pos, int64_t nq, int32_t *hard,
                                    int32_t *n_hard)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nq)
        return;
    if (flag[s])
        hard[pos[s]] = (int32_t)s;
    if (s == nq - 1)
        *n_hard = pos[s] + flag[s];
}

// Query telemetry of a mixed call: pairs screened (every valid candidate of
// every query), candidates listed for rescoring, hard queries and their
// pairs, the longest list.
__global__ void mixed_telemetry_kernel(const int64_t *__restrict__ qsorted,
                                       const int32_t *__restrict__ cnt, int64_t nq, int64_t n,
                                       int32_t pred, int64_t *telem)
{
    __shared__ unsigned long long s_acc[4];
    __shared__ int s_max;
    if (threadIdx.x < 4)
        s_acc[threadIdx.x] = 0ull;
    if (threadIdx.x == 0)
        s_max = 0;
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nq) {
        const int64_t qid = qsorted[s];
        const int64_t lim = pred ? (qid < 0 ? n : (qid < n ? qid : n)) : n;
        const int32_t c = cnt[s];
        const bool hard = c > kCap;
        atomicAdd(&s_acc[0], (unsigned long long)lim);
        atomicAdd(&s_acc[1], hard ? 0ull : (unsigned long long)c);
        atomicAdd(&s_acc[2], hard ? 1ull : 0ull);
        atomicAdd(&s_acc[3], hard ? (unsigned long long)lim : 0ull);
        atomicMax(&s_max, c);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd((unsigned long long *)&telem[CX_QT_DISTANCE_EVALS], s_acc[0]);
        atomicAdd((unsigned long long *)&telem[CX_QT_LISTED], s_acc[1]);
        atomicAdd((unsigned long long *)&telem[CX_QT_HARD_QUERIES], s_acc[2]);
        atomicAdd((unsigned long long *)&telem[CX_QT_HARD_PAIRS], s_acc[3]);
        atomicMax((unsigned long long *)&telem[CX_QT_MAX_LIST], (unsigned long long)s_max);
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
    MetaScale sc{};
    for (int c = 0; c < t.d; c++)
        sc.il[c] = t.inv_ls[c];
    gather_meta_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
        t.raw.ptr, t.p.ptr, t.inv_order.ptr, n, t.d, sc, t.x_ds.ptr, t.isp_ds.ptr);
    CX_HIP(hipGetLastError());
    CX_HIP(hipStreamSynchronize(st));
    t.mixed_ready = true;
}

// pair codes kept per launch group of query tiles (bytes)
constexpr int64_t kCodeBudget = (int64_t)4 << 30;
// segment top-k lists of the hard queries per launch batch (bytes)
constexpr int64_t kPartBudget = (int64_t)512 << 20;

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

    // workspace: sorted queries, permutation, rows, tau, counts, histograms,
    // lists, and the hard-query bookkeeping
    auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += align(bytes);
        return o;
    };
    const size_t o_qs = take(sizeof(int64_t) * nq);
    const size_t o_perm = take(sizeof(int64_t) * nq);
    const size_t o_iota = take(sizeof(int64_t) * nq);
    const size_t o_row = take(sizeof(int32_t) * nq);
    const size_t o_tau = take(sizeof(float) * nq);
    const size_t o_cnt = take(sizeof(int32_t) * nq);
    const size_t o_hist = take(sizeof(int32_t) * nq * kBins);
    const size_t o_ctr = take(256);
    const size_t o_list = take(sizeof(int32_t) * (size_t)nq * kCap);
    const size_t o_flag = take(sizeof(int32_t) * nq);
    const size_t o_pos = take(sizeof(int32_t) * nq);
    const size_t o_hard = take(sizeof(int32_t) * nq);
    size_t sort_bytes = 0, scan_bytes = 0;
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const int64_t *)nullptr,
                                              (int64_t *)nullptr, (const int64_t *)nullptr,
                                              (int64_t *)nullptr, (int)nq, 0, 64, st));
    CX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const int32_t *)nullptr,
                                            (int32_t *)nullptr, (int)nq, st));
    const size_t o_tmp = take(std::max(sort_bytes, scan_bytes));
    uint8_t *ws = workspace(t, off);
    int64_t *qs = (int64_t *)(ws + o_qs), *perm = (int64_t *)(ws + o_perm);
    int64_t *iota = (int64_t *)(ws + o_iota);
    int32_t *qrow = (int32_t *)(ws + o_row), *cnt = (int32_t *)(ws + o_cnt);
    int32_t *ctr = (int32_t *)(ws + o_ctr), *list = (int32_t *)(ws + o_list);
    float *tau = (float *)(ws + o_tau);
    int32_t *hist = (int32_t *)(ws + o_hist);
    int32_t *flag = (int32_t *)(ws + o_flag), *pos = (int32_t *)(ws + o_pos);
    int32_t *hard = (int32_t *)(ws + o_hard);
    // counters: ctr[0..1] rescore items (int64), ctr[2..3] hard items (int64),
    // ctr[4] hard count
    int64_t *rs_items = (int64_t *)ctr, *hs_items = (int64_t *)(ctr + 2);
    int32_t *n_hard_d = ctr + 4;

    const unsigned qb = (unsigned)((nq + 255) / 256);
    iota64_kernel<<<qb, 256, 0, st>>>(iota, nq);
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(ws + o_tmp, sort_bytes, d_queries, qs, iota, perm,
                                              (int)nq, 0, 64, st));
    query_rows_kernel<<<qb, 256, 0, st>>>(qs, nq, n, t.order.ptr, qrow);
    prof_mark(t, 2);
    CX_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * nq, st));
    CX_HIP(hipMemsetAsync(hist, 0, sizeof(int32_t) * nq * kBins, st));
    CX_HIP(hipMemsetAsync(ctr, 0, 256, st));

    ScreenArgs sa{};
    sa.vhi = t.vhi.ptr;
    sa.vlo = t.vlo.ptr;
    sa.x = t.x_ds.ptr; // scaled by 1/lengthscale
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
    // |computed V.V - exact| <= cD |V_q||V_j|: the hi/lo split drops at most
    // 3.1 * 2^-18 of each product, the hi.hi FP32 accumulation adds gamma_mp,
    // the cross terms (2^-8 smaller) gamma_2mp, and the final sum one rounding.
    const double u = 1.0 / 16777216.0;
    const double g1 = (t.mp + 16) * u / (1.0 - (t.mp + 16) * u);
    const double g2 = (2 * t.mp + 16) * u / (1.0 - (2 * t.mp + 16) * u);
    sa.cD = (float)((3.1 / 262144.0 + g1 * 1.01 + g2 * 0.01 + 2 * u) * 1.05);
    // kernel value error: the FP32 kernel (exp2/sqrt, folded constants) stays
    // within 32 u var; scaled coordinates |x/l| <= R add at most
    // u var (1.3 R sqrt(d) + 0.4 (d + 2)) through d^2 (both kernels)
    sa.eK = (float)(std::fabs((double)t.var) * u *
                    (64.0 + 1.5 * t.xs_max * std::sqrt((double)t.d) + (t.d + 2.0)));
    const double l2e = 1.4426950408889634;
    if (t.kernel == 0) {
        sa.kA = (float)(-0.5 * l2e);
    } else {
        sa.kA = (float)(5.0 * l2e * l2e);
        sa.kB = (float)((double)t.var / l2e);
        sa.kC = (float)((double)t.var / (3.0 * l2e * l2e));
    }
    sa.tau = tau;
    sa.hist = hist;
    sa.cnt = cnt;
    sa.list = list;
    const int32_t n_tiles = (int32_t)((nq + kQT - 1) / kQT);
    sa.n_seg = (int32_t)((n + kSeg - 1) / kSeg);
    const size_t dyn_lds = sizeof(float) * 2 * (size_t)t.d * kCT;
    // launch groups of query tiles whose pair codes fit the code budget
    const int64_t tile_bytes = (int64_t)sa.n_seg * kSeg * kQT;
    const int32_t group =
        (int32_t)std::max<int64_t>(1, std::min<int64_t>(n_tiles, kCodeBudget / tile_bytes));
    if (t.codes.count < (size_t)(group * tile_bytes)) {
        CX_HIP(hipStreamSynchronize(st));
        t.codes.alloc((size_t)(group * tile_bytes));
    }
    sa.codes = (uint32_t *)t.codes.ptr;
    const bool one_group = group >= n_tiles;
    for (int32_t t0 = 0; t0 < n_tiles; t0 += group) {
        sa.tile0 = t0;
        sa.n_tiles = std::min(group, n_tiles - t0);
        launch_screen(sa, dyn_lds, st);
        if (one_group)
            prof_mark(t, 3);
        const int64_t s_begin = (int64_t)t0 * kQT;
        const int64_t s_end = std::min<int64_t>(nq, (int64_t)(t0 + sa.n_tiles) * kQT);
        tau_kernel<<<(unsigned)((s_end - s_begin + 255) / 256), 256, 0, st>>>(hist, s_begin,
                                                                               s_end, k, tau);
        if (one_group)
            prof_mark(t, 4);
        filter_kernel<<<dim3((unsigned)sa.n_tiles, (unsigned)sa.n_seg), kScreenThreads, 0, st>>>(
            sa);
        CX_HIP(hipGetLastError());
    }
    if (!one_group) {
        prof_mark(t, 3);
        prof_mark(t, 4);
    }
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
    ra.counter = rs_items;

    // hard queries: ordered list (read back once) ...
    hard_flag_kernel<<<qb, 256, 0, st>>>(cnt, nq, flag);
    CX_HIP(hipcub::DeviceScan::ExclusiveSum(ws + o_tmp, scan_bytes, flag, pos, (int)nq, st));
    hard_scatter_kernel<<<qb, 256, 0, st>>>(flag, pos, nq, hard, n_hard_d);
    if (!t.host_word)
        CX_HIP(hipHostMalloc((void **)&t.host_word, 64, hipHostMallocDefault));
    CX_HIP(hipMemcpyAsync(t.host_word, n_hard_d, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (!t.sync_event)
        CX_HIP(hipEventCreateWithFlags(&t.sync_event, hipEventDisableTiming));
    hipEvent_t ev = t.sync_event;
    CX_HIP(hipEventRecord(ev, st));
    // ... while the listed queries are rescored
    const int64_t blocks = nq < 2048 ? nq : 2048;
    rescore_kernel<<<(unsigned)blocks, kRsThreads, 0, st>>>(ra);
    CX_HIP(hipGetLastError());
    prof_mark(t, 6);
    CX_HIP(hipEventSynchronize(ev));
    const int32_t n_hard = *t.host_word;
    t.last_hard = n_hard;
    if (n_hard > 0) {
        HardArgs ha{};
        ha.r = ra;
        ha.hard = hard;
        ha.n_hard = n_hard;
        ha.nseg = (int32_t)((n + kHSeg - 1) / kHSeg);
        ha.counter = hs_items;
        const int32_t n_groups = (n_hard + kHQ - 1) / kHQ;
        const int64_t group_bytes = (int64_t)ha.nseg * kHQ * k * sizeof(HardEntry);
        const int32_t batch = (int32_t)std::max<int64_t>(
            1, std::min<int64_t>(n_groups, kPartBudget / group_bytes));
        if (t.part.count < (size_t)(batch * group_bytes)) {
            CX_HIP(hipStreamSynchronize(st));
            t.part.alloc((size_t)(batch * group_bytes));
        }
        ha.part = (HardEntry *)t.part.ptr;
        for (int32_t g0 = 0; g0 < n_groups; g0 += batch) {
            ha.g0 = g0;
            ha.gn = std::min(batch, n_groups - g0);
            CX_HIP(hipMemsetAsync(hs_items, 0, sizeof(int64_t), st));
            const int64_t items = (int64_t)ha.gn * ha.nseg;
            hard_tile_kernel<<<(unsigned)std::min<int64_t>(items, 2048), kRsThreads, 0, st>>>(ha);
            CX_HIP(hipGetLastError());
            const int64_t hq = std::min<int64_t>((int64_t)ha.gn * kHQ, 2048);
            hard_merge_kernel<<<(unsigned)hq, kRsThreads, 0, st>>>(ha);
            CX_HIP(hipGetLastError());
        }
    }
    if (t.telem)
        mixed_telemetry_kernel<<<qb, 256, 0, st>>>(qs, cnt, nq, n, predecessor ? 1 : 0, t.telem);
    prof_mark(t, 7);
}

} // namespace cx
