// Mixed-precision batched residual kNN (CX_KNN_MIXED): exact neighbours under
// d = sqrt(1 - |rho|) evaluated on the stored inputs, with a certified screen.
//
// Work is laid out as query tiles x candidate tiles in dataset order, so the
// Vecchia mask j < i (src/algo.rs:918-925 / src/lib.rs:444-448) becomes a
// triangle: a query tile only visits candidates below its largest index.
//
//   prep      V -> dataset-order BF16 hi/lo split (V = hi + lo + r, |r| <=
//             2^-18 |V|), row norms |V_i| rounded up.
//   screen    one workgroup per 64-query tile x 2048-candidate segment, four
//             wavefronts of 32 candidates x 64 queries; V.V on MFMA
//             (v_mfma_f32_32x32x16_bf16) as hi.hi + (hi.lo + lo.hi) in two
//             FP32 accumulators, kernel value in FP32, and a certified
//             interval [lb2, ub2] containing the exact squared distance
//             1 - |rho|.
//             pass 1: per-query histogram of ub2 -> tau_q, an upper bound on
//                     the k-th exact squared distance;
//             pass 2: every candidate with lb2 <= tau_q is appended to the
//                     query's list (a superset of the exact k nearest).
//   rescore   one wavefront per query: compensated FP32 (Dot2 dot,
//             double-float distance) for each listed candidate, exact top-k
//             in registers, ascending output with a per-result error bound.
//   hard      a query whose list overflowed its kCap slots is scored over all
//             its valid candidates the same way, one wavefront per query and
//             candidate segment, the segments' top-k lists merged after.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>

#include "cx_internal.hpp"
#include "df.hpp"

namespace cx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kQT = 64;      // queries per workgroup tile
constexpr int kCW = 64;      // candidates per wavefront block
constexpr int kCT = 4 * kCW; // candidates per workgroup tile
constexpr int kMpAlign = 32; // V columns padded to this (int8 slice K steps too)
constexpr int kKC = 64;      // K chunk of query rows staged in LDS
constexpr int kKP = kKC + 8; // padded LDS row (bf16)
constexpr int kBins = 128;   // histogram bins of the squared upper bound over [0, 1)
constexpr float kHistCap = 0.99f; // squared bounds at or above this are not counted
constexpr int kCap = 2048;   // listed candidates per query before overflow
// second listing of queries still overflowing after the exact-product screen
constexpr int kCapBig = 8192;
constexpr int64_t kBigListBudget = (int64_t)1 << 30; // bytes of those lists per call
constexpr int kMaxD = 16;
constexpr int kMixMaxK = 256;
// candidates per workgroup of the screen (a segment; grid y of the filter):
// a multiple of kCT in [kSegMin, kSegMax], chosen per tree (see screen_segment)
constexpr int32_t kSegMin = kCT, kSegMax = 64 * kCT;
constexpr int kScreenWaves = 4;
constexpr int kScreenThreads = 64 * kScreenWaves;
// Workgroup b of a launch runs on XCD b % kXcds (round-robin dispatch)
constexpr int kXcds = 8;

// ---- prep ------------------------------------------------------------------

__device__ __forceinline__ uint16_t bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Offset (in bf16 elements) of V[row][k], part h (0 hi, 1 lo), in the
// MFMA-fragment layout: rows in blocks of 32, K in steps of 16; block R, step
// s and part h form one 1 KiB fragment whose lane l holds the 8 values of row
// 32R + (l & 31) at k = 16s + 8(l >> 5) .. + 7, the operand layout of
// v_mfma_f32_32x32x16_bf16, so a wavefront loads a fragment as one
// contiguous 1 KiB read.
__device__ __forceinline__ int64_t frag_off(int64_t row, int32_t k, int h, int32_t steps)
{
    const int64_t chunk = ((row >> 5) * steps + (k >> 4)) * 2 + h;
    const int lane = (int)(row & 31) + 32 * ((k >> 3) & 1);
    return (chunk * 64 + lane) * 8 + (k & 7);
}

// V (tree order) -> dataset-order hi/lo in the fragment layout (zero padded
// to mp columns; rows past n stay zero), row norms |V_i| rounded up.
__global__ void split_kernel(const float *__restrict__ v, const int64_t *__restrict__ inv_order,
                             int64_t n, int32_t m, int32_t mp, uint16_t *__restrict__ vsw,
                             float *__restrict__ vnorm)
{
    const int64_t row = blockIdx.x; // dataset index
    if (row >= n)
        return;
    const float *src = v + inv_order[row] * (int64_t)m;
    const int32_t steps = mp / 16;
    float acc = 0.0f;
    for (int32_t c = threadIdx.x; c < mp; c += blockDim.x) {
        float x = c < m ? src[c] : 0.0f;
        uint16_t h = bf16_rne(x);
        uint16_t l = bf16_rne(x - bf16_to_f(h));
        vsw[frag_off(row, c, 0, steps)] = h;
        vsw[frag_off(row, c, 1, steps)] = l;
        acc = __builtin_fmaf(x, x, acc);
    }
    __shared__ float s_acc[64];
    for (int o = 32; o > 0; o >>= 1)
        acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0)
        s_acc[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        const int nw = (int)(blockDim.x >> 6);
        for (int w = 0; w < nw; w++)
            t += s_acc[w];
        // a sum of non-negative terms through `depth` roundings is within
        // gamma_depth of the exact sum: scale it up by that, then round the
        // norm up so products of norms bound |V_i . V_j| from above
        const int depth = (mp + (int)blockDim.x - 1) / (int)blockDim.x + 6 + nw;
        t *= 1.0f + (float)(depth + 2) * 1.2e-7f;
        vnorm[row] = __builtin_sqrtf(t) * (1.0f + 1.2e-7f) + 1e-30f;
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
    const uint16_t *vsw; // npad x mp hi and lo, dataset order, fragment layout
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
    int32_t seg; // candidates per segment
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
// lane): sixteen consecutive words per lane, word (4 rb + g) 2 + b holding
// the wavefront's candidate rows 32 rb + 8 g + 4 (lane >> 5) .. + 3 (one byte
// each) against query column 32 b + (lane & 31), so a wavefront stores and
// reloads 4 KiB contiguous per candidate tile.
__device__ __forceinline__ int64_t code_word(const ScreenArgs &A, int64_t blk, int ct, int wid,
                                             int lane)
{
    return (((blk * (A.seg / kCT) + ct) * kScreenWaves + wid) * 64 + lane) * 16;
}

// Per-lane query columns of a query tile (shared by the screen and the
// filter): column, Vecchia limit on the candidate index, own dataset row.
__device__ __forceinline__ void tile_queries(const ScreenArgs &A, int64_t s0, int lane,
                                             int32_t *qcol, int32_t *qlim, int32_t *qself,
                                             int32_t *qrowv)
{
#pragma unroll
    for (int b = 0; b < 2; b++) {
        qcol[b] = 32 * b + (lane & 31);
        const int64_t s = s0 + qcol[b];
        const bool live = s < A.nq;
        const int64_t qid = live ? A.qsorted[s] : 0;
        const int32_t row = live ? A.qrow[s] : 0;
        qlim[b] = !live ? 0 : (A.pred ? (qid < 0 ? A.n : (qid < A.n ? (int32_t)qid : A.n)) : A.n);
        qself[b] = live ? row : -1;
        qrowv[b] = row;
    }
}

// The screen's workgroup -> (query tile, segment) map.  Each XCD takes whole
// candidate segments (so a segment's 2048 candidate rows, 1 MiB at rank
// 256, stay in that XCD's L2 while every query tile streams them) and runs
// all query tiles of one segment, heaviest first, before its next one.
// Segments go to XCDs in snake order (0..7, then 7..0, ...): under the
// Vecchia mask segment s is needed by the tiles above it, so pairing early
// and late segments balances the XCDs.  The grid holds kXcds x
// ceil(n_seg / kXcds) x n_tiles workgroups; those past the last segment and
// tiles with no candidate in the segment exit at once.
__device__ __forceinline__ bool screen_block(const ScreenArgs &A, int32_t *tl, int32_t *seg)
{
    const int32_t bid = (int32_t)blockIdx.x;
    const int32_t xcd = bid % kXcds, k = bid / kXcds;
    const int32_t round = k / A.n_tiles;
    const int32_t pos = (round & 1) ? kXcds - 1 - xcd : xcd;
    const int32_t sg = round * kXcds + pos;
    if (sg >= A.n_seg)
        return false;
    *seg = sg;
    *tl = A.n_tiles - 1 - k % A.n_tiles;
    return true;
}

// candidate range [c_begin, c_end) of this (tile, segment); empty if none
__device__ __forceinline__ void tile_range(const ScreenArgs &A, int64_t s0, int32_t seg,
                                           int32_t *c_begin, int32_t *c_end)
{
    int32_t e = A.n;
    if (A.pred) {
        // [0, largest query index) under the mask
        const int64_t last = s0 + kQT - 1 < A.nq - 1 ? s0 + kQT - 1 : A.nq - 1;
        const int64_t qmax = A.qsorted[last];
        e = qmax < 0 ? A.n : (qmax < A.n ? (int32_t)qmax : A.n);
    }
    const int32_t b = seg * A.seg;
    *c_begin = b;
    *c_end = e < b + A.seg ? e : b + A.seg;
}

// Query rows of a K chunk: 64 rows x kKC bf16, hi and lo, as 16-byte pieces
// of the fragment layout; thread t moves row t / 4, K16 step t % 4 of the
// chunk (both 8-wide halves, hi and lo).
struct BPieces {
    uint4 h0, h1, l0, l1;
};

__device__ __forceinline__ BPieces load_b_chunk(const ScreenArgs &A, const int32_t *sQrow,
                                                int32_t kc, int tid)
{
    // the fragment layout's columns are padded to whole chunks
    const int32_t steps = A.mp >> 4;
    const int32_t s = (kc >> 4) + (tid & 3);
    const int64_t row = sQrow[tid >> 2];
    const int64_t base = (((row >> 5) * steps + s) * 2) * 64 + (row & 31);
    const uint4 *v = (const uint4 *)A.vsw;
    BPieces p;
    p.h0 = v[base];
    p.h1 = v[base + 32];
    p.l0 = v[base + 64];
    p.l1 = v[base + 96];
    return p;
}

__device__ __forceinline__ void store_b_chunk(uint16_t *sBh, uint16_t *sBl, const BPieces &p,
                                              int tid)
{
    const int row = tid >> 2, col = 16 * (tid & 3);
    *(uint4 *)(sBh + row * kKP + col) = p.h0;
    *(uint4 *)(sBh + row * kKP + col + 8) = p.h1;
    *(uint4 *)(sBl + row * kKP + col) = p.l0;
    *(uint4 *)(sBl + row * kKP + col + 8) = p.l1;
}

// Pass 1.  A workgroup owns 64 queries (MFMA columns) and loops over a
// 2048-candidate segment in tiles of 256 candidates; each of its four
// wavefronts computes a 64 x 64 block (candidates 64 w .. of the tile: four
// 32 x 32 MFMA blocks, each with a hi.hi and a cross accumulator, 128
// accumulator registers).  Candidate fragments come straight from global
// memory in the fragment layout (one 1 KiB load per fragment, two K16 steps
// ahead); query rows go through LDS in double-buffered K chunks of 64, one
// barrier per chunk.  Per pair: the certified interval [rlo, rhi] of |rho|;
// 1 - rlo feeds the query's histogram, 1 - rhi is stored as a one-byte code
// for the filter.  DD > 0 fixes the coordinate dimension (query coordinates
// in registers); DD == 0 reads them from LDS for any d <= kMaxD.  KT is the
// kernel (0 RBF, 1 Matern-5/2).  Coordinates arrive scaled by 1/lengthscale.
constexpr int kHistRow = kBins / 2 + 1; // words per query column, padded

template <int DD, int KT>
__global__ void __launch_bounds__(kScreenThreads, 2) screen_kernel(ScreenArgs A)
{
    __shared__ __attribute__((aligned(16))) uint16_t sB[2][2][kQT * kKP]; // [buf][hi, lo]
    __shared__ int32_t sQrow[kQT];
    // histogram: one 32-bit counter per bin when d is fixed (d <= 5 leaves
    // room for two workgroups per CU), else two 16-bit bins per word (at
    // most kSegMax counts per bin); rows padded by one word
    constexpr bool kWide = DD > 0;
    constexpr int kRow = kWide ? kBins + 1 : kHistRow;
    __shared__ uint32_t sHist[kQT * kRow];
    // dynamic: per wavefront its 64 candidates' coordinates [d][64], 1/sqrt(p)
    // and norms; then the query coordinates [d][kQT]
    extern __shared__ __attribute__((aligned(16))) float sDyn[];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int32_t tl, seg;
    if (!screen_block(A, &tl, &seg))
        return;
    const int64_t s0 = (int64_t)(A.tile0 + tl) * kQT;
    const int64_t blk = (int64_t)tl * A.n_seg + seg;
    const int32_t mp = A.mp, steps = mp >> 4;
    const int32_t d = DD > 0 ? DD : A.d;
    int32_t c_begin, c_end;
    tile_range(A, s0, seg, &c_begin, &c_end);
    if (c_begin >= c_end)
        return;
    float *sCx = sDyn + wid * (d + 2) * kCW, *sCw = sCx + d * kCW, *sCn = sCw + kCW;
    float *sQx = sDyn + kScreenWaves * (d + 2) * kCW;

    for (int i = tid; i < kQT; i += kScreenThreads) {
        const int64_t s = s0 + i;
        const int32_t row = s < A.nq ? A.qrow[s] : 0;
        sQrow[i] = row;
        for (int c = 0; c < d; c++)
            sQx[c * kQT + i] = A.x[(int64_t)row * d + c];
    }
    for (int i = tid; i < kQT * kRow; i += kScreenThreads)
        sHist[i] = 0u;
    __syncthreads();

    int32_t qcol[2], qlim[2], qself[2], qrowv[2];
    tile_queries(A, s0, lane, qcol, qlim, qself, qrowv);
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
    // smallest Vecchia limit over the tile's 64 queries (every wavefront
    // holds all 64 columns): candidate blocks below it are valid for all
    int32_t lim_min = qlim[0] < qlim[1] ? qlim[0] : qlim[1];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t t = __shfl_xor(lim_min, o, 64);
        lim_min = t < lim_min ? t : lim_min;
    }
    const float eKp = A.eK * 1.000002f;

    // B chunk 0 into buffer 0; later chunks cycle through the K range (the
    // query rows are the same for every candidate tile)
    const int32_t n_chunks = mp / kKC;
    {
        const BPieces p = load_b_chunk(A, sQrow, 0, tid);
        store_b_chunk(sB[0][0], sB[0][1], p, tid);
    }
    __syncthreads();
    int buf = 0;
    int32_t chunk = 0; // chunk held in sB[buf]

    const uint4 *va = (const uint4 *)A.vsw;
    // this lane's 16-byte piece of fragment (row block R, step s, part h)
    auto a_piece = [&](int64_t R, int32_t s, int h) -> uint4 {
        return va[((R * steps + s) * 2 + h) * 64 + lane];
    };

    // candidate fragments two K16 steps ahead: a[s & 1] = {hi R0, lo R0, hi
    // R0+1, lo R0+1}.  Every step refills its slot unconditionally (no branch
    // for the wait counts to merge over): past a tile's last step with the
    // next tile's first steps (or, after the wavefront's last tile, the same
    // block again, unused).
    uint4 a[2][4];
    if (c_begin + kCW * wid < c_end) {
        const int64_t R = (c_begin + kCW * wid) >> 5;
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int f = 0; f < 4; f++)
                a[q][f] = a_piece(R + (f >> 1), q, f & 1);
    }

    for (int32_t c0 = c_begin; c0 < c_end; c0 += kCT) {
        const int ct = (c0 - c_begin) / kCT;
        const int32_t cw0 = c0 + kCW * wid;   // this wavefront's first candidate
        const bool live = cw0 < c_end;        // uniform per wavefront
        const int64_t R0 = cw0 >> 5;
        // the block whose first steps follow this tile's last ones
        const int64_t Rn = cw0 + kCT < c_end ? R0 + kCT / 32 : R0;
        // this wavefront's candidate data (read after the K loop)
        {
            const int32_t j = cw0 + lane < A.n ? cw0 + lane : 0;
            for (int c = 0; c < d; c++)
                sCx[c * kCW + lane] = A.x[(int64_t)j * d + c];
            sCw[lane] = A.isp[j];
            sCn[lane] = A.vnorm[j];
        }
        // per 32 x 32 block: hi.hi and the cross terms hi.lo + lo.hi
        floatx16 acc_h[2][2], acc_x[2][2]; // [row block][query block]
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                acc_h[rb][b] = floatx16{};
                acc_x[rb][b] = floatx16{};
            }
        for (int32_t ci = 0; ci < n_chunks; ci++) {
            // the next chunk's query pieces fly while this one is multiplied
            const int32_t nxt = chunk + 1 < n_chunks ? chunk + 1 : 0;
            const BPieces pb = load_b_chunk(A, sQrow, nxt * kKC, tid);
            const uint16_t *bh = sB[buf][0], *bl = sB[buf][1];
            if (live) {
                __builtin_amdgcn_s_setprio(1);
                // query fragments one K16 step ahead of the MFMAs
                bf16x8 fb[2][2][2]; // [slot][query block][hi, lo]
                auto read_b = [&](int j, int slot) {
                    const int koff = 16 * j + 8 * (lane >> 5);
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        fb[slot][b][0] = __builtin_bit_cast(
                            bf16x8, *(const uint4 *)(bh + qcol[b] * kKP + koff));
                        fb[slot][b][1] = __builtin_bit_cast(
                            bf16x8, *(const uint4 *)(bl + qcol[b] * kKP + koff));
                    }
                };
                read_b(0, 0);
#pragma unroll
                for (int j = 0; j < kKC / 16; j++) {
                    const int32_t s = chunk * (kKC / 16) + j;
                    const int q = j & 1;
                    if (j + 1 < kKC / 16)
                        read_b(j + 1, (j + 1) & 1);
#pragma unroll
                    for (int rb = 0; rb < 2; rb++) {
                        const bf16x8 ah = __builtin_bit_cast(bf16x8, a[q][2 * rb]);
                        const bf16x8 al = __builtin_bit_cast(bf16x8, a[q][2 * rb + 1]);
#pragma unroll
                        for (int b = 0; b < 2; b++) {
                            acc_h[rb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                ah, fb[j & 1][b][0], acc_h[rb][b], 0, 0, 0);
                            acc_x[rb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                ah, fb[j & 1][b][1], acc_x[rb][b], 0, 0, 0);
                            acc_x[rb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                al, fb[j & 1][b][0], acc_x[rb][b], 0, 0, 0);
                        }
                    }
                    // refill this slot with step s + 2 (of this tile or the next)
                    const bool wrap = s + 2 >= steps;
                    const int64_t Rf = wrap ? Rn : R0;
                    const int32_t sf = wrap ? s + 2 - steps : s + 2;
#pragma unroll
                    for (int f = 0; f < 4; f++)
                        a[q][f] = a_piece(Rf + (f >> 1), sf, f & 1);
                    // keep the refill here: the scheduler would otherwise sink
                    // it next to its use two steps on, exposing its latency
                    __builtin_amdgcn_sched_barrier(0);
                }
                __builtin_amdgcn_s_setprio(0);
            }
            // every wavefront has finished reading sB[buf ^ 1] (the chunk
            // before this one) at the previous barrier
            store_b_chunk(sB[buf ^ 1][0], sB[buf ^ 1][1], pb, tid);
            __syncthreads();
            buf ^= 1;
            chunk = nxt;
        }
        if (!live)
            continue;
        // epilogue: register r of block (rb, b) is candidate cw0 + 32 rb +
        // 8 (r / 4) + r % 4 + 4 (lane >> 5), query column qcol[b].  FULL:
        // every pair of the block lies below every query's Vecchia limit.
        auto epilogue = [&](auto full_tag) {
            constexpr bool FULL = decltype(full_tag)::value;
            uint32_t words[16];
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int g = 0; g < 4; g++) {
                    const int rg = 32 * rb + 8 * g + 4 * (lane >> 5); // within the wavefront
                    const float4 cw = *(const float4 *)(sCw + rg);
                    const float4 cn = *(const float4 *)(sCn + rg);
                    float d2[2][4];
#pragma unroll
                    for (int b = 0; b < 2; b++)
#pragma unroll
                        for (int rr = 0; rr < 4; rr++)
                            d2[b][rr] = 0.0f;
                    for (int c = 0; c < d; c++) {
                        const float4 cx = *(const float4 *)(sCx + c * kCW + rg);
#pragma unroll
                        for (int b = 0; b < 2; b++) {
                            const float q = DD > 0 ? qx[b][DD > 0 ? c : 0] : sQx[c * kQT + qcol[b]];
                            const float t0 = q - cx.x, t1 = q - cx.y;
                            const float t2 = q - cx.z, t3 = q - cx.w;
                            d2[b][0] = __builtin_fmaf(t0, t0, d2[b][0]);
                            d2[b][1] = __builtin_fmaf(t1, t1, d2[b][1]);
                            d2[b][2] = __builtin_fmaf(t2, t2, d2[b][2]);
                            d2[b][3] = __builtin_fmaf(t3, t3, d2[b][3]);
                        }
                    }
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        uint32_t word = 0u;
#pragma unroll
                        for (int rr = 0; rr < 4; rr++) {
                            const int r = 4 * g + rr;
                            const float cwv = rr == 0 ? cw.x : rr == 1 ? cw.y : rr == 2 ? cw.z : cw.w;
                            const float cnv = rr == 0 ? cn.x : rr == 1 ? cn.y : rr == 2 ? cn.z : cn.w;
                            const float kv = fast_kernel<KT>(d2[b][rr], A);
                            const float an =
                                __builtin_fabsf(kv - (acc_h[rb][b][r] + acc_x[rb][b][r]));
                            // 1/sqrt(max(p_q p_j, 1e-18)) from per-row 1/sqrt(p)
                            const float w = __builtin_fminf(qw[b] * cwv, 1e9f);
                            // certified |rho| error w (eK + cD |V_q||V_j|) (1 + 2e-6)
                            // for the kernel, the dot and the final sum, plus
                            // 9.6e-7 |num| w for the roundings of num, w and
                            // the products here: rlo <= |rho| <= rhi
                            const float ew = w * __builtin_fmaf(qcn[b], cnv, eKp);
                            const float aw = an * w;
                            const float rlo = __builtin_fmaf(aw, 1.0f - 9.6e-7f, -ew);
                            const float rhi = __builtin_fmaf(aw, 1.0f + 9.6e-7f, ew);
                            bool valid = true;
                            if (!FULL)
                                valid = cw0 + rg + rr < qlim[b];
                            // histogram bin of the squared upper bound 1 - rlo:
                            // 128 (1 + 1.5e-7 - rlo) in one rounding (7.6e-6
                            // at most, inside the 1.5e-7 margin), counted
                            // below kHistCap; rlo above 1 (var or p not
                            // matching V, |rho| clamped to 1) lands in bin 0.
                            // The pair's own row may be counted at its bound
                            // or not at all: tau stays an upper bound
                            const float binf = __builtin_fmaf(rlo, -(float)kBins,
                                                              (float)kBins * (1.0f + 1.5e-7f));
                            // (one atomic per pair, adding zero when not
                            // counted: no branch between the pairs)
                            {
                                const bool in = valid && binf < (float)kBins * kHistCap;
                                const int bin = __builtin_amdgcn_fmed3f(binf, 0.0f, 127.0f);
                                if (kWide)
                                    atomicAdd(&sHist[qcol[b] * kRow + bin], in ? 1u : 0u);
                                else
                                    atomicAdd(&sHist[qcol[b] * kRow + (bin >> 1)],
                                              in ? ((bin & 1) ? 65536u : 1u) : 0u);
                            }
                            // code = floor(256 * squared lower bound) with the
                            // bound 1 - rhi - 2.3e-7 (the 2.3e-7 covers the
                            // rounding of this fma), 255 if invalid
                            const float cf = __builtin_fmaf(-256.0f, rhi, 255.99994f);
                            const uint32_t code =
                                valid ? (uint32_t)__builtin_amdgcn_fmed3f(cf, 0.0f, 255.0f) : 255u;
                            word |= code << (8 * rr);
                        }
                        words[(4 * rb + g) * 2 + b] = word;
                    }
                }
            uint4 *cdst = (uint4 *)(A.codes + code_word(A, blk, ct, wid, lane));
#pragma unroll
            for (int i = 0; i < 4; i++)
                cdst[i] = make_uint4(words[4 * i], words[4 * i + 1], words[4 * i + 2],
                                     words[4 * i + 3]);
        };
        if (cw0 + kCW <= lim_min)
            epilogue(std::true_type{});
        else
            epilogue(std::false_type{});
    }
    __syncthreads();
    if (kWide) {
        for (int i = tid; i < kQT * kBins; i += kScreenThreads) {
            const int q = i / kBins, bin = i % kBins;
            const int64_t s = s0 + q;
            const uint32_t h = sHist[q * kRow + bin];
            if (s < A.nq && h)
                atomicAdd(&A.hist[s * kBins + bin], (int32_t)h);
        }
    } else {
        for (int i = tid; i < kQT * kBins / 2; i += kScreenThreads) {
            const int q = i / (kBins / 2), wd = i % (kBins / 2);
            const int64_t s = s0 + q;
            const uint32_t h = sHist[q * kRow + wd];
            if (s < A.nq && h) {
                const int bin = 2 * wd;
                if (h & 0xffffu)
                    atomicAdd(&A.hist[s * kBins + bin], (int32_t)(h & 0xffffu));
                if (h >> 16)
                    atomicAdd(&A.hist[s * kBins + bin + 1], (int32_t)(h >> 16));
            }
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
    const int32_t tl = A.n_tiles - 1 - (int32_t)blockIdx.x;
    const int64_t s0 = (int64_t)(A.tile0 + tl) * kQT;
    const int64_t blk = (int64_t)tl * A.n_seg + blockIdx.y;
    int32_t c_begin, c_end;
    tile_range(A, s0, (int32_t)blockIdx.y, &c_begin, &c_end);
    if (c_begin >= c_end)
        return;
    int32_t qcol[2], qlim[2], qself[2], qrowv[2];
    tile_queries(A, s0, lane, qcol, qlim, qself, qrowv);
    uint32_t qt[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
        const int64_t s = s0 + qcol[b];
        const float t = s < A.nq ? A.tau[s] : 0.0f;
        qt[b] = t >= 2.0f ? 255u : (uint32_t)(t * 256.0f);
    }
    const int rl = 4 * (lane >> 5);
    for (int32_t c0 = c_begin; c0 < c_end; c0 += kCT) {
        const int ct = (c0 - c_begin) / kCT;
        const int32_t cw0 = c0 + kCW * wid;
        if (cw0 >= c_end)
            continue;
        const uint4 *csrc = (const uint4 *)(A.codes + code_word(A, blk, ct, wid, lane));
        uint32_t words[16];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint4 w = csrc[i];
            words[4 * i] = w.x;
            words[4 * i + 1] = w.y;
            words[4 * i + 2] = w.z;
            words[4 * i + 3] = w.w;
        }
        uint32_t emit[2] = {0u, 0u};
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int g = 0; g < 4; g++)
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    const uint32_t word = words[(4 * rb + g) * 2 + b];
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        const int32_t j = cw0 + 32 * rb + 8 * g + rl + rr;
                        const uint32_t code = (word >> (8 * rr)) & 255u;
                        if (j < qlim[b] && code <= qt[b])
                            emit[b] |= 1u << ((4 * rb + g) * 4 + rr);
                    }
                }
#pragma unroll
        for (int b = 0; b < 2; b++) {
            // the query's own row (distance 0) is always listed
            const int32_t dr = qself[b] - cw0 - rl;
            if (dr >= 0 && dr < kCW - rl && (dr & 7) < 4 && qself[b] < qlim[b])
                emit[b] |= 1u << ((dr >> 3) * 4 + (dr & 3));
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
                    dst[p] = cw0 + rl + 8 * (r >> 2) + (r & 3);
                p++;
            }
        }
    }
}

template <int KT> static void launch_screen_kt(const ScreenArgs &A, size_t lds, hipStream_t st)
{
    const int64_t rounds = (A.n_seg + kXcds - 1) / kXcds;
    const int64_t blocks = (int64_t)kXcds * rounds * A.n_tiles;
    CX_REQUIRE(blocks < ((int64_t)1 << 31), "screen grid exceeds 2^31 workgroups");
    const dim3 grid((unsigned)blocks);
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

// ---- exact scoring -------------------------------------------------------------
// Listed candidates (rescore) and the candidates of hard queries are scored
// exactly in FP32 with compensated arithmetic: a Dot2 dot (TwoProduct +
// TwoSum per term) kept as a double-float, the kernel value, numerator,
// division and square root in double-float (df.hpp), and each query keeps its
// top-k in the registers of one wavefront (WaveTopK).  Every kernel here runs
// its wavefronts independently: no workgroup barriers.

// wavefronts per workgroup: one, so that a slot frees as soon as its query
// is scored (the C2 call 1.8% faster with two than with four, 0.6% faster
// again with one)
constexpr int kWaves = 1;
constexpr int kRsThreads = 64 * kWaves;

// the scoring kernels give each XCD its own range of work
constexpr int64_t kXcdQueries = 256; // consecutive sorted queries per XCD chunk


// one scored pair: distance as a double-float (h, l), candidate (dataset
// index), error bound
struct Ent {
    float h, l;
    int32_t j;
    float e;
};
using HardEntry = Ent;

struct RescoreArgs {
    const float *v;       // n x m tree order
    const float *raw;     // n x d tree order
    const float *p;       // n tree order
    const int64_t *inv_order;
    const int64_t *qsorted;
    const int64_t *qperm; // sorted position -> original query position
    const int32_t *qrow;
    const int32_t *cnt;   // listed candidates per sorted query (per qlist entry if set)
    const int32_t *list;  // cap slots per sorted query (per qlist entry if list_by_ix)
    const int32_t *qlist; // optional: rescore only these sorted positions
    int64_t nlist;
    int32_t cap;          // list slots per query; a larger count is not listed
    int32_t list_by_ix;
    const float *vnorm;   // |V_i| rounded up, dataset order
    int64_t n, nq;
    int32_t d, m, kernel, pred, k;
    float var;
    float c_dot;          // dot error bound per |V_q||V_j| (dot_error_coeff)
    float e_kern;         // double-float kernel value error bound (kernel_error)
    float inv_ls_hi[kMaxD], inv_ls_lo[kMaxD];
    int64_t *out_idx;
    float *out_dist, *out_err;
};

// Error of the compensated dot per |V_q||V_j| (Cauchy-Schwarz bounds
// sum|V_q[c] V_j[c]|): each lane runs Dot2 (Ogita, Rump and Oishi 2005) over
// at most n = ceil(m/8) terms in each of its chains, whose unrounded result
// s + c is within gamma_n^2 sum|products| (gamma_n = n u / (1 - n u),
// u = 2^-24); the chains and the eight lanes are then added as double-floats
// (four additions, each within 2^-44 of the magnitudes added).
static inline float dot_error_coeff(int32_t m)
{
    const double u = 1.0 / 16777216.0;
    const double n = (double)((m + 7) / 8);
    const double g = n * u / (1.0 - n * u);
    return (float)((g * g + 4.0 * std::ldexp(1.0, -44)) * 1.01);
}

// Error of the double-float kernel value: d^2 from TwoSum differences times
// 1/l as a double-float (relative (d + 2) 2^-45), df_exp within 2^-43
// relative plus the argument's error |x| (d + 2) 2^-45 (RBF: max x e^-x =
// 1/e; Matern-5/2: a = sqrt(5 d^2) within (d + 3) 2^-46, the polynomial and
// product within 4 2^-44, and max (1 + a + a^2/3)(2 + a) e^-a < 2.6), the
// product by var 2^-46: both kernels within |var| ((1.3 d + 4) 2^-45 + 8 2^-44).
static inline float kernel_error(float var, int32_t d)
{
    return (float)(std::fabs((double)var) *
                   ((1.3 * d + 4.0) * std::ldexp(1.0, -45) + 8.0 * std::ldexp(1.0, -44)) * 1.01);
}

// The residual distance from the double-float dot V_q.V_j and a bound on that
// dot's error: kernel value, numerator, division and square root in
// double-float.  Returns the distance and a bound on its error (err, which
// also covers rounding the distance to FP32).
// Error: e_num = dot_err + e_kern + 2^-44 (|K| + |dot|) (the subtraction);
// the square root of max(p_q p_j, 1e-18) and the division add 2^-42 |rho|;
// s = 1 - |rho| adds 2^-43; then |sqrt(s) - sqrt(s')| <= min(sqrt(e_s),
// e_s / (sqrt(s) + sqrt(max(s - e_s, 0)))), plus 2^-43 sqrt(s) for df_sqrt.
// var * k(x_q, x_j) in double-float from raw coordinates (within A.e_kern)
__device__ __forceinline__ df kernel_value_df(const RescoreArgs &A, const float *xq,
                                              const float *xj)
{
    const int32_t d = A.d;
    df d2 = {0.0f, 0.0f};
    for (int c = 0; c < d; c++) {
        const df diff = two_sum(xq[c], -xj[c]);
        const df s = df_mul(diff, df{A.inv_ls_hi[c], A.inv_ls_lo[c]});
        d2 = df_add(d2, df_mul(s, s));
    }
    df kv;
    if (A.kernel == 0) {
        kv = df_exp(df_mul_f(d2, -0.5f));
    } else {
        const df a = df_sqrt(df_mul_f(d2, 5.0f));
        constexpr DfConst third = split_const(1.0 / 3.0);
        const df poly =
            df_add(df_add(df{1.0f, 0.0f}, a), df_mul(df_mul(a, a), df{third.hi, third.lo}));
        kv = df_mul(poly, df_exp(df_neg(a)));
    }
    return df_mul_f(kv, A.var);
}

__device__ __forceinline__ df finish_distance_at(const RescoreArgs &A, const float *xq, float pq,
                                                 const float *xj, float pj, df dot, float dot_err,
                                                 float *err)
{
    const df kv = kernel_value_df(A, xq, xj);
    const df num = df_sub(kv, dot);
    df pp = two_prod(pq, pj);
    // floor max(p_q p_j, 1e-18) with 1e-18 as a double-float pair (the
    // FP32 literal alone is 1.0000000458e-18); hi + lo is within 8e-34 of it
    constexpr float kFloorHi = 0x1.2725dep-60f, kFloorLo = -0x1.c5b78ap-85f;
    if (pp.hi < kFloorHi || (pp.hi == kFloorHi && pp.lo < kFloorLo))
        pp = {kFloorHi, kFloorLo};
    const df den = df_sqrt(pp);
    df rho = df_div(num, den);
    if (rho.hi < 0.0f)
        rho = df_neg(rho);
    if (rho.hi > 1.0f || (rho.hi == 1.0f && rho.lo > 0.0f)) // |rho| clamped to 1
        rho = {1.0f, 0.0f};
    const df sv = df_sub(df{1.0f, 0.0f}, rho);
    const df dist = df_sqrt(sv);
    // the bound itself in FP32 with hardware reciprocal and root (about
    // 1 ulp each), covered by the 1.01 and 1.001 factors
    const float e_num = dot_err + A.e_kern +
                        5.7e-14f * (__builtin_fabsf(kv.hi) + __builtin_fabsf(dot.hi));
    const float e_s = e_num * __builtin_amdgcn_rcpf(den.hi) * 1.001f + rho.hi * 2.3e-13f +
                      1.2e-13f + 1e-37f;
    const float sd = dist.hi, s_hi = sv.hi;
    const float lo = s_hi - e_s > 0.0f ? __builtin_amdgcn_sqrtf(s_hi - e_s) * 0.999f : 0.0f;
    const float e_sq = __builtin_amdgcn_sqrtf(e_s) * 1.001f;
    float e_d = sd > 0.0f ? e_s * __builtin_amdgcn_rcpf(sd + lo) * 1.001f : e_sq;
    e_d = e_d < e_sq ? e_d : e_sq;
    *err = (e_d + sd * 1.2e-13f) * 1.01f + sd * 6e-8f;
    return dist;
}

// the same for tree rows qt and jt
__device__ __forceinline__ df finish_distance(const RescoreArgs &A, int64_t qt, int64_t jt, df dot,
                                              float dot_err, float *err)
{
    return finish_distance_at(A, A.raw + qt * A.d, A.p[qt], A.raw + jt * A.d, A.p[jt], dot,
                              dot_err, err);
}

// One Dot2 step: TwoProduct a*b, TwoSum into s, the two errors into c.
__device__ __forceinline__ void dot2_step(float a, float b, float &s, float &c)
{
    const float ph = a * b;
    const float pl = __builtin_fmaf(a, b, -ph);
    const float t = s + ph;
    const float bb = t - s;
    c += ((s - (t - bb)) + (ph - bb)) + pl;
    s = t;
}

// two Dot2 chains of a lane, then the eight lanes of the group, as one
// double-float
__device__ __forceinline__ df dot2_reduce(float s0, float c0, float s1, float c1)
{
    df x = df_add(two_sum(s0, c0), two_sum(s1, c1));
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        const df y = {__shfl_xor(x.hi, o, 64), __shfl_xor(x.lo, o, 64)};
        x = df_add(x, y);
    }
    return x;
}

// row[col .. col + 3] as a float4, zero when col >= m (m % 4 == 0): the load
// itself is unconditional (column 0 when past m) so that it stays one
// 16-byte load rather than four predicated ones
__device__ __forceinline__ float4 load_col4(const float *__restrict__ row, int32_t col, int32_t m)
{
    const bool in = col < m;
    const float4 v = *(const float4 *)(row + (in ? col : 0));
    return in ? v : make_float4(0.f, 0.f, 0.f, 0.f);
}

// orders the wavefront's own LDS accesses (no workgroup barrier)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// entries order by (distance hi, lo, index); the sentinel is last
__device__ __forceinline__ bool ent_lt(const Ent &a, const Ent &b)
{
    return a.h < b.h || (a.h == b.h && (a.l < b.l || (a.l == b.l && a.j < b.j)));
}

__device__ __forceinline__ Ent ent_none()
{
    return Ent{__builtin_huge_valf(), 0.0f, 0x7fffffff, 0.0f};
}

__device__ __forceinline__ Ent ent_shfl(const Ent &a, int src)
{
    return Ent{__shfl(a.h, src, 64), __shfl(a.l, src, 64), __shfl(a.j, src, 64),
               __shfl(a.e, src, 64)};
}

// compare-exchange with lane ^ mask: this lane keeps the smaller of the two
// when take_min, else the larger
__device__ __forceinline__ void ent_xchg(Ent &a, int mask, bool take_min)
{
    const Ent o = Ent{__shfl_xor(a.h, mask, 64), __shfl_xor(a.l, mask, 64),
                      __shfl_xor(a.j, mask, 64), __shfl_xor(a.e, mask, 64)};
    if (ent_lt(o, a) == take_min)
        a = o;
}

__device__ __forceinline__ void ent_order(Ent &lo, Ent &hi)
{
    if (ent_lt(hi, lo)) {
        const Ent t = lo;
        lo = hi;
        hi = t;
    }
}

// ascending bitonic sort of one entry per lane across the wavefront
__device__ __forceinline__ void sort64(Ent &a, int lane)
{
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const bool up = size == 64 || (lane & size) == 0;
            ent_xchg(a, stride, ((lane & stride) == 0) == up);
        }
}

// The k smallest entries seen (k <= 64 KB), sorted, held by one wavefront:
// rank r = lane + 64 b in p[b].  KB + 1 must be a power of two.
template <int KB> struct WaveTopK {
    Ent p[KB];

    __device__ __forceinline__ void init()
    {
#pragma unroll
        for (int b = 0; b < KB; b++)
            p[b] = ent_none();
    }

    // offer one entry per lane (ent_none() for none): entries that do not
    // rank ahead of the current last kept entry are dropped; if any remain,
    // the batch is sorted and merged (bitonic merge of the kept entries with
    // the reversed batch, the lower 64 KB kept)
    __device__ __forceinline__ void offer(Ent c, int lane)
    {
        const Ent last = ent_shfl(p[KB - 1], 63);
        const bool in = ent_lt(c, last);
        if (!__any(in))
            return;
        if (!in)
            c = ent_none();
        sort64(c, lane);
        Ent q[KB + 1];
#pragma unroll
        for (int b = 0; b < KB; b++)
            q[b] = p[b];
        q[KB] = ent_shfl(c, 63 - lane);
#pragma unroll
        for (int sb = (KB + 1) / 2; sb >= 1; sb >>= 1)
#pragma unroll
            for (int b = 0; b < KB + 1; b++)
                if ((b & sb) == 0)
                    ent_order(q[b], q[b + sb]);
#pragma unroll
        for (int stride = 32; stride >= 1; stride >>= 1)
#pragma unroll
            for (int b = 0; b < KB + 1; b++)
                ent_xchg(q[b], stride, (lane & stride) == 0);
#pragma unroll
        for (int b = 0; b < KB; b++)
            p[b] = q[b];
    }

    // the first k entries: rank r to `put(r, entry)`
    template <typename F> __device__ __forceinline__ void emit(int k, int lane, F put) const
    {
#pragma unroll
        for (int b = 0; b < KB; b++) {
            const int r = lane + 64 * b;
            if (r < k)
                put(r, p[b]);
        }
    }
};

__device__ __forceinline__ void put_output(const RescoreArgs &A, int64_t pos_out, int r,
                                           const Ent &e)
{
    const int64_t o = pos_out * A.k + r;
    const bool ok = e.h < __builtin_huge_valf();
    A.out_idx[o] = ok ? (int64_t)e.j : -1;
    A.out_dist[o] = ok ? e.h + e.l : 3.40282347e38f;
    if (A.out_err)
        A.out_err[o] = ok ? e.e : 0.0f;
}

// query rows staged per wavefront (floats, zero past m)
constexpr int kStageM = 512;

__device__ __forceinline__ bool staged_rank(int32_t m) { return (m & 3) == 0 && m <= kStageM; }

__device__ __forceinline__ void stage_row(const float *__restrict__ src, int32_t m, float *dst,
                                          int lane)
{
    if (staged_rank(m))
        for (int c = lane; c < kStageM; c += 64)
            dst[c] = c < m ? src[c] : 0.0f;
}

// Dots of one wavefront's batch of up to 64 candidates (tree rows jt, one per
// lane) against the staged query row (rank <= 256, m % 4 == 0): lane group g
// (eight lanes) takes candidates g, g + 8, ..., g + 56, each lane covering
// columns 4 sub + 32 t (t < 8) as float4 pieces in two Dot2 chains.  The next
// candidate's row is loaded while the current one is multiplied, so a batch
// waits on memory about once.  Dots go to (s_dh, s_dl)[e]; error bound
// c_dot |V_q||V_j|.
__device__ __forceinline__ void batch_dots_256(const float *__restrict__ v, int32_t m,
                                               const float *vq, int32_t jt, int n_in, float *s_dh,
                                               float *s_dl, int lane)
{
    const int grp = lane >> 3, sub = lane & 7;
    int32_t jts[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
        jts[u] = __shfl(jt, grp + 8 * u, 64);
    float4 cur[8], nxt[8];
    if (grp < n_in) {
        const float *row = v + (int64_t)jts[0] * m;
#pragma unroll
        for (int t = 0; t < 8; t++)
            cur[t] = load_col4(row, 4 * sub + 32 * t, m);
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int e = grp + 8 * u;
        if (u < 7 && e + 8 < n_in) {
            const float *row = v + (int64_t)jts[u + 1] * m;
#pragma unroll
            for (int t = 0; t < 8; t++)
                nxt[t] = load_col4(row, 4 * sub + 32 * t, m);
        }
        if (e < n_in) {
            float s0 = 0.0f, c0 = 0.0f, s1 = 0.0f, c1 = 0.0f;
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const float4 a = *(const float4 *)(vq + 4 * sub + 32 * t);
                dot2_step(a.x, cur[t].x, s0, c0);
                dot2_step(a.y, cur[t].y, s1, c1);
                dot2_step(a.z, cur[t].z, s0, c0);
                dot2_step(a.w, cur[t].w, s1, c1);
            }
            const df x = dot2_reduce(s0, c0, s1, c1);
            if (sub == 0) {
                s_dh[e] = x.hi;
                s_dl[e] = x.lo;
            }
        }
        if (u < 7) {
#pragma unroll
            for (int t = 0; t < 8; t++)
                cur[t] = nxt[t];
        }
    }
}

// Dots of a batch for the other ranks: the staged query row (rank <= 512,
// m % 4 == 0) one candidate per lane group at a time in 128-column steps, or
// the query row read from memory (any rank, columns sub + 8 t).
__device__ __forceinline__ void batch_dots_other(const float *__restrict__ v, int32_t m,
                                                 const float *vq, const float *vq_mem,
                                                 int32_t jt, int n_in, float *s_dh, float *s_dl,
                                                 int lane)
{
    const int grp = lane >> 3, sub = lane & 7;
    const bool staged = staged_rank(m);
#pragma unroll 1
    for (int u = 0; u < 64; u += 8) {
        const int e = u + grp;
        const int32_t je = __shfl(jt, e, 64);
        if (e >= n_in)
            continue;
        const float *row = v + (int64_t)je * m;
        float s0 = 0.0f, c0 = 0.0f, s1 = 0.0f, c1 = 0.0f;
        if (staged) {
#pragma unroll 1
            for (int32_t cb = 0; cb < m; cb += 128) {
                float4 x[4];
#pragma unroll
                for (int t = 0; t < 4; t++)
                    x[t] = load_col4(row, cb + 4 * sub + 32 * t, m);
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const float4 a = *(const float4 *)(vq + cb + 4 * sub + 32 * t);
                    dot2_step(a.x, x[t].x, s0, c0);
                    dot2_step(a.y, x[t].y, s1, c1);
                    dot2_step(a.z, x[t].z, s0, c0);
                    dot2_step(a.w, x[t].w, s1, c1);
                }
            }
        } else {
            for (int32_t col = sub; col < m; col += 8)
                dot2_step(vq_mem[col], row[col], s0, c0);
        }
        const df x = dot2_reduce(s0, c0, s1, c1);
        if (sub == 0) {
            s_dh[e] = x.hi;
            s_dl[e] = x.lo;
        }
    }
}

// One wavefront scores candidates for one query: `count` candidates, the
// list `lst` (dataset indices) when given, else the range base + [0, count),
// those at or past `lim` skipped; batches of 64 offered to `top`.
template <int KB>
__device__ __forceinline__ void score_query(const RescoreArgs &A, int64_t qt, int32_t qrow,
                                            float qn, const int32_t *lst, int64_t base,
                                            int64_t count, int64_t lim, const float *vq,
                                            float *s_dh, float *s_dl, WaveTopK<KB> &top, int lane)
{
    const int32_t m = A.m;
    const bool fast = (m & 3) == 0 && m <= 256;
    for (int64_t c0 = 0; c0 < count; c0 += 64) {
        const int n_in = count - c0 < 64 ? (int)(count - c0) : 64;
        const bool valid = lane < n_in;
        const int64_t j = valid ? (lst ? (int64_t)lst[c0 + lane] : base + c0 + lane) : 0;
        const int32_t jt = valid ? (int32_t)A.inv_order[j] : 0;
        if (fast)
            batch_dots_256(A.v, m, vq, jt, n_in, s_dh, s_dl, lane);
        else
            batch_dots_other(A.v, m, vq, A.v + qt * m, jt, n_in, s_dh, s_dl, lane);
        wave_sync();
        Ent c = ent_none();
        if (valid && j < lim) {
            c.j = (int32_t)j;
            if (j == qrow) {
                c.h = 0.0f;
                c.l = 0.0f;
                c.e = 0.0f;
            } else {
                const df dd = finish_distance(A, qt, jt, df{s_dh[lane], s_dl[lane]},
                                              A.c_dot * qn * A.vnorm[j], &c.e);
                c.h = dd.hi;
                c.l = dd.lo;
            }
        }
        wave_sync();
        top.offer(c, lane);
    }
}

// ---- listed candidates ---------------------------------------------------------
// One wavefront per query (from its XCD's chunk of sorted queries) scores its
// list.  Queries whose list overflowed kCap are left to the hard kernels.

// four wavefronts per SIMD for one kept entry per lane (k <= 64): the row
// gathers are latency-bound; three for the wider top-k
template <int KB>
__global__ void __launch_bounds__(kRsThreads, KB == 1 ? 4 : 3) rescore_kernel(RescoreArgs A)
{
    __shared__ __attribute__((aligned(16))) float s_vq[kWaves][kStageM];
    __shared__ float s_dh[kWaves][64], s_dl[kWaves][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t m = A.m, k = A.k;
    // one query per wavefront: workgroup b runs on XCD x = b % kXcds, and the
    // XCD's workgroups take its chunks x, x + 8, ... of kXcdQueries
    // consecutive sorted queries in order, so its L2 holds the candidate rows
    // of neighbouring queries (a plain grid: other streams' kernels interleave)
    const int x = blockIdx.x % kXcds;
    {
        const int64_t local = (int64_t)(blockIdx.x / kXcds) * kWaves + w;
        const int64_t ix = (x + kXcds * (local / kXcdQueries)) * kXcdQueries + local % kXcdQueries;
        if (ix >= (A.qlist ? A.nlist : A.nq))
            return;
        const int64_t s = A.qlist ? (int64_t)A.qlist[ix] : ix;
        const int32_t count = A.cnt[ix];
        if (count > A.cap)
            return; // listed again with more slots, or scored by the hard kernels
        const int32_t qrow = A.qrow[s];
        const int64_t qt = A.inv_order[qrow];
        wave_sync();
        stage_row(A.v + qt * m, m, s_vq[w], lane);
        wave_sync();
        WaveTopK<KB> top;
        top.init();
        score_query<KB>(A, qt, qrow, A.vnorm[qrow], A.list + (A.list_by_ix ? ix : s) * A.cap, 0,
                        count, A.n, s_vq[w],
                        s_dh[w], s_dl[w], top, lane);
        const int64_t pos = A.qperm[s];
        top.emit(k, lane, [&](int r, const Ent &e) { put_output(A, pos, r, e); });
    }
}

// ---- hard queries ------------------------------------------------------------
// A query whose list overflowed kCap (a wide screen interval, e.g. tiny p) is
// scored over all its valid candidates.  One wavefront item = one hard query
// x a kHSeg-candidate segment, scored as a list (the segment's rows stay in
// the XCD's L2 for the other queries); a merge kernel combines the segments'
// top-k lists.

constexpr int kHSeg = 512;   // candidates per item

struct HardArgs {
    RescoreArgs r;
    const int32_t *hard;  // sorted positions of the hard queries
    int32_t n_hard;
    int32_t h0, hn;       // this batch: hard queries [h0, h0 + hn)
    int32_t nseg;         // segments per query (ceil(n / kHSeg))
    HardEntry *part;      // hn x nseg x k segment top-k lists
};

__device__ __forceinline__ int64_t hard_limit(const RescoreArgs &A, int64_t s)
{
    const int64_t qid = A.qsorted[s];
    // Vecchia bound on the original index (negative -> unbounded)
    return A.pred ? (qid < 0 ? A.n : (qid < A.n ? qid : A.n)) : A.n;
}

template <int KB>
__global__ void __launch_bounds__(kRsThreads, 3) hard_tile_kernel(HardArgs H)
{
    __shared__ __attribute__((aligned(16))) float s_vq[kWaves][kStageM];
    __shared__ float s_dh[kWaves][64], s_dl[kWaves][64];
    const RescoreArgs &A = H.r;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t m = A.m, k = A.k;
    // one item per wavefront: workgroup b runs on XCD x = b % kXcds, and the
    // XCD takes the segments x, x + 8, ..., each for every query of the batch
    // before the next, so its L2 holds the segment's rows
    const int x = blockIdx.x % kXcds;
    {
        const int64_t local = (int64_t)(blockIdx.x / kXcds) * kWaves + w;
        const int64_t seg_l = x + kXcds * (local / H.hn);
        if (seg_l >= H.nseg)
            return;
        const int32_t hl = (int32_t)(local % H.hn), seg = (int32_t)seg_l;
        const int64_t s = H.hard[H.h0 + hl];
        const int64_t lim = hard_limit(A, s);
        const int64_t beg = (int64_t)seg * kHSeg;
        if (beg >= lim)
            return; // past the query's candidates (never read)
        const int64_t end = beg + kHSeg < lim ? beg + kHSeg : lim;
        const int32_t qrow = A.qrow[s];
        const int64_t qt = A.inv_order[qrow];
        wave_sync();
        stage_row(A.v + qt * m, m, s_vq[w], lane);
        wave_sync();
        WaveTopK<KB> top;
        top.init();
        score_query<KB>(A, qt, qrow, A.vnorm[qrow], nullptr, beg, end - beg, lim, s_vq[w],
                        s_dh[w], s_dl[w], top, lane);
        HardEntry *out = H.part + ((int64_t)hl * H.nseg + seg) * k;
        top.emit(k, lane, [&](int r, const Ent &e) { out[r] = e; });
    }
}

// Merge the segment lists of each hard query of the batch (one wavefront per
// query); write its row.
template <int KB>
__global__ void __launch_bounds__(kRsThreads) hard_merge_kernel(HardArgs H)
{
    const RescoreArgs &A = H.r;
    const int lane = threadIdx.x & 63;
    const int32_t k = A.k;
    const int32_t stride = gridDim.x * kWaves;
    for (int32_t hl = blockIdx.x * kWaves + (threadIdx.x >> 6); hl < H.hn; hl += stride) {
        const int64_t s = H.hard[H.h0 + hl];
        const int64_t limit = hard_limit(A, s);
        const int32_t nl = (int32_t)((limit + kHSeg - 1) / kHSeg);
        WaveTopK<KB> top;
        top.init();
        const int64_t total = (int64_t)nl * k;
        const HardEntry *src = H.part + (int64_t)hl * H.nseg * k;
        for (int64_t e0 = 0; e0 < total; e0 += 64) {
            const int64_t e = e0 + lane;
            top.offer(e < total ? src[e] : ent_none(), lane);
        }
        const int64_t pos = A.qperm[s];
        top.emit(k, lane, [&](int r, const Ent &e) { put_output(A, pos, r, e); });
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

// ordered compaction over the overflowed-query list: the queries whose new
// list still overflows (cnt2 > cap) go to dst; the count to n_out
__global__ void over_flag_kernel(const int32_t *__restrict__ cnt2, int64_t n, int32_t cap,
                                 int32_t *flag)
{
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h < n)
        flag[h] = cnt2[h] > cap ? 1 : 0;
}

__global__ void over_scatter_kernel(const int32_t *__restrict__ flag,
                                    const int32_t *__restrict__ pos, int64_t n,
                                    const int32_t *__restrict__ src, int32_t *dst,
                                    int32_t *n_out)
{
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= n)
        return;
    if (flag[h])
        dst[pos[h]] = src[h];
    if (h == n - 1)
        *n_out = pos[h] + flag[h];
}

// after the main rescore: the overflowed queries' final counts
__global__ void over_commit_kernel(const int32_t *__restrict__ hard,
                                   const int32_t *__restrict__ cnt2, int64_t n, int32_t *cnt)
{
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h < n)
        cnt[hard[h]] = cnt2[h];
}

// Query telemetry of a mixed call: pairs screened (every valid candidate of
// every query), candidates listed for rescoring, hard queries and their
// pairs, the longest list.
__global__ void mixed_telemetry_kernel(const int64_t *__restrict__ qsorted,
                                       const int32_t *__restrict__ cnt, int64_t nq, int64_t n,
                                       int32_t pred, int32_t hard_above, int64_t *telem)
{
    __shared__ unsigned long long s_acc[5];
    __shared__ int s_max;
    if (threadIdx.x < 5)
        s_acc[threadIdx.x] = 0ull;
    if (threadIdx.x == 0)
        s_max = 0;
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nq) {
        const int64_t qid = qsorted[s];
        const int64_t lim = pred ? (qid < 0 ? n : (qid < n ? qid : n)) : n;
        const int32_t c = cnt[s];
        const bool hard = c > hard_above; // scored over all its candidates
        atomicAdd(&s_acc[0], (unsigned long long)lim);
        atomicAdd(&s_acc[1], hard ? 0ull : (unsigned long long)c);
        atomicAdd(&s_acc[2], hard ? 1ull : 0ull);
        atomicAdd(&s_acc[3], hard ? (unsigned long long)lim : 0ull);
        atomicAdd(&s_acc[4], hard ? (unsigned long long)c : 0ull);
        atomicMax(&s_max, c);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd((unsigned long long *)&telem[CX_QT_DISTANCE_EVALS], s_acc[0]);
        atomicAdd((unsigned long long *)&telem[CX_QT_LISTED], s_acc[1]);
        atomicAdd((unsigned long long *)&telem[CX_QT_HARD_QUERIES], s_acc[2]);
        atomicAdd((unsigned long long *)&telem[CX_QT_HARD_PAIRS], s_acc[3]);
        atomicAdd((unsigned long long *)&telem[CX_QT_HARD_LISTED], s_acc[4]);
        atomicMax((unsigned long long *)&telem[CX_QT_MAX_LIST], (unsigned long long)s_max);
    }
}

// ---- exact-product screen of overflowed queries --------------------------------
//
// A query whose list overflowed (tiny p makes w = 1/sqrt(p_q p_j) large, so
// the BF16 interval is wide) gets a second certified screen before the
// rescore when such queries are few.  Each V row is split into six int8
// slices under a power-of-two scale 2^e, |V_i| < 2^e:
//   V_i = 2^e (A1 2^-6 + A2 2^-13 + ... + A6 2^-41) + r_i,  |A| <= 64, |r_i| <= 2^(e-42)
// (each slice is the nearest integer of the scaled remainder; every step is
// exact in FP32).  The int8 MFMA sums slice products exactly in int32; with
// G_t = sum_{a+b=t} A_a(q) . A_b(j) for t = 2..7,
//   V_q . V_j = 2^(e_q+e_j) sum_t 2^-(12+7(t-2)) G_t + E,
//   |E| <= 2^(e_q+e_j-42) m (5.04 [pairs a+b >= 8] + 2.01 [r terms]
//                            + 1.6 [double-float sum of the six terms]),
// |A_a . A_b| <= 2^12 m.  With the kernel value in double-float
// (kernel_value_df, the rescore's evaluation) the numerator k - dot carries
// a certified error; |rho| = |k - dot| / sqrt(p_q p_j) follows in FP32 with
// its relative rounding, which bounds d^2 = 1 - |rho| from both sides.  Pass 1 stores the squared
// bounds of every valid pair; tau, the k-th smallest upper bound rounded up
// to a 2^-19 bin edge, bounds the k-th exact d^2; pass 2 lists every pair
// whose lower bound is <= tau.  A query whose list
// now fits kCap is rescored with the others; one that still overflows takes
// the exhaustive path.

constexpr int kSlices = 6;
constexpr int kHsBins = 1024;          // bins per level of the tau selection
constexpr int kHsTile = 128;           // candidates per tile (four wavefronts x 32)
constexpr int kHsSeg = 2 * kHsTile;    // candidates per workgroup
constexpr int kHsRow = 48;             // LDS bytes per staged 32-byte slice row (padded)
constexpr int64_t kHsBudget = (int64_t)512 << 20; // bytes of stored bounds per batch
// list counters one 128-byte line apart: each takes one atomic per listed
// pair, and counters sharing a line would serialize on one L2 channel
constexpr int kHsCntStride = 32;
// overflowed queries above which the exact-product screen takes the wide
// tiles (hs_bounds_wide_kernel, next K step prefetched); fewer run beside the
// rescore in the small tiles, which slow it less
constexpr int32_t kHsWideMin = 1024;
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Slices and row exponent of dataset row blockIdx.x (tree row inv_order[.]).
__global__ void slice_kernel(const float *__restrict__ v, const int64_t *__restrict__ inv_order,
                             int64_t n, int32_t m, int32_t mp, int8_t *__restrict__ sl,
                             int32_t *__restrict__ sexp)
{
    __shared__ float s_max[4];
    const int64_t row = blockIdx.x;
    if (row >= n)
        return;
    const float *src = v + inv_order[row] * (int64_t)m;
    float mx = 0.0f;
    for (int32_t c = threadIdx.x; c < m; c += blockDim.x)
        mx = __builtin_fmaxf(mx, __builtin_fabsf(src[c]));
    for (int o = 32; o > 0; o >>= 1)
        mx = __builtin_fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0)
        s_max[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = __builtin_fmaxf(__builtin_fmaxf(s_max[0], s_max[1]), __builtin_fmaxf(s_max[2], s_max[3]));
    int e = 0;
    if (mx > 0.0f)
        (void)__builtin_frexpf(mx, &e); // mx = f 2^e, f in [0.5, 1): |V| < 2^e
    if (threadIdx.x == 0)
        sexp[row] = e;
    int8_t *dst = sl + row * (int64_t)kSlices * mp;
    for (int32_t c = threadIdx.x; c < mp; c += blockDim.x) {
        float r = c < m ? __builtin_ldexpf(src[c], -e) : 0.0f;
        float s = 64.0f;
#pragma unroll
        for (int a = 0; a < kSlices; a++) {
            const float y = r * s;
            const float q = __builtin_rintf(y);
            dst[(int64_t)a * mp + c] = (int8_t)(int)q;
            r = y - q; // exact: the fractional part of y
            s = 128.0f;
        }
    }
}

struct HsArgs {
    RescoreArgs r;         // tree-order data (finish_distance), queries, lists
    const int8_t *sl;      // npad x kSlices x mp, dataset order
    const int32_t *sexp;   // npad row exponents
    const float *isp;      // n: 1/sqrt(p), dataset order
    const int32_t *hard;   // sorted positions of the overflowed queries
    int32_t h0, hn;        // this batch: hard[h0 .. h0 + hn)
    int32_t mp;
    float2 *bounds;        // hn x n squared (lower, upper) bounds
    float *tau;            // hn
    int32_t *hcnt;         // hn counters, kHsCntStride ints apart
    int32_t *list;         // cap slots per sorted query, or per position (list_by_pos)
    int32_t cap, list_by_pos;
    int32_t *cnt;          // new count per position of the overflowed-query list
    int64_t *telem;        // query counters or null
};

// double-float of an int32 (exact: both parts are integers below 2^24)
__device__ __forceinline__ df df_of_int(int32_t g)
{
    const float hi = (float)g;
    return {hi, (float)(g - (int32_t)hi)};
}

// Pass 1a.  Workgroup (segment x, query tile y) = 32 overflowed queries (MFMA
// columns) x a segment of 256 candidates in tiles of 128 (four wavefronts x
// 32 MFMA rows); stores each valid pair's dot as a double-float.
__global__ void __launch_bounds__(256, 2) hs_bounds_kernel(HsArgs H)
{
    __shared__ __attribute__((aligned(16))) int8_t sA[kSlices * kHsTile * kHsRow];
    __shared__ __attribute__((aligned(16))) int8_t sB[kSlices * 32 * kHsRow];
    __shared__ int32_t sQrow[32], sEj[kHsTile];
    const RescoreArgs &A = H.r;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t n = A.n;
    const int32_t mp = H.mp;
    const int hq = (int)blockIdx.y * 32 + (lane & 31); // this lane's query column
    const bool qlive = hq < H.hn;
    const int64_t s = qlive ? H.hard[H.h0 + hq] : 0;
    const int64_t lim = qlive ? hard_limit(A, s) : 0;
    const int32_t qrow = qlive ? A.qrow[s] : 0;
    const int32_t eq = H.sexp[qrow];
    if (tid < 32) {
        const int h = (int)blockIdx.y * 32 + tid;
        sQrow[tid] = h < H.hn ? A.qrow[H.hard[H.h0 + h]] : 0;
    }
    // the segment ends at the largest limit of the 32 queries
    int64_t lmax = lim;
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
        const int64_t u = __shfl_xor(lmax, o, 64);
        lmax = u > lmax ? u : lmax;
    }
    __shared__ int64_t s_lmax[4];
    if (lane == 0)
        s_lmax[w] = lmax;
    __syncthreads();
    lmax = s_lmax[0];
    for (int i = 1; i < 4; i++)
        lmax = s_lmax[i] > lmax ? s_lmax[i] : lmax;
    const int64_t seg0 = (int64_t)blockIdx.x * kHsSeg;
    const int64_t seg1 = seg0 + kHsSeg < lmax ? seg0 + kHsSeg : lmax;
    for (int64_t c0 = seg0; c0 < seg1; c0 += kHsTile) {
        __syncthreads(); // the previous tile's epilogue still reads sEj
        if (tid < kHsTile) {
            const int64_t j = c0 + tid < n ? c0 + tid : 0;
            sEj[tid] = H.sexp[j];
        }
        v16i G[6];
#pragma unroll
        for (int t = 0; t < 6; t++)
            G[t] = v16i{};
        // candidate slices: 128 rows x 6 slices x 2 halves of 16 bytes,
        // query slices: 32 rows x 6 x 2
        for (int32_t kc = 0; kc < mp; kc += 32) {
            __syncthreads();
            for (int u = tid; u < kHsTile * kSlices * 2; u += 256) {
                const int row = u / (kSlices * 2), rem = u % (kSlices * 2);
                const int a = rem >> 1, half = rem & 1;
                const int64_t j = c0 + row;
                const int64_t jr = j < n ? j : 0;
                const uint4 val =
                    *(const uint4 *)(H.sl + (jr * kSlices + a) * mp + kc + 16 * half);
                *(uint4 *)(sA + (a * kHsTile + row) * kHsRow + 16 * half) = val;
            }
            for (int u = tid; u < 32 * kSlices * 2; u += 256) {
                const int row = u / (kSlices * 2), rem = u % (kSlices * 2);
                const int a = rem >> 1, half = rem & 1;
                const uint4 val =
                    *(const uint4 *)(H.sl + ((int64_t)sQrow[row] * kSlices + a) * mp + kc +
                                     16 * half);
                *(uint4 *)(sB + (a * 32 + row) * kHsRow + 16 * half) = val;
            }
            __syncthreads();
            // any consistent K order works for A and B: the sums are exact
            const int koff = 16 * (lane >> 5);
            v4i av[kSlices], bv[kSlices];
#pragma unroll
            for (int a = 0; a < kSlices; a++) {
                av[a] = *(const v4i *)(sA + (a * kHsTile + 32 * w + (lane & 31)) * kHsRow + koff);
                bv[a] = *(const v4i *)(sB + (a * 32 + (lane & 31)) * kHsRow + koff);
            }
#pragma unroll
            for (int t = 2; t <= 7; t++)
#pragma unroll
                for (int a = 1; a <= 6; a++) {
                    const int b = t - a;
                    if (b >= 1 && b <= 6)
                        G[t - 2] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[a - 1], bv[b - 1],
                                                                         G[t - 2], 0, 0, 0);
                }
        }
        // register r of lane l: candidate row 8(r/4) + 4(l/32) + r%4 of the
        // wavefront's 32, query column l%32
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int rw = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int64_t j = c0 + 32 * w + rw;
            if (!qlive || j >= lim)
                continue;
            const int32_t ej = sEj[32 * w + rw];
            const int es = eq + ej;
            df dot = {0.0f, 0.0f};
#pragma unroll
            for (int t = 0; t < 6; t++) {
                const df g = df_of_int(G[t][r]);
                const int sh = es - (12 + 7 * t);
                dot = df_add(dot, df{__builtin_ldexpf(g.hi, sh), __builtin_ldexpf(g.lo, sh)});
            }
            H.bounds[(int64_t)hq * n + j] = make_float2(dot.hi, dot.lo);
        }
    }
}

// Pass 1a for many overflowed queries: the same pairs and sums with larger
// workgroup tiles, 128 queries x 64 candidates (eight wavefronts, two
// candidate rows of 32 by four query columns of 32), so each staged
// candidate slice serves 128 queries instead of 32 and a K step stages 4.5
// bytes per pair instead of 7.5.  Segments of 256 candidates.
constexpr int kHwQ = 128, kHwC = 64, kHwSeg = 256;

__global__ void __launch_bounds__(512, 1) hs_bounds_wide_kernel(HsArgs H)
{
    __shared__ __attribute__((aligned(16))) int8_t sA[kSlices * kHwC * kHsRow];
    __shared__ __attribute__((aligned(16))) int8_t sB[kSlices * kHwQ * kHsRow];
    __shared__ int32_t sQrow[kHwQ], sEj[kHwC];
    __shared__ int64_t s_lmax[8];
    const RescoreArgs &A = H.r;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int cb = w & 1, qb = w >> 1; // candidate row block, query column block
    const int64_t n = A.n;
    const int32_t mp = H.mp;
    const int hq = (int)blockIdx.y * kHwQ + 32 * qb + (lane & 31);
    const bool qlive = hq < H.hn;
    const int64_t s = qlive ? H.hard[H.h0 + hq] : 0;
    const int64_t lim = qlive ? hard_limit(A, s) : 0;
    const int32_t qrow = qlive ? A.qrow[s] : 0;
    const int32_t eq = H.sexp[qrow];
    if (tid < kHwQ) {
        const int h = (int)blockIdx.y * kHwQ + tid;
        sQrow[tid] = h < H.hn ? A.qrow[H.hard[H.h0 + h]] : 0;
    }
    // the segment ends at the largest limit of the 128 queries
    int64_t lmax = lim;
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
        const int64_t u = __shfl_xor(lmax, o, 64);
        lmax = u > lmax ? u : lmax;
    }
    if (lane == 0)
        s_lmax[w] = lmax;
    __syncthreads();
    lmax = s_lmax[0];
    for (int i = 1; i < 8; i++)
        lmax = s_lmax[i] > lmax ? s_lmax[i] : lmax;
    const int64_t seg0 = (int64_t)blockIdx.x * kHwSeg;
    const int64_t seg1 = seg0 + kHwSeg < lmax ? seg0 + kHwSeg : lmax;
    for (int64_t c0 = seg0; c0 < seg1; c0 += kHwC) {
        __syncthreads(); // the previous tile's epilogue still reads sEj
        if (tid < kHwC) {
            const int64_t j = c0 + tid < n ? c0 + tid : 0;
            sEj[tid] = H.sexp[j];
        }
        v16i G[6];
#pragma unroll
        for (int t = 0; t < 6; t++)
            G[t] = v16i{};
        // staged 16-byte pieces: candidates 64 rows x 6 slices x 2 halves
        // (768: pieces tid and tid + 512 when below), queries 128 x 12 (1536:
        // tid, tid + 512, tid + 1024); the next K step's loads fly while this
        // one is multiplied
#define HW_GA(u) ((((c0 + (u) / (kSlices * 2)) < n ? c0 + (u) / (kSlices * 2) : (int64_t)0) *      \
                       kSlices + ((u) % (kSlices * 2) >> 1)) * mp + 16 * ((u) & 1))
#define HW_LA(u) ((((u) % (kSlices * 2) >> 1) * kHwC + (u) / (kSlices * 2)) * kHsRow + 16 * ((u) & 1))
#define HW_GB(u) (((int64_t)sQrow[(u) / (kSlices * 2)] * kSlices + ((u) % (kSlices * 2) >> 1)) * mp + \
                  16 * ((u) & 1))
#define HW_LB(u) ((((u) % (kSlices * 2) >> 1) * kHwQ + (u) / (kSlices * 2)) * kHsRow + 16 * ((u) & 1))
        const bool a1 = tid + 512 < kHwC * kSlices * 2;
        uint4 pa0, pa1 = make_uint4(0u, 0u, 0u, 0u), pb0, pb1, pb2;
#define HW_LOAD(o)                                                                                \
    do {                                                                                          \
        pa0 = *(const uint4 *)(H.sl + HW_GA(tid) + (o));                                          \
        if (a1)                                                                                   \
            pa1 = *(const uint4 *)(H.sl + HW_GA(tid + 512) + (o));                                \
        pb0 = *(const uint4 *)(H.sl + HW_GB(tid) + (o));                                          \
        pb1 = *(const uint4 *)(H.sl + HW_GB(tid + 512) + (o));                                    \
        pb2 = *(const uint4 *)(H.sl + HW_GB(tid + 1024) + (o));                                   \
    } while (0)
        HW_LOAD(0);
        for (int32_t kc = 0; kc < mp; kc += 32) {
            __syncthreads();
            *(uint4 *)(sA + HW_LA(tid)) = pa0;
            if (a1)
                *(uint4 *)(sA + HW_LA(tid + 512)) = pa1;
            *(uint4 *)(sB + HW_LB(tid)) = pb0;
            *(uint4 *)(sB + HW_LB(tid + 512)) = pb1;
            *(uint4 *)(sB + HW_LB(tid + 1024)) = pb2;
            if (kc + 32 < mp)
                HW_LOAD(kc + 32);
            __syncthreads();
            // any consistent K order works for A and B: the sums are exact
            const int koff = 16 * (lane >> 5);
            v4i av[kSlices], bv[kSlices];
#pragma unroll
            for (int a = 0; a < kSlices; a++) {
                av[a] = *(const v4i *)(sA + (a * kHwC + 32 * cb + (lane & 31)) * kHsRow + koff);
                bv[a] = *(const v4i *)(sB + (a * kHwQ + 32 * qb + (lane & 31)) * kHsRow + koff);
            }
#pragma unroll
            for (int t = 2; t <= 7; t++)
#pragma unroll
                for (int a = 1; a <= 6; a++) {
                    const int b = t - a;
                    if (b >= 1 && b <= 6)
                        G[t - 2] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[a - 1], bv[b - 1],
                                                                         G[t - 2], 0, 0, 0);
                }
        }
#undef HW_LOAD
#undef HW_GA
#undef HW_LA
#undef HW_GB
#undef HW_LB
        // register r of lane l: candidate row 8(r/4) + 4(l/32) + r%4 of the
        // wavefront's 32, query column l%32
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int rw = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int64_t j = c0 + 32 * cb + rw;
            if (!qlive || j >= lim)
                continue;
            const int32_t ej = sEj[32 * cb + rw];
            const int es = eq + ej;
            df dot = {0.0f, 0.0f};
#pragma unroll
            for (int t = 0; t < 6; t++) {
                const df g = df_of_int(G[t][r]);
                const int sh = es - (12 + 7 * t);
                dot = df_add(dot, df{__builtin_ldexpf(g.hi, sh), __builtin_ldexpf(g.lo, sh)});
            }
            H.bounds[(int64_t)hq * n + j] = make_float2(dot.hi, dot.lo);
        }
    }
}

// Pass 1b, one thread per (query, candidate): the interval of d^2 from the
// stored dot, in place.  |rho| = |k - dot| w with w = min(1/sqrt(p_q p_j),
// 1e9): the numerator in double-float (its error: the dot's, the kernel
// value's and the subtraction's, as in finish_distance), w and the product in
// FP32 (relative error < 6e-7 together); d^2 = 1 - |rho|, with 1.2e-7 for
// the rounding of 1 - r.
__global__ void __launch_bounds__(256) hs_interval_kernel(HsArgs H)
{
    const RescoreArgs &A = H.r;
    const int64_t n = A.n;
    const int h = (int)blockIdx.y;
    const int64_t s = H.hard[H.h0 + h];
    const int64_t lim = hard_limit(A, s);
    const int32_t qrow = A.qrow[s];
    const float *xq = A.raw + A.inv_order[qrow] * A.d;
    const float qisp = H.isp[qrow];
    const int32_t eq = H.sexp[qrow];
    const float m9 = 9.0f * (float)A.m;
    float2 *row = H.bounds + (int64_t)h * n;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < lim;
         j += (int64_t)gridDim.x * blockDim.x) {
        const float2 dv = row[j];
        const df dot = {dv.x, dv.y};
        const float e_dot = __builtin_ldexpf(m9, eq + H.sexp[j] - 42) + 1e-37f;
        const df kv = kernel_value_df(A, xq, A.raw + A.inv_order[j] * A.d);
        const df num = df_sub(kv, dot);
        const float en = __builtin_fabsf(num.lo) + e_dot + A.e_kern +
                         5.7e-14f * (__builtin_fabsf(kv.hi) + __builtin_fabsf(dot.hi));
        const float wq = __builtin_fminf(qisp * H.isp[j], 1e9f);
        const float ar = __builtin_fabsf(num.hi) * wq;
        const float er = en * wq * 1.000001f + ar * 6e-7f;
        const float rlo = __builtin_fminf(__builtin_fmaxf(ar - er, 0.0f), 1.0f);
        const float rhi = __builtin_fminf(ar + er, 1.0f);
        const float lb2 = __builtin_fmaxf((1.0f - rhi) - 1.2e-7f, 0.0f);
        const float ub2 = (1.0f - rlo) + 1.2e-7f;
        row[j] = make_float2(lb2, ub2);
    }
}

// tau per query (one workgroup each): the k-th smallest upper bound, rounded
// up to a bin edge, from two histogram levels over the stored bounds -- 1024
// bins of width 2^-9 over [0, 2), then 1024 sub-bins of the bin holding the
// k-th (width 2^-19).  Binning is exact: v * 2^9 and (v - b/2^9) * 2^19 are
// exact in FP32 for v in bin b, and tau = (1024 b1 + b2 + 1) 2^-19 is
// representable, so every bound counted at or below the chosen sub-bin is
// <= tau.  +inf when the query has fewer than k pairs.
constexpr int kHsTauThreads = 1024; // one bin per thread in hs_find

__device__ __forceinline__ int hs_find(const int32_t *hist, int32_t need, int32_t *below, int *s_w,
                                       int32_t *s_tot)
{
    // first bin where the running count reaches need; *below = count before it
    static_assert(kHsBins == kHsTauThreads, "one bin per thread");
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int32_t c = hist[tid];
    int32_t incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t u = __shfl_up(incl, o, 64);
        if (lane >= o)
            incl += u;
    }
    if (lane == 63)
        s_tot[wid] = incl;
    __syncthreads();
    int32_t off = 0;
    for (int w = 0; w < wid; w++)
        off += s_tot[w];
    incl += off;
    const int32_t excl = incl - c;
    if (tid == 0)
        *s_w = kHsBins;
    __syncthreads();
    if (excl < need && incl >= need) {
        *s_w = tid;
        *below = excl;
    }
    __syncthreads();
    const int r = *s_w;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kHsTauThreads) hs_tau_kernel(HsArgs H, int32_t k)
{
    __shared__ int32_t s_hist[kHsBins];
    __shared__ int32_t s_tot[kHsTauThreads / 64], s_below;
    __shared__ int s_bin;
    const RescoreArgs &A = H.r;
    const int h = (int)blockIdx.x, tid = threadIdx.x;
    const int64_t n = A.n;
    const int64_t lim = hard_limit(A, H.hard[H.h0 + h]);
    const float2 *bd = H.bounds + (int64_t)h * n;
    for (int i = tid; i < kHsBins; i += kHsTauThreads)
        s_hist[i] = 0;
    __syncthreads();
    for (int64_t j = tid; j < lim; j += kHsTauThreads) {
        const float v = bd[j].y;
        const int b = (v >= 0.0f && v < 2.0f) ? (int)(v * 512.0f) : kHsBins - 1;
        atomicAdd(&s_hist[b], 1);
    }
    __syncthreads();
    if (tid == 0)
        s_below = 0;
    const int b1 = hs_find(s_hist, k, &s_below, &s_bin, s_tot);
    if (b1 >= kHsBins - 1) { // fewer than k pairs, or the k-th is past 2 - 2^-9
        if (tid == 0)
            H.tau[h] = __builtin_inff();
        return;
    }
    const int32_t below1 = s_below;
    const float lo1 = (float)b1 * (1.0f / 512.0f);
    for (int i = tid; i < kHsBins; i += kHsTauThreads)
        s_hist[i] = 0;
    __syncthreads();
    for (int64_t j = tid; j < lim; j += kHsTauThreads) {
        const float v = bd[j].y;
        if (v >= lo1 && v < lo1 + (1.0f / 512.0f)) {
            const int b = min((int)((v - lo1) * 524288.0f), kHsBins - 1);
            atomicAdd(&s_hist[b], 1);
        }
    }
    __syncthreads();
    const int b2 = hs_find(s_hist, k - below1, &s_below, &s_bin, s_tot);
    if (tid == 0)
        H.tau[h] = (float)(1024 * b1 + min(b2, kHsBins - 1) + 1) * (1.0f / 524288.0f);
}

// Pass 2: list every valid pair with lower bound <= tau (slots past cap are
// counted, not written).
__global__ void hs_list_kernel(HsArgs H)
{
    const RescoreArgs &A = H.r;
    const int64_t n = A.n;
    const int h = (int)blockIdx.y;
    const int64_t s = H.hard[H.h0 + h];
    const int64_t lim = hard_limit(A, s);
    const float tau = H.tau[h];
    int32_t *dst = H.list + (H.list_by_pos ? (int64_t)(H.h0 + h) : s) * H.cap;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < lim;
         j += (int64_t)gridDim.x * blockDim.x) {
        if (H.bounds[(int64_t)h * n + j].x <= tau) {
            const int32_t p = atomicAdd(&H.hcnt[h * kHsCntStride], 1);
            if (p < H.cap)
                dst[p] = (int32_t)j;
        }
    }
}

// The new counts, per position of the overflowed-query list.
__global__ void hs_commit_kernel(HsArgs H)
{
    const int h = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (h < H.hn)
        H.cnt[H.h0 + h] = H.hcnt[h * kHsCntStride];
    if (h == 0 && H.telem)
        atomicAdd((unsigned long long *)&H.telem[CX_QT_RESCREENED], (unsigned long long)H.hn);
}

// The overflowed queries hard[0 .. n_hard) get the exact-product screen;
// the ones whose new list fits kCap become listed queries (cnt rewritten).
// Lists of cap slots, per sorted query or (list_by_pos) per position of
// hard; the second form lists queries already counted as rescreened.
// after_first runs on the host once the first batch's hs_bounds_kernel is
// queued (the caller queues work of its own there).
static void rescreen_hard(TreeState &t, const RescoreArgs &ra, int32_t *hard, int32_t n_hard,
                          int32_t k, int32_t *cnt, int32_t *list, int32_t cap,
                          int32_t list_by_pos, hipStream_t st,
                          const std::function<void()> &after_first)
{
    const int64_t n = t.n;
    if (!t.slices_ready) {
        const int64_t npad = ((n + kCT - 1) / kCT) * kCT + kCT;
        t.vsl.alloc((size_t)npad * kSlices * t.mp);
        t.vexp.alloc(npad);
        CX_HIP(hipMemsetAsync(t.vsl.ptr, 0, t.vsl.bytes(), st));
        CX_HIP(hipMemsetAsync(t.vexp.ptr, 0, t.vexp.bytes(), st));
        slice_kernel<<<(unsigned)n, 256, 0, st>>>(t.v.ptr, t.inv_order.ptr, n, t.m, t.mp,
                                                  t.vsl.ptr, t.vexp.ptr);
        CX_HIP(hipGetLastError());
        t.slices_ready = true;
    }
    const int64_t per_query = n * (int64_t)sizeof(float2);
    const int32_t batch =
        (int32_t)std::max<int64_t>(1, std::min<int64_t>(n_hard, kHsBudget / per_query));
    const size_t b_bounds = (size_t)batch * per_query;
    const size_t b_small = (size_t)batch * sizeof(int32_t);
    const size_t need = b_bounds + b_small * (1 + kHsCntStride) + 1024;
    if (t.hs_buf.count < need) {
        CX_HIP(hipStreamSynchronize(st));
        t.hs_buf.alloc(need);
    }
    uint8_t *base = t.hs_buf.ptr;
    HsArgs H{};
    H.r = ra;
    H.sl = t.vsl.ptr;
    H.sexp = t.vexp.ptr;
    H.isp = t.isp_ds.ptr;
    H.hard = hard;
    H.mp = t.mp;
    H.bounds = (float2 *)base;
    H.tau = (float *)(base + b_bounds);
    H.hcnt = (int32_t *)(base + b_bounds + b_small);
    H.list = list;
    H.cnt = cnt;
    H.cap = cap;
    H.list_by_pos = list_by_pos;
    H.telem = list_by_pos ? nullptr : t.telem;
    for (int32_t h0 = 0; h0 < n_hard; h0 += batch) {
        H.h0 = h0;
        H.hn = std::min(batch, n_hard - h0);
        CX_HIP(hipMemsetAsync(H.hcnt, 0, (size_t)H.hn * kHsCntStride * sizeof(int32_t), st));
        // x: candidate segments (up to 2^31 - 1), y: tiles of 32 queries
        const dim3 g1((unsigned)((n + kHsSeg - 1) / kHsSeg), (unsigned)((H.hn + 31) / 32));
        if (n_hard > kHsWideMin) {
            const dim3 gw((unsigned)((n + kHwSeg - 1) / kHwSeg),
                          (unsigned)((H.hn + kHwQ - 1) / kHwQ));
            hs_bounds_wide_kernel<<<gw, 512, 0, st>>>(H);
        } else {
            hs_bounds_kernel<<<g1, 256, 0, st>>>(H);
        }
        if (h0 == 0)
            after_first();
        const dim3 g2((unsigned)std::min<int64_t>((n + 255) / 256, 64), (unsigned)H.hn);
        hs_interval_kernel<<<g2, 256, 0, st>>>(H);
        hs_tau_kernel<<<(unsigned)H.hn, kHsTauThreads, 0, st>>>(H, k);
        hs_list_kernel<<<g2, 256, 0, st>>>(H);
        hs_commit_kernel<<<(unsigned)((H.hn + 255) / 256), 256, 0, st>>>(H);
        CX_HIP(hipGetLastError());
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
    t.mp = ((t.m + kMpAlign - 1) / kMpAlign) * kMpAlign;
    // rows padded past the last candidate tile a screen workgroup may read
    const int64_t npad = ((n + kCT - 1) / kCT) * kCT + kCT;
    // the screen's fragment layout pads the columns to whole K chunks
    t.mps = ((t.m + kKC - 1) / kKC) * kKC;
    t.vsw.alloc((size_t)npad * t.mps * 2);
    CX_HIP(hipMemsetAsync(t.vsw.ptr, 0, t.vsw.bytes(), st));
    t.vnorm.alloc(npad);
    CX_HIP(hipMemsetAsync(t.vnorm.ptr, 0, t.vnorm.bytes(), st));
    t.slices_ready = false; // the exact-product screen's slices follow V
    split_kernel<<<(unsigned)n, 256, 0, st>>>(t.v.ptr, t.inv_order.ptr, n, t.m, t.mps, t.vsw.ptr,
                                              t.vnorm.ptr);
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

// pair codes kept per launch group of query tiles (bytes);
// COVERTREEX_HIP_CODE_BUDGET overrides it (the tests split small calls into
// several groups with it)
constexpr int64_t kCodeBudget = (int64_t)4 << 30;
static int64_t code_budget()
{
    if (const char *e = getenv("COVERTREEX_HIP_CODE_BUDGET")) {
        const long long v = atoll(e);
        if (v > 0)
            return (int64_t)v;
    }
    return kCodeBudget;
}
// segment top-k lists of the hard queries per launch batch (bytes)
constexpr int64_t kPartBudget = (int64_t)512 << 20;

// Candidates per screen segment for nq queries.  Each XCD streams a
// segment's candidate fragments (hi and lo, 4 * mp bytes per row) for every
// query tile while they sit in its L2 (4 MiB), so a segment holds about
// 2 MiB of them: 2 MiB / (4 mp) rows, a power of two in [kSegMin, 4096]
// (mp 64 and 128: 4096, 256: 2048, 512: 1024); it is halved (down to
// kSegMin) while the grid would hold fewer than 512 workgroups, so small
// calls still spread over the CUs.  These are the picks of the tuner
// (covertreex_amd/tuner.py) on the workload shapes.  A tree's own choice
// (cx_tree_set_screen_segment) overrides the rule.
int32_t screen_segment(const TreeState &t, int64_t nq)
{
    if (t.screen_seg > 0)
        return t.screen_seg;
    const int64_t rows = ((int64_t)2 << 20) / (4 * (int64_t)std::max(t.mps, 1));
    int32_t seg = kSegMin;
    while (seg * 2 <= rows && seg * 2 <= 4096)
        seg *= 2;
    const int64_t tiles = (nq + kQT - 1) / kQT;
    while (seg > kSegMin && tiles * ((t.n + seg - 1) / seg) < 512)
        seg /= 2;
    return seg;
}

bool screen_segment_valid(int32_t seg)
{
    return seg == 0 || (seg >= kSegMin && seg <= kSegMax && seg % kCT == 0);
}

void knn_mixed(TreeState &t, const int64_t *d_queries, int64_t nq, int32_t k, int32_t predecessor,
               int64_t *d_idx, float *d_dist, float *d_err)
{
    CX_REQUIRE(k >= 1 && k <= kMixMaxK, "mixed-precision kNN supports 1 <= k <= 256");
    CX_REQUIRE(t.d <= kMaxD, "mixed-precision kNN supports d <= 16");
    const int32_t seg_len = screen_segment(t, nq);
    t.last_seg = seg_len;
    CX_REQUIRE(t.n < ((int64_t)1 << 31) - 4 * (int64_t)kSegMax,
               "mixed-precision kNN supports fewer than 2^31 - 65536 points");
    // the filter's grid y counts segments (at most 65535)
    CX_REQUIRE(t.n <= (int64_t)65535 * seg_len,
               "mixed-precision kNN supports at most 65535 segments of candidates");
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
    const size_t o_hard2 = take(sizeof(int32_t) * nq);
    const size_t o_cnt2 = take(sizeof(int32_t) * nq);
    const size_t o_hard3 = take(sizeof(int32_t) * nq);
    const size_t o_cnt3 = take(sizeof(int32_t) * nq);
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
    // overflowed queries still overflowing after the exact-product screen,
    // and the new counts per position of the overflowed-query list
    int32_t *hard2 = (int32_t *)(ws + o_hard2), *cnt2 = (int32_t *)(ws + o_cnt2);
    // the same after the second listing (counts per position of hard2)
    int32_t *hard3 = (int32_t *)(ws + o_hard3), *cnt3 = (int32_t *)(ws + o_cnt3);
    // ctr[0]: the overflowed-query count; ctr[1]: those still overflowing;
    // ctr[2]: those still overflowing the second listing
    int32_t *n_hard_d = ctr, *n_hard2_d = ctr + 1, *n_hard3_d = ctr + 2;

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
    sa.vsw = t.vsw.ptr;
    sa.x = t.x_ds.ptr; // scaled by 1/lengthscale
    sa.isp = t.isp_ds.ptr;
    sa.vnorm = t.vnorm.ptr;
    sa.qsorted = qs;
    sa.qrow = qrow;
    sa.n = (int32_t)n;
    sa.nq = nq;
    sa.d = t.d;
    sa.mp = t.mps;
    sa.kernel = t.kernel;
    sa.pred = predecessor ? 1 : 0;
    sa.k = k;
    sa.var = t.var;
    // |computed V.V - exact| <= cD |V_q||V_j|: the hi/lo split drops at most
    // 3.1 * 2^-18 of each product, the hi.hi FP32 accumulation adds gamma_mps,
    // the cross terms (2^-8 smaller) gamma_2mps, and the final sum one rounding.
    const double u = 1.0 / 16777216.0;
    const double g1 = (t.mps + 16) * u / (1.0 - (t.mps + 16) * u);
    const double g2 = (2 * t.mps + 16) * u / (1.0 - (2 * t.mps + 16) * u);
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
    sa.seg = seg_len;
    sa.n_seg = (int32_t)((n + seg_len - 1) / seg_len);
    // per wavefront: its candidates' coordinates, 1/sqrt(p) and norms; the
    // query coordinates
    const size_t dyn_lds =
        sizeof(float) * ((size_t)kScreenWaves * (t.d + 2) * kCW + (size_t)t.d * kQT);
    // launch groups of query tiles whose pair codes fit the code budget
    const int64_t tile_bytes = (int64_t)sa.n_seg * seg_len * kQT;
    const int32_t group =
        (int32_t)std::max<int64_t>(1, std::min<int64_t>(n_tiles, code_budget() / tile_bytes));
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
    ra.vnorm = t.vnorm.ptr;
    ra.raw = t.raw.ptr;
    ra.p = t.p.ptr;
    ra.inv_order = t.inv_order.ptr;
    ra.qsorted = qs;
    ra.qperm = perm;
    ra.qrow = qrow;
    ra.cnt = cnt;
    ra.list = list;
    ra.cap = kCap;
    ra.list_by_ix = 0;
    ra.n = n;
    ra.nq = nq;
    ra.d = t.d;
    ra.m = t.m;
    ra.kernel = t.kernel;
    ra.pred = predecessor ? 1 : 0;
    ra.k = k;
    ra.var = t.var;
    ra.c_dot = dot_error_coeff(t.m);
    ra.e_kern = kernel_error(t.var, t.d);
    for (int c = 0; c < t.d; c++) {
        const double il = 1.0 / (double)t.ls[c];
        ra.inv_ls_hi[c] = (float)il;
        ra.inv_ls_lo[c] = (float)(il - (double)(float)il);
    }
    ra.out_idx = d_idx;
    ra.out_dist = d_dist;
    ra.out_err = d_err;

    // hard queries: ordered list (read back once) ...
    hard_flag_kernel<<<qb, 256, 0, st>>>(cnt, nq, flag);
    CX_HIP(hipcub::DeviceScan::ExclusiveSum(ws + o_tmp, scan_bytes, flag, pos, (int)nq, st));
    hard_scatter_kernel<<<qb, 256, 0, st>>>(flag, pos, nq, hard, n_hard_d);
    if (!t.host_word)
        t.host_word = acquire_host_word();
    CX_HIP(hipMemcpyAsync(t.host_word, n_hard_d, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (!t.sync_event)
        t.sync_event = acquire_event();
    hipEvent_t ev = t.sync_event;
    CX_HIP(hipEventRecord(ev, st));
    // ... while the listed queries are rescored
    // top-k held as KB entries per lane: k <= 64 KB
    const int kb = k <= 64 ? 1 : k <= 192 ? 3 : 7;
    auto launch_rescore = [&](const RescoreArgs &r, int64_t count, hipStream_t sm) {
        // workgroups: each XCD gets ceil(chunks / 8) chunks of kXcdQueries queries
        const int64_t n_chunks = (count + kXcdQueries - 1) / kXcdQueries;
        const unsigned grid =
            (unsigned)(kXcds * ((n_chunks + kXcds - 1) / kXcds) * (kXcdQueries / kWaves));
        if (kb == 1)
            rescore_kernel<1><<<grid, kRsThreads, 0, sm>>>(r);
        else if (kb == 3)
            rescore_kernel<3><<<grid, kRsThreads, 0, sm>>>(r);
        else
            rescore_kernel<7><<<grid, kRsThreads, 0, sm>>>(r);
        CX_HIP(hipGetLastError());
    };
    CX_HIP(hipEventSynchronize(ev));
    const int32_t n_over = *t.host_word;
    int32_t n_hard = n_over, n_big = 0;
    if (n_over == 0) {
        launch_rescore(ra, nq, st);
        prof_mark(t, 6);
    } else {
        // The overflowed queries run on a second stream beside the rescore,
        // which skips them (disjoint output rows; everything they read is
        // ordered by ev).  They get the exact-product screen and are
        // rescored from their new lists; the ones still overflowing are
        // listed again with more slots, and only those past these are scored
        // exhaustively.  Their final counts are written after the rescore,
        // on the tree's stream.
        if (!t.aux_stream)
            t.aux_stream = acquire_stream(t.device);
        if (!t.join_event)
            t.join_event = acquire_event();
        const hipStream_t hs = t.aux_stream;
        StreamScope hs_scope(hs);
        CX_HIP(hipStreamWaitEvent(hs, ev, 0));
        const unsigned ob = (unsigned)((n_over + 255) / 256);
        {
            // the main rescore is queued right after the first exact-product
            // screen kernel: that kernel's workgroups reach the CUs first,
            // and the rest of this chain is queued while both run
            rescreen_hard(t, ra, hard, n_over, k, cnt2, list, kCap, 0, hs, [&] {
                launch_rescore(ra, nq, st);
                prof_mark(t, 6);
            });
            RescoreArgs sub = ra;
            sub.cnt = cnt2;
            sub.qlist = hard;
            sub.nlist = n_over;
            launch_rescore(sub, n_over, hs);
            over_flag_kernel<<<ob, 256, 0, hs>>>(cnt2, n_over, kCap, flag);
            CX_HIP(hipcub::DeviceScan::ExclusiveSum(ws + o_tmp, scan_bytes, flag, pos,
                                                    (int)n_over, hs));
            over_scatter_kernel<<<ob, 256, 0, hs>>>(flag, pos, n_over, hard, hard2, n_hard2_d);
            if (!t.host_word2)
                t.host_word2 = acquire_host_word();
            if (!t.sync_event2)
                t.sync_event2 = acquire_event();
            CX_HIP(hipMemcpyAsync(t.host_word2, n_hard2_d, sizeof(int32_t),
                                  hipMemcpyDeviceToHost, hs));
            CX_HIP(hipEventRecord(t.sync_event2, hs));
            CX_HIP(hipEventSynchronize(t.sync_event2));
            n_hard = *t.host_word2;
        }
        // the queries still overflowing are listed again into kCapBig slots
        // each (by position) and rescored from there; beyond those, or
        // when the lists would not fit kBigListBudget, exhaustively
        const int32_t *exhaustive = hard2;
        if (n_hard > 0 && (int64_t)n_hard * kCapBig * (int64_t)sizeof(int32_t) <= kBigListBudget) {
            n_big = n_hard;
            const size_t need = (size_t)n_big * kCapBig;
            if (t.list_big.count < need) {
                CX_HIP(hipStreamSynchronize(hs));
                t.list_big.alloc(need);
            }
            rescreen_hard(t, ra, hard2, n_big, k, cnt3, t.list_big.ptr, kCapBig, 1, hs, [] {});
            RescoreArgs big = ra;
            big.cnt = cnt3;
            big.qlist = hard2;
            big.nlist = n_big;
            big.list = t.list_big.ptr;
            big.cap = kCapBig;
            big.list_by_ix = 1;
            launch_rescore(big, n_big, hs);
            const unsigned bb = (unsigned)((n_big + 255) / 256);
            over_flag_kernel<<<bb, 256, 0, hs>>>(cnt3, n_big, kCapBig, flag);
            CX_HIP(hipcub::DeviceScan::ExclusiveSum(ws + o_tmp, scan_bytes, flag, pos,
                                                    (int)n_big, hs));
            over_scatter_kernel<<<bb, 256, 0, hs>>>(flag, pos, n_big, hard2, hard3, n_hard3_d);
            CX_HIP(hipMemcpyAsync(t.host_word2, n_hard3_d, sizeof(int32_t),
                                  hipMemcpyDeviceToHost, hs));
            CX_HIP(hipEventRecord(t.sync_event2, hs));
            CX_HIP(hipEventSynchronize(t.sync_event2));
            n_hard = *t.host_word2;
            exhaustive = hard3;
        }
        if (n_hard > 0) {
            HardArgs ha{};
            ha.r = ra;
            ha.hard = exhaustive;
            ha.n_hard = n_hard;
            ha.nseg = (int32_t)((n + kHSeg - 1) / kHSeg);
            const int64_t query_bytes = (int64_t)ha.nseg * k * sizeof(HardEntry);
            const int32_t batch = (int32_t)std::max<int64_t>(
                1, std::min<int64_t>(n_hard, kPartBudget / query_bytes));
            if (t.part.count < (size_t)(batch * query_bytes)) {
                CX_HIP(hipStreamSynchronize(hs));
                t.part.alloc((size_t)(batch * query_bytes));
            }
            ha.part = (HardEntry *)t.part.ptr;
            for (int32_t h0 = 0; h0 < n_hard; h0 += batch) {
                ha.h0 = h0;
                ha.hn = std::min(batch, n_hard - h0);
                // workgroups: each XCD gets ceil(nseg / 8) segments x hn queries
                const int64_t per_xcd = ((ha.nseg + kXcds - 1) / kXcds) * (int64_t)ha.hn;
                const unsigned tg = (unsigned)(kXcds * ((per_xcd + kWaves - 1) / kWaves));
                // one wavefront per query of the batch
                const unsigned mg =
                    (unsigned)std::min<int64_t>(((int64_t)ha.hn + kWaves - 1) / kWaves, 2048);
                if (kb == 1) {
                    hard_tile_kernel<1><<<tg, kRsThreads, 0, hs>>>(ha);
                    hard_merge_kernel<1><<<mg, kRsThreads, 0, hs>>>(ha);
                } else if (kb == 3) {
                    hard_tile_kernel<3><<<tg, kRsThreads, 0, hs>>>(ha);
                    hard_merge_kernel<3><<<mg, kRsThreads, 0, hs>>>(ha);
                } else {
                    hard_tile_kernel<7><<<tg, kRsThreads, 0, hs>>>(ha);
                    hard_merge_kernel<7><<<mg, kRsThreads, 0, hs>>>(ha);
                }
                CX_HIP(hipGetLastError());
            }
        }
        CX_HIP(hipEventRecord(t.join_event, hs));
        CX_HIP(hipStreamWaitEvent(st, t.join_event, 0));
        over_commit_kernel<<<ob, 256, 0, st>>>(hard, cnt2, n_over, cnt);
        if (n_big > 0)
            over_commit_kernel<<<(unsigned)((n_big + 255) / 256), 256, 0, st>>>(hard2, cnt3, n_big,
                                                                               cnt);
    }
    t.last_hard = n_hard;
    if (t.telem)
        mixed_telemetry_kernel<<<qb, 256, 0, st>>>(qs, cnt, nq, n, predecessor ? 1 : 0,
                                                   n_big > 0 ? kCapBig : kCap, t.telem);
    prof_mark(t, 7);
}

} // namespace cx
