// Euclidean kNN over the device cover tree.
//
// CX_KNN_REFERENCE_F32: single_knn_query (src/algo.rs:208-323) restated, one
//   wavefront per query: every root seeded into a candidate min-heap, best-first
//   pops, prune when the result heap is full and d - 2^(level+1) > kth, all
//   children pushed.  Both heaps follow std::collections::BinaryHeap's sift
//   rules so ties resolve as in the reference; the candidate heap lives in a
//   per-wavefront HBM buffer.  Predecessor mode compares node index < query
//   position, as batch_knn_query does (src/algo.rs:194).
// CX_KNN_EXACT: the k smallest FP32 fold distances, ties broken by the lower
//   index, found by a certified box-pruned sweep over Hilbert-ordered blocks
//   of 64 points (one wavefront per query; details above the kernel).
#include "cx_internal.hpp"
#include "hilbert.hpp"

#include <hipcub/hipcub.hpp>

namespace cx {

constexpr int kEuMaxD = 16;
constexpr int kEuMaxK = 64;

__device__ __forceinline__ float euclid_dist(const float *__restrict__ a,
                                             const float *__restrict__ b, int32_t d)
{
    // Euclidean::distance = sqrt of the in-order fold of squared differences
    float acc = 0.0f;
    for (int32_t c = 0; c < d; c++) {
        float df = a[c] - b[c];
        acc = acc + df * df;
    }
    return __builtin_sqrtf(acc);
}

struct Cand {
    float d;
    int32_t node;
};

// ---- candidate heap (Candidate: reversed order -> min-heap) ----------------
__device__ void cand_push(Cand *h, int64_t &len, Cand x)
{
    int64_t pos = len++;
    while (pos > 0) {
        int64_t parent = (pos - 1) / 2;
        if (!(x.d < h[parent].d)) // stop when x <= parent in reversed order
            break;
        h[pos] = h[parent];
        pos = parent;
    }
    h[pos] = x;
}

__device__ Cand cand_pop(Cand *h, int64_t &len)
{
    Cand top = h[0];
    Cand last = h[--len];
    if (len > 0) {
        int64_t end = len, pos = 0, child = 1;
        const int64_t lim = end >= 2 ? end - 2 : 0;
        while (end >= 2 && child <= lim) {
            child += (!(h[child + 1].d > h[child].d)) ? 1 : 0;
            h[pos] = h[child];
            pos = child;
            child = 2 * pos + 1;
        }
        if (child == end - 1) {
            h[pos] = h[child];
            pos = child;
        }
        while (pos > 0) {
            int64_t parent = (pos - 1) / 2;
            if (!(last.d < h[parent].d))
                break;
            h[pos] = h[parent];
            pos = parent;
        }
        h[pos] = last;
    }
    return top;
}

// ---- result heap (Neighbor: max-heap by distance) -----------------------------
__device__ void res_push(Cand *h, int &len, Cand x)
{
    int pos = len++;
    while (pos > 0) {
        int parent = (pos - 1) / 2;
        if (!(x.d > h[parent].d))
            break;
        h[pos] = h[parent];
        pos = parent;
    }
    h[pos] = x;
}

__device__ void res_pop(Cand *h, int &len)
{
    Cand last = h[--len];
    if (len == 0)
        return;
    int end = len, pos = 0, child = 1;
    const int lim = end >= 2 ? end - 2 : 0;
    while (end >= 2 && child <= lim) {
        child += (!(h[child].d > h[child + 1].d)) ? 1 : 0;
        h[pos] = h[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (child == end - 1) {
        h[pos] = h[child];
        pos = child;
    }
    while (pos > 0) {
        int parent = (pos - 1) / 2;
        if (!(last.d > h[parent].d))
            break;
        h[pos] = h[parent];
        pos = parent;
    }
    h[pos] = last;
}

__device__ void res_sort(Cand *h, int len)
{
    // into_sorted_vec: swap the max to the end and sift down
    for (int end = len - 1; end >= 1; end--) {
        Cand tmp = h[0];
        h[0] = h[end];
        h[end] = tmp;
        Cand hole = h[0];
        int pos = 0, child = 1;
        const int lim = end >= 2 ? end - 2 : 0;
        bool placed = false;
        while (end >= 2 && child <= lim) {
            child += (!(h[child].d > h[child + 1].d)) ? 1 : 0;
            if (!(hole.d < h[child].d)) {
                h[pos] = hole;
                placed = true;
                break;
            }
            h[pos] = h[child];
            pos = child;
            child = 2 * pos + 1;
        }
        if (placed)
            continue;
        if (child == end - 1 && hole.d < h[child].d) {
            h[pos] = h[child];
            pos = child;
        }
        h[pos] = hole;
    }
}

__global__ void __launch_bounds__(64)
    knn_euclid_ref_kernel(const float *__restrict__ pts, int32_t d, int64_t n,
                          const int32_t *__restrict__ roots, int64_t n_roots,
                          const int32_t *__restrict__ child_off,
                          const int32_t *__restrict__ child_list,
                          const int32_t *__restrict__ levels, const float *__restrict__ queries,
                          int64_t nq, int32_t k, int32_t pred, int64_t *out_idx, float *out_dist,
                          int32_t *counter, Cand *ws, int64_t cap)
{
    __shared__ Cand s_res[kEuMaxK + 1];
    __shared__ float s_q[kEuMaxD];
    __shared__ int s_qi;
    __shared__ Cand s_kids[64];
    const int lane = threadIdx.x;
    Cand *heap = ws + (int64_t)blockIdx.x * cap;
    for (;;) {
        if (lane == 0)
            s_qi = atomicAdd(counter, 1);
        __syncthreads();
        const int64_t qi = s_qi;
        __syncthreads();
        if (qi >= nq)
            break;
        for (int j = lane; j < k; j += 64) {
            out_idx[qi * k + j] = -1;
            out_dist[qi * k + j] = 3.40282347e38f;
        }
        if (lane < d)
            s_q[lane] = queries[qi * d + lane];
        __syncthreads();
        if (n == 0 || (pred && qi == 0))
            continue;
        int64_t hlen = 0;
        int rlen = 0;
        float kth = 3.40282347e38f;
        // seed every root in index order
        for (int64_t b = 0; b < n_roots; b += 64) {
            const int64_t ri = b + lane;
            float dd = 0.0f;
            int32_t r = 0;
            if (ri < n_roots) {
                r = roots[ri];
                dd = euclid_dist(s_q, pts + (int64_t)r * d, d);
                s_kids[lane] = Cand{dd, r};
            }
            __syncthreads();
            if (lane == 0) {
                const int cnt = (int)((n_roots - b) < 64 ? (n_roots - b) : 64);
                for (int u = 0; u < cnt; u++)
                    cand_push(heap, hlen, s_kids[u]);
            }
            __syncthreads();
        }
        // best-first (src/algo.rs:249-313); lane 0 owns both heaps
        for (;;) {
            __shared__ int s_node;
            __shared__ float s_nd;
            if (lane == 0) {
                int node = -1;
                float nd = 0.0f;
                while (hlen > 0) {
                    Cand c = cand_pop(heap, hlen);
                    const bool can_prune = pred ? rlen >= k : rlen == k;
                    if (can_prune) {
                        const int32_t lvl = levels[c.node];
                        const int32_t e = (int32_t)((uint32_t)lvl + 1u);
                        const float radius = e < -160 ? 0.0f
                                                      : (e > 127 ? __builtin_huge_valf()
                                                                 : __builtin_ldexpf(1.0f, e));
                        if (c.d - radius > kth)
                            continue;
                    }
                    const bool ok = pred ? (int64_t)c.node < qi : true;
                    if (ok) {
                        res_push(s_res, rlen, c);
                        if (rlen > k)
                            res_pop(s_res, rlen);
                        if (rlen == k)
                            kth = s_res[0].d;
                    }
                    node = c.node;
                    nd = c.d;
                    break;
                }
                s_node = node;
                s_nd = nd;
            }
            __syncthreads();
            const int node = s_node;
            __syncthreads();
            if (node < 0)
                break;
            // expand children in list order
            const int32_t c0 = child_off[node], c1 = child_off[node + 1];
            for (int32_t cb = c0; cb < c1; cb += 64) {
                const int32_t ci = cb + lane;
                if (ci < c1) {
                    const int32_t ch = child_list[ci];
                    s_kids[lane] = Cand{euclid_dist(s_q, pts + (int64_t)ch * d, d), ch};
                }
                __syncthreads();
                if (lane == 0) {
                    const int cnt = (c1 - cb) < 64 ? (c1 - cb) : 64;
                    for (int u = 0; u < cnt; u++)
                        cand_push(heap, hlen, s_kids[u]);
                }
                __syncthreads();
            }
        }
        if (lane == 0) {
            res_sort(s_res, rlen);
            for (int j = 0; j < rlen && j < k; j++) {
                out_idx[qi * k + j] = s_res[j].node;
                out_dist[qi * k + j] = s_res[j].d;
            }
        }
        __syncthreads();
    }
}

// ---- exact search over Hilbert blocks ------------------------------------------
//
// The points are sorted by Hilbert code and cut into blocks of 64 rows (one per
// lane) with an axis-aligned box each; 64 consecutive blocks form a superblock
// with its own box.  A box's lower bound folds max(lo - q, q - hi, 0)^2 over
// the axes in the order the distance folds (q - p)^2.  Subtraction, squaring
// and addition of non-negative terms are monotone under round-to-nearest, so
// the FP32 bound never exceeds the FP32 fold distance of any point in the box:
// a box whose bound exceeds the current k-th distance holds no candidate, and
// skipping it changes nothing.  The result is the k smallest (d^2, index)
// pairs of the FP32 fold distances, the same list a scan of every point makes.

constexpr int kEuBlock = 64;         // rows per block, one per lane
constexpr int kEuQueryWaves = 4;     // queries per workgroup, one per wave
constexpr int kEuXcds = 8;           // workgroup b runs on XCD b % 8
constexpr int32_t kEuNone = 0x7fffffff; // empty list slot: sorts after every point

__global__ void __launch_bounds__(256)
    eu_gather_kernel(const float *__restrict__ raw, const int64_t *__restrict__ perm, int64_t n,
                     int32_t d, int64_t npad, float *__restrict__ pts, int32_t *__restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad)
        return;
    if (i < n) {
        const int64_t src = perm[i];
        for (int32_t c = 0; c < d; c++)
            pts[i * d + c] = raw[src * d + c];
        idx[i] = (int32_t)src;
    } else {
        for (int32_t c = 0; c < d; c++)
            pts[i * d + c] = 0.0f;
        idx[i] = -1;
    }
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int32_t wave_min_i(int32_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const int32_t w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// Box of each group of 64 entries: with rows != nullptr the entries are the
// point rows of one block (valid where idx >= 0); otherwise they are 64 block
// boxes (in_lo / in_hi / in_min, axis-major over n_in) forming one superblock.
// Boxes are stored axis-major: lo[c * n_out + g].
__global__ void __launch_bounds__(64)
    eu_box_kernel(const float *__restrict__ rows, const int32_t *__restrict__ idx,
                  const float *__restrict__ in_lo, const float *__restrict__ in_hi,
                  const int32_t *__restrict__ in_min, int64_t n_in, int64_t n_out, int32_t d,
                  float *__restrict__ lo, float *__restrict__ hi, int32_t *__restrict__ gmin)
{
    const int64_t g = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t e = g * kEuBlock + lane;
    bool valid;
    int32_t mi;
    if (rows) {
        const int32_t j = idx[e]; // the rows are padded to whole blocks
        valid = j >= 0;
        mi = valid ? j : kEuNone;
    } else {
        valid = e < n_in;
        mi = valid ? in_min[e] : kEuNone;
    }
    for (int32_t c = 0; c < d; c++) {
        float a = __builtin_huge_valf(), b = -__builtin_huge_valf();
        if (valid) {
            if (rows) {
                a = b = rows[e * d + c];
            } else {
                a = in_lo[(int64_t)c * n_in + e];
                b = in_hi[(int64_t)c * n_in + e];
            }
        }
        a = wave_min(a);
        b = wave_max(b);
        if (lane == 0) {
            lo[(int64_t)c * n_out + g] = a;
            hi[(int64_t)c * n_out + g] = b;
        }
    }
    mi = wave_min_i(mi);
    if (lane == 0)
        gmin[g] = mi;
}

// Hilbert code of each query (same bounds and bits as the points), the block
// holding its position in the sorted codes, and the identity order.
__global__ void __launch_bounds__(256)
    eu_seed_kernel(const float *__restrict__ queries, int64_t nq, int32_t d, int32_t bits,
                   const double *__restrict__ bounds, const uint64_t *__restrict__ codes,
                   int64_t n, uint64_t *__restrict__ qcodes, int64_t *__restrict__ qorder,
                   int32_t *__restrict__ seed)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq)
        return;
    uint64_t v[kEuMaxD];
    for (int32_t c = 0; c < d; c++)
        v[c] = hilbert_quantize((double)queries[i * d + c], bounds[c], bounds[d + c], bits);
    const uint64_t code = hilbert_from_quantized(v, d, bits);
    int64_t a = 0, b = n; // lower bound of code in codes[0, n)
    while (a < b) {
        const int64_t mid = (a + b) >> 1;
        if (codes[mid] < code)
            a = mid + 1;
        else
            b = mid;
    }
    const int64_t pos = a < n ? a : n - 1;
    seed[i] = (int32_t)(pos / kEuBlock);
    qcodes[i] = code;
    qorder[i] = i;
}

struct EuArgs {
    const float *pts;    // npad x d, Hilbert order
    const int32_t *idx;  // npad, -1 for padding
    const float *lo, *hi;     // block boxes, axis-major over nb
    const int32_t *bmin;      // smallest index per block
    const float *slo, *shi;   // superblock boxes, axis-major over nsb
    const int32_t *sbmin;
    const int32_t *seed;      // per query: block of its Hilbert position
    const int64_t *order;     // processing order of the queries (Hilbert)
    const float *queries;
    int64_t nq, nb, nsb;
    int32_t d, k, pred;
    int64_t *out_idx;
    float *out_dist;
};

__device__ __forceinline__ bool eu_less(float da, int32_t ia, float db, int32_t ib)
{
    return da < db || (da == db && ia < ib);
}

__device__ __forceinline__ float rdl(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// box lower bound, folded in axis order like the distance
__device__ __forceinline__ float eu_box_lb(const float *__restrict__ lo,
                                           const float *__restrict__ hi, int64_t stride,
                                           int64_t g, const float *q, int32_t d)
{
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < kEuMaxD; c++) {
        if (c < d) {
            const float a = lo[(int64_t)c * stride + g] - q[c];
            const float b = q[c] - hi[(int64_t)c * stride + g];
            const float gap = fmaxf(fmaxf(a, b), 0.0f);
            acc = acc + gap * gap;
        }
    }
    return acc;
}

// The wave's sorted list: lane j holds the j-th smallest (d^2, index) pair;
// lanes past k hold what was shifted out and never sort before lane k-1.
struct EuList {
    float d;
    int32_t i;
    float kd;   // lane k-1's pair, the admission bound
    int32_t ki;
};

__device__ __forceinline__ void eu_insert(EuList &L, float cd, int32_t ci, int lane, int32_t k)
{
    const bool before = eu_less(L.d, L.i, cd, ci);
    const int pos = __builtin_popcountll(__ballot(before));
    const float up_d = __shfl_up(L.d, 1, 64);
    const int32_t up_i = __shfl_up(L.i, 1, 64);
    if (lane == pos) {
        L.d = cd;
        L.i = ci;
    } else if (lane > pos) {
        L.d = up_d;
        L.i = up_i;
    }
    L.kd = rdl(L.d, k - 1);
    L.ki = __builtin_amdgcn_readlane(L.i, k - 1);
}

__device__ __forceinline__ void eu_scan_block(const EuArgs &A, EuList &L, int64_t b,
                                              const float *q, int64_t qi, int lane)
{
    const int64_t row = b * kEuBlock + lane;
    const int32_t j = A.idx[row];
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < kEuMaxD; c++) {
        if (c < A.d) {
            const float df = q[c] - A.pts[row * A.d + c];
            acc = acc + df * df;
        }
    }
    const bool ok = j >= 0 && (!A.pred || (int64_t)j < qi);
    uint64_t mask = __ballot(ok && eu_less(acc, j, L.kd, L.ki));
    while (mask) {
        const int l = __builtin_ctzll(mask);
        mask &= mask - 1;
        const float cd = rdl(acc, l);
        const int32_t ci = __builtin_amdgcn_readlane(j, l);
        if (eu_less(cd, ci, L.kd, L.ki))
            eu_insert(L, cd, ci, lane, A.k);
    }
}

__global__ void __launch_bounds__(64 * kEuQueryWaves) knn_euclid_exact_kernel(EuArgs A)
{
    const int lane = threadIdx.x & 63;
    // consecutive Hilbert-ordered queries go to one XCD (workgroups are dealt
    // to the 8 XCDs round robin), so neighbouring queries share its L2
    // the grid is a whole number of rounds of kEuXcds workgroups
    const int64_t per_xcd = gridDim.x / kEuXcds;
    const int64_t gx = (int64_t)(blockIdx.x % kEuXcds) * per_xcd + blockIdx.x / kEuXcds;
    const int64_t w = gx * kEuQueryWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (w >= A.nq)
        return;
    const int64_t qi = A.order[w];
    float q[kEuMaxD];
#pragma unroll
    for (int c = 0; c < kEuMaxD; c++)
        q[c] = c < A.d ? A.queries[qi * A.d + c] : 0.0f;
    EuList L{__builtin_huge_valf(), kEuNone, __builtin_huge_valf(), kEuNone};
    if (!(A.pred && qi == 0)) {
        // seed: the blocks around the query's Hilbert position
        const int64_t b0 = A.seed[qi];
        const int64_t s0 = b0 > 0 ? b0 - 1 : 0;
        const int64_t s1 = b0 + 1 < A.nb ? b0 + 1 : A.nb - 1;
        for (int64_t b = s0; b <= s1; b++)
            if (!A.pred || (int64_t)A.bmin[b] < qi)
                eu_scan_block(A, L, b, q, qi, lane);
        // sweep: superblocks whose box can hold a candidate, then their blocks
        for (int64_t sc = 0; sc < A.nsb; sc += 64) {
            const int64_t sb = sc + lane;
            bool live = sb < A.nsb && (!A.pred || (int64_t)A.sbmin[sb] < qi);
            const float slb = live ? eu_box_lb(A.slo, A.shi, A.nsb, sb, q, A.d) : 0.0f;
            uint64_t smask = __ballot(live && slb <= L.kd);
            while (smask) {
                const int l = __builtin_ctzll(smask);
                smask &= smask - 1;
                if (!(rdl(slb, l) <= L.kd))
                    continue;
                const int64_t b = (sc + l) * kEuBlock + lane;
                bool bl = b < A.nb && (b < s0 || b > s1) &&
                          (!A.pred || (int64_t)A.bmin[b] < qi);
                const float blb = bl ? eu_box_lb(A.lo, A.hi, A.nb, b, q, A.d) : 0.0f;
                uint64_t bmask = __ballot(bl && blb <= L.kd);
                while (bmask) {
                    const int l2 = __builtin_ctzll(bmask);
                    bmask &= bmask - 1;
                    if (rdl(blb, l2) <= L.kd)
                        eu_scan_block(A, L, (sc + l) * kEuBlock + l2, q, qi, lane);
                }
            }
        }
    }
    if (lane < A.k) {
        const bool none = L.i == kEuNone;
        A.out_idx[qi * A.k + lane] = none ? -1 : L.i;
        A.out_dist[qi * A.k + lane] = none ? 3.40282347e38f : __builtin_sqrtf(L.d);
    }
}

__global__ void __launch_bounds__(256)
    fill_empty_kernel(int64_t count, int64_t *__restrict__ idx, float *__restrict__ dist)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        idx[i] = -1;
        dist[i] = 3.40282347e38f;
    }
}

// Hilbert-sorted copy of the points with block and superblock boxes (once per
// point set; finalize_tree clears eu_ready).
static void prepare_exact(TreeState &t, hipStream_t st)
{
    if (t.eu_ready)
        return;
    const int64_t n = t.n;
    const int32_t d = t.d;
    const int64_t nb = (n + kEuBlock - 1) / kEuBlock;
    const int64_t nsb = (nb + kEuBlock - 1) / kEuBlock;
    const int64_t npad = nb * kEuBlock;
    DevBuf<int64_t> perm(n);
    t.eu_codes.alloc(n);
    t.eu_bounds.alloc(2 * (size_t)d);
    hilbert_sort_points_f32(t.raw.ptr, n, d, t.eu_codes.ptr, perm.ptr, t.eu_bounds.ptr,
                            t.eu_bounds.ptr + d, st);
    t.eu_pts.alloc((size_t)npad * d);
    t.eu_idx.alloc(npad);
    eu_gather_kernel<<<(unsigned)((npad + 255) / 256), 256, 0, st>>>(t.raw.ptr, perm.ptr, n, d,
                                                                      npad, t.eu_pts.ptr,
                                                                      t.eu_idx.ptr);
    // one allocation for both levels: boxes axis-major, then the minima
    t.eu_lo.alloc((size_t)(nb + nsb) * d);
    t.eu_hi.alloc((size_t)(nb + nsb) * d);
    t.eu_min.alloc(nb + nsb);
    eu_box_kernel<<<(unsigned)nb, 64, 0, st>>>(t.eu_pts.ptr, t.eu_idx.ptr, nullptr, nullptr,
                                                nullptr, npad, nb, d, t.eu_lo.ptr, t.eu_hi.ptr,
                                                t.eu_min.ptr);
    eu_box_kernel<<<(unsigned)nsb, 64, 0, st>>>(
        nullptr, nullptr, t.eu_lo.ptr, t.eu_hi.ptr, t.eu_min.ptr, nb, nsb, d,
        t.eu_lo.ptr + nb * d, t.eu_hi.ptr + nb * d, t.eu_min.ptr + nb);
    CX_HIP(hipGetLastError());
    CX_HIP(hipStreamSynchronize(st)); // perm is freed on return
    t.eu_nb = nb;
    t.eu_ready = true;
}

void knn_euclidean(TreeState &t, const float *d_queries, int64_t nq, int32_t k, int32_t pred,
                   int32_t mode, int64_t *d_idx, float *d_dist)
{
    CX_REQUIRE(t.d <= kEuMaxD, "Euclidean kNN supports d <= 16");
    CX_REQUIRE(k >= 1 && k <= kEuMaxK, "Euclidean kNN supports 1 <= k <= 64");
    if (nq <= 0)
        return;
    hipStream_t st = t.stream;
    prof_mark(t, 0);
    if (mode == CX_KNN_REFERENCE_F32) {
        const int64_t waves = nq < 1024 ? nq : 1024;
        const int64_t cap = t.n + 64;
        uint8_t *ws = workspace(t, 256 + (size_t)waves * cap * sizeof(Cand));
        int32_t *counter = (int32_t *)ws;
        CX_HIP(hipMemsetAsync(counter, 0, sizeof(int32_t), st));
        knn_euclid_ref_kernel<<<(unsigned)waves, 64, 0, st>>>(
            t.raw.ptr, t.d, t.n, t.roots.ptr, t.n_roots, t.child_off.ptr, t.child_list.ptr,
            t.levels.ptr, d_queries, nq, k, pred, d_idx, d_dist, counter, (Cand *)(ws + 256), cap);
    } else if (t.n == 0) {
        fill_empty_kernel<<<(unsigned)((nq * k + 255) / 256), 256, 0, st>>>(nq * k, d_idx,
                                                                            d_dist);
    } else {
        prepare_exact(t, st);
        const int64_t nb = t.eu_nb, nsb = (nb + kEuBlock - 1) / kEuBlock;
        // query codes, seeds and the Hilbert processing order
        const size_t a_codes = (size_t)nq * 8, a_order = (size_t)nq * 8, a_seed = (size_t)nq * 4;
        size_t sort_bytes = 0;
        CX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint64_t *)nullptr,
                                                  (uint64_t *)nullptr, (int64_t *)nullptr,
                                                  (int64_t *)nullptr, (int)nq, 0, 64, st));
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        uint8_t *ws = workspace(t, 2 * al(a_codes) + 2 * al(a_order) + al(a_seed) + al(sort_bytes));
        uint64_t *qc = (uint64_t *)ws, *qc_sorted = (uint64_t *)(ws + al(a_codes));
        int64_t *qo = (int64_t *)(ws + 2 * al(a_codes));
        int64_t *qo_sorted = (int64_t *)(ws + 2 * al(a_codes) + al(a_order));
        int32_t *seed = (int32_t *)(ws + 2 * al(a_codes) + 2 * al(a_order));
        uint8_t *sort_tmp = ws + 2 * al(a_codes) + 2 * al(a_order) + al(a_seed);
        eu_seed_kernel<<<(unsigned)((nq + 255) / 256), 256, 0, st>>>(
            d_queries, nq, t.d, hilbert_bits(t.d), t.eu_bounds.ptr, t.eu_codes.ptr, t.n, qc, qo,
            seed);
        CX_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, qc, qc_sorted, qo,
                                                  qo_sorted, (int)nq, 0, 64, st));
        EuArgs A{t.eu_pts.ptr, t.eu_idx.ptr, t.eu_lo.ptr, t.eu_hi.ptr, t.eu_min.ptr,
                 t.eu_lo.ptr + nb * t.d, t.eu_hi.ptr + nb * t.d, t.eu_min.ptr + nb, seed,
                 qo_sorted, d_queries, nq, nb, nsb, t.d, k, pred, d_idx, d_dist};
        const int64_t groups = (nq + kEuQueryWaves - 1) / kEuQueryWaves;
        const int64_t grid = (groups + kEuXcds - 1) / kEuXcds * kEuXcds;
        knn_euclid_exact_kernel<<<(unsigned)grid, 64 * kEuQueryWaves, 0, st>>>(A);
    }
    CX_HIP(hipGetLastError());
    prof_mark(t, 1);
}

} // namespace cx
