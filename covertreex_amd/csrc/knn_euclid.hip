// Euclidean kNN over the device cover tree.
//
// CX_KNN_REFERENCE_F32: single_knn_query (src/algo.rs:208-323) restated, one
//   wavefront per query: every root seeded into a candidate min-heap, best-first
//   pops, prune when the result heap is full and d - 2^(level+1) > kth, all
//   children pushed.  Both heaps follow std::collections::BinaryHeap's sift
//   rules so ties resolve as in the reference; the candidate heap lives in a
//   per-wavefront HBM buffer.  Predecessor mode compares node index < query
//   position, as batch_knn_query does (src/algo.rs:194).
// CX_KNN_EXACT: exact k nearest by a tiled scan: each workgroup holds 128
//   queries (one per thread) and streams candidate points through LDS; every
//   thread keeps its own sorted top-k in LDS.  FP32 differences, ties broken by
//   the lower index.
#include "cx_internal.hpp"

namespace cx {

constexpr int kEuMaxD = 16;
constexpr int kEuMaxK = 64;
constexpr int kEuThreads = 128;
constexpr int kEuChunk = 256;

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

__global__ void __launch_bounds__(kEuThreads)
    knn_euclid_exact_kernel(const float *__restrict__ pts, int32_t d, int64_t n,
                            const float *__restrict__ queries, int64_t nq, int32_t k, int32_t pred,
                            int64_t *out_idx, float *out_dist)
{
    __shared__ float s_c[kEuChunk * kEuMaxD];
    __shared__ float s_td[kEuMaxK * kEuThreads];
    __shared__ int32_t s_ti[kEuMaxK * kEuThreads];
    const int t = threadIdx.x;
    const int64_t qi = (int64_t)blockIdx.x * kEuThreads + t;
    const bool live = qi < nq;
    float q[kEuMaxD];
#pragma unroll
    for (int c = 0; c < kEuMaxD; c++)
        q[c] = (live && c < d) ? queries[qi * d + c] : 0.0f;
    int count = 0;
    float kth2 = __builtin_huge_valf();
    // candidates below this block's largest query position suffice under the mask
    int64_t c_end = n;
    if (pred) {
        const int64_t last = ((int64_t)blockIdx.x + 1) * kEuThreads;
        c_end = last < n ? last : n;
    }
    for (int64_t c0 = 0; c0 < c_end; c0 += kEuChunk) {
        __syncthreads();
        const int cn = (int)((c_end - c0) < kEuChunk ? (c_end - c0) : kEuChunk);
        for (int i = t; i < cn * d; i += kEuThreads)
            s_c[i] = pts[c0 * d + i];
        __syncthreads();
        if (!live)
            continue;
        for (int u = 0; u < cn; u++) {
            const int64_t j = c0 + u;
            if (pred && j >= qi)
                break;
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < kEuMaxD; c++) {
                if (c < d) {
                    float df = q[c] - s_c[u * d + c];
                    acc = acc + df * df;
                }
            }
            if (count == k && !(acc < kth2))
                continue;
            // insert into the ascending list (ties: lower index first)
            int pos = count < k ? count : k - 1;
            while (pos > 0) {
                const float pv = s_td[(pos - 1) * kEuThreads + t];
                if (!(acc < pv))
                    break;
                s_td[pos * kEuThreads + t] = pv;
                s_ti[pos * kEuThreads + t] = s_ti[(pos - 1) * kEuThreads + t];
                pos--;
            }
            s_td[pos * kEuThreads + t] = acc;
            s_ti[pos * kEuThreads + t] = (int32_t)j;
            if (count < k)
                count++;
            if (count == k)
                kth2 = s_td[(k - 1) * kEuThreads + t];
        }
    }
    if (!live)
        return;
    for (int j = 0; j < k; j++) {
        out_idx[qi * k + j] = j < count ? s_ti[j * kEuThreads + t] : -1;
        out_dist[qi * k + j] =
            j < count ? __builtin_sqrtf(s_td[j * kEuThreads + t]) : 3.40282347e38f;
    }
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
    } else {
        const unsigned blocks = (unsigned)((nq + kEuThreads - 1) / kEuThreads);
        knn_euclid_exact_kernel<<<blocks, kEuThreads, 0, st>>>(t.raw.ptr, t.d, t.n, d_queries, nq,
                                                                k, pred, d_idx, d_dist);
    }
    CX_HIP(hipGetLastError());
    prof_mark(t, 1);
}

} // namespace cx
