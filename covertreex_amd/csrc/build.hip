// Device tree build for the residual-correlation (and Euclidean) cover tree.
//
//   morton order  -> hilbert_like_order, src/pcct.rs:27-67
//   gather        -> build_pcct2_residual_tree reorder, src/lib.rs:1404-1443
//                    ResidualMetric::new scaled coords / norms, src/metric.rs:84-177
//   batch insert  -> batch_insert_collect, src/algo/batch/mod.rs:447-800
//   si_cache      -> compute_si_cache_residual, src/algo.rs:52-106
//   subtree bounds-> compute_subtree_index_bounds, src/tree.rs:144-185
//
// Insertion is sequential over batches and levels, so one persistent
// workgroup (16 wavefronts) walks the whole build while the tree stays in HBM
// and each batch's candidate state stays in LDS.  Within a level the work is
// spread over the workgroup: one candidate per thread for the near/far scan,
// one far candidate per thread for each greedy-MIS step and for the
// non-leader assignment.  Distances use the reference f32 arithmetic
// (device_common.hpp) so every comparison matches the reference bit for bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "cx_internal.hpp"
#include "device_common.hpp"

namespace cx {

constexpr int kBiThreads = 1024;
constexpr int kBiMaxChunk = 1024;
constexpr int kBiWaves = kBiThreads / 64;

enum { kMetricResidual = 0, kMetricEuclidean = 1 };

struct TreeLinks {
    int64_t *parents;
    int64_t *children;
    int64_t *next;
    int32_t *levels;
};

// Euclidean::distance_sq, src/metric.rs:33-41.
__device__ __forceinline__ float euclid_sq(const float *__restrict__ pts, int32_t d, int64_t i,
                                           int64_t j)
{
    const float *x = pts + i * d, *y = pts + j * d;
    float acc = 0.0f;
    for (int32_t c = 0; c < d; c++) {
        float df = x[c] - y[c];
        acc = acc + df * df;
    }
    return acc;
}

template <int KIND>
__device__ __forceinline__ float build_dist_sq(const MetricView &mv, int64_t i, int64_t j)
{
    if constexpr (KIND == kMetricEuclidean)
        return euclid_sq(mv.scaled, mv.d, i, j);
    else
        return dist_sq_ref(mv, i, j);
}

__device__ __forceinline__ void link_child(const TreeLinks &t, int64_t parent, int64_t child)
{
    // CoverTreeData::link_child, src/tree.rs:98-103: prepend.
    int64_t old = ld_rx(&t.children[parent]);
    st_rx(&t.children[parent], child);
    st_rx(&t.next[child], old);
}

__device__ __forceinline__ int64_t grid_key(float v, float s)
{
    // (v * scale).floor().to_i64().unwrap_or(0), src/algo/batch/mod.rs:232-235
    float f = floorf(v * s);
    if (f != f || f >= 9.2233720368547758e18f || f < -9.2233720368547758e18f)
        return 0;
    return (int64_t)f;
}

// grid_conflicts_for_candidates (src/algo/batch/mod.rs:212-327) as a
// predicate: equal cell keys, or Chebyshev-1 neighbours when D <= 3 (the
// neighbour keys span min(D,3) axes and only match full keys when D <= 3).
__device__ __forceinline__ bool grid_adjacent(const float *__restrict__ raw, int32_t gd, float gs,
                                              int64_t a, int64_t b)
{
    if (gd <= 3) {
        for (int32_t c = 0; c < gd; c++) {
            int64_t dk = grid_key(raw[a * gd + c], gs) - grid_key(raw[b * gd + c], gs);
            if (dk > 1 || dk < -1)
                return false;
        }
        return true;
    }
    for (int32_t c = 0; c < gd; c++)
        if (grid_key(raw[a * gd + c], gs) != grid_key(raw[b * gd + c], gs))
            return false;
    return true;
}

// Ordered block compaction: returns the position of this thread's flagged
// item among flagged items of lower thread id; *total receives the count.
__device__ __forceinline__ int block_prefix(bool flag, int *s_cnt, int *total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t mask = __ballot(flag);
    int pre = __popcll(mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (lane == 0)
        s_cnt[wid] = __popcll(mask);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBiWaves; w++) {
        int c = s_cnt[w];
        off += (w < wid) ? c : 0;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + pre;
}

__device__ __forceinline__ int block_min_int(int v, int *s_red)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int u = __shfl_xor(v, o, 64);
        v = u < v ? u : v;
    }
    if (lane == 0)
        s_red[wid] = v;
    __syncthreads();
    int r = s_red[0];
#pragma unroll
    for (int w = 1; w < kBiWaves; w++)
        r = s_red[w] < r ? s_red[w] : r;
    __syncthreads();
    return r;
}

// Inserts rows [s_begin, n) in batches of `chunk`.  Per batch b (counted from
// s_begin) it records bstats[b * kBatchFields + f] (see cx_internal.hpp):
// start row, size, leaders, bottom-level survivors, dropped orphans, levels
// with a far set, distance evaluations of the near/far scan, the MIS and the
// non-leader step, conflict edges found through the grid and through
// distance, and wall-clock ticks of the three steps and of the whole batch.
template <int KIND>
__global__ void __launch_bounds__(kBiThreads)
    batch_insert_kernel(MetricView mv, TreeLinks t, const float *__restrict__ raw, int32_t gd,
                        float gs, int32_t use_grid, int64_t s_begin, int64_t n, int64_t chunk,
                        int64_t *bstats)
{
    __shared__ int32_t s_ca[kBiMaxChunk], s_cb[kBiMaxChunk];
    __shared__ int32_t s_va[kBiMaxChunk], s_vb[kBiMaxChunk];
    __shared__ int32_t s_best[kBiMaxChunk], s_far[kBiMaxChunk], s_farorig[kBiMaxChunk];
    __shared__ int32_t s_leader[kBiMaxChunk], s_nlbest[kBiMaxChunk];
    __shared__ int8_t s_state[kBiMaxChunk];
    __shared__ int s_cnt[kBiWaves], s_red[kBiWaves];
    __shared__ int s_nl;

    const int tid = threadIdx.x;

    for (int64_t s = s_begin; s < n; s += chunk) {
        const int64_t e = s + chunk < n ? s + chunk : n;
        const int nb = (int)(e - s);
        int64_t ev_scan = 0, ev_mis = 0, ev_nl = 0, n_orphans = 0, e_grid = 0, e_dist = 0;
        int64_t n_leaders = 0, n_far_levels = 0;
        uint64_t tk_scan = 0, tk_mis = 0, tk_nl = 0;
        const uint64_t tk_begin = wall_clock64();
        // add_point(row, i32::MIN, -1), src/algo/batch/mod.rs:511-513
        for (int i = tid; i < nb; i += kBiThreads) {
            st_rx(&t.parents[s + i], (int64_t)-1);
            st_rx(&t.children[s + i], (int64_t)-1);
            st_rx(&t.next[s + i], (int64_t)-1);
            st_rx(&t.levels[s + i], kLevelUnset);
        }
        int32_t *cur_c = s_ca, *cur_v = s_va, *nxt_c = s_cb, *nxt_v = s_vb;
        int nc;
        if (s == 0) { // root bootstrap at level 10, levels -10..10 (521-531)
            if (tid == 0)
                st_rx(&t.levels[0], (int32_t)10);
            nc = nb - 1;
            for (int i = tid; i < nc; i += kBiThreads) {
                cur_c[i] = (int32_t)(s + 1 + i);
                cur_v[i] = 0;
            }
        } else {
            nc = nb;
            for (int i = tid; i < nc; i += kBiThreads) {
                cur_c[i] = (int32_t)(s + i);
                cur_v[i] = 0; // active sets start as [root_idx = 0] (539-540)
            }
        }
        __syncthreads();
        for (int32_t level = 9; level >= -10 && nc > 0; --level) {
            const float r = ldexpf(1.0f, level);
            const float r2 = r * r;
            const uint64_t tk0 = wall_clock64();
            // near/far over each candidate's cover and the cover's children (580-627)
            if (tid < nc) {
                const int64_t q = cur_c[tid], p = cur_v[tid];
                float bd = 3.40282347e38f;
                int64_t bn = -1;
                float dp = build_dist_sq<KIND>(mv, q, p);
                ev_scan++;
                if (dp <= r2 && dp < bd) {
                    bd = dp;
                    bn = p;
                }
                const int64_t first = ld_rx(&t.children[p]);
                int64_t c = first;
                while (c != -1) {
                    float dc = build_dist_sq<KIND>(mv, q, c);
                    ev_scan++;
                    if (dc <= r2 && dc < bd) {
                        bd = dc;
                        bn = c;
                    }
                    int64_t nx = ld_rx(&t.next[c]);
                    if (nx == c)
                        break;
                    c = nx;
                    if (c == first)
                        break;
                }
                s_best[tid] = (int32_t)bn;
            }
            __syncthreads();
            const uint64_t tk1 = wall_clock64();
            tk_scan += tk1 - tk0;
            bool active = tid < nc;
            bool near = active && s_best[tid] >= 0;
            int nn = 0;
            int pos = block_prefix(near, s_cnt, &nn);
            if (near) {
                nxt_c[pos] = cur_c[tid];
                nxt_v[pos] = s_best[tid];
            } else if (active) {
                int fpos = tid - pos;
                s_far[fpos] = cur_c[tid];
                s_farorig[fpos] = tid;
            }
            const int nf = nc - nn;
            __syncthreads();
            if (nf > 0) {
                n_far_levels++;
                // greedy MIS in far-list order (compute_mis_greedy, mis.rs:9-59)
                // over pairwise d_sq <= r^2 conflicts (370-400) and grid edges
                // (662-674): only rows of selected candidates are ever needed.
                if (tid < nf)
                    s_state[tid] = 0;
                __syncthreads();
                int start = 0;
                for (;;) {
                    int cand = (tid < nf && tid >= start && s_state[tid] == 0) ? tid : INT32_MAX;
                    int istar = block_min_int(cand, s_red);
                    if (istar == INT32_MAX)
                        break;
                    if (tid == istar)
                        s_state[tid] = 1;
                    if (tid > istar && tid < nf && s_state[tid] == 0) {
                        const int64_t a = s_far[istar], b = s_far[tid];
                        bool adj = false;
                        if (use_grid)
                            adj = grid_adjacent(raw, gd, gs, a, b);
                        if (adj) {
                            e_grid++;
                        } else {
                            adj = build_dist_sq<KIND>(mv, a, b) <= r2;
                            ev_mis++;
                            e_dist += adj ? 1 : 0;
                        }
                        if (adj)
                            s_state[tid] = -1;
                    }
                    __syncthreads();
                    start = istar + 1;
                }
                // leaders: level and parent = previous cover, in far order (694-714)
                if (tid == 0) {
                    int nl = 0;
                    for (int i = 0; i < nf; i++) {
                        if (s_state[i] != 1)
                            continue;
                        const int64_t qn = s_far[i];
                        s_leader[nl++] = (int32_t)qn;
                        st_rx(&t.levels[qn], level);
                        const int64_t par = cur_v[s_farorig[i]];
                        st_rx(&t.parents[qn], par);
                        link_child(t, par, qn);
                    }
                    s_nl = nl;
                    n_leaders += nl;
                }
                __syncthreads();
                const uint64_t tk2 = wall_clock64();
                tk_mis += tk2 - tk1;
                const int nl = s_nl;
                // non-leaders: nearest leader within r, strict <, first wins (716-769)
                if (tid < nf) {
                    int32_t bl = -1;
                    if (s_state[tid] != 1) {
                        float bd = 3.40282347e38f;
                        const int64_t qn = s_far[tid];
                        for (int l = 0; l < nl; l++) {
                            float dl = build_dist_sq<KIND>(mv, qn, s_leader[l]);
                            if (dl <= r2 && dl < bd) {
                                bd = dl;
                                bl = s_leader[l];
                            }
                        }
                        ev_nl += nl;
                        if (bl < 0)
                            n_orphans++;
                    }
                    s_nlbest[tid] = bl;
                }
                __syncthreads();
                bool keep = tid < nf && s_state[tid] != 1 && s_nlbest[tid] >= 0;
                int nk = 0;
                int kp = block_prefix(keep, s_cnt, &nk);
                if (keep) {
                    nxt_c[nn + kp] = s_far[tid];
                    nxt_v[nn + kp] = s_nlbest[tid];
                }
                nn += nk;
                __syncthreads();
                tk_nl += wall_clock64() - tk2;
            }
            int32_t *tc = cur_c, *tv = cur_v;
            cur_c = nxt_c;
            cur_v = nxt_v;
            nxt_c = tc;
            nxt_v = tv;
            nc = nn;
            __syncthreads();
        }
        // survivors at the bottom level (791-800)
        if (tid == 0) {
            for (int i = 0; i < nc; i++) {
                const int64_t qn = cur_c[i];
                st_rx(&t.levels[qn], (int32_t)-10);
                st_rx(&t.parents[qn], (int64_t)cur_v[i]);
                link_child(t, cur_v[i], qn);
            }
        }
        __syncthreads();
        if (bstats) {
            int64_t *bs = bstats + ((s - s_begin) / chunk) * kBatchFields;
            atomicAdd((unsigned long long *)&bs[6], (unsigned long long)ev_scan);
            atomicAdd((unsigned long long *)&bs[7], (unsigned long long)ev_mis);
            atomicAdd((unsigned long long *)&bs[8], (unsigned long long)ev_nl);
            atomicAdd((unsigned long long *)&bs[4], (unsigned long long)n_orphans);
            atomicAdd((unsigned long long *)&bs[9], (unsigned long long)e_grid);
            atomicAdd((unsigned long long *)&bs[10], (unsigned long long)e_dist);
            if (tid == 0) {
                bs[0] = s;
                bs[1] = nb;
                bs[2] = n_leaders;
                bs[3] = nc;
                bs[5] = n_far_levels;
                bs[11] = (int64_t)tk_scan;
                bs[12] = (int64_t)tk_mis;
                bs[13] = (int64_t)tk_nl;
                bs[14] = (int64_t)(wall_clock64() - tk_begin);
            }
        }
    }
}

// ---- Morton order ---------------------------------------------------------

__global__ void minmax3_kernel(const float *__restrict__ x, int64_t n, float *out)
{
    // per-axis min/max; NaN never updates (matches `if v < mins[i]`)
    __shared__ float smin[3][256], smax[3][256];
    float mn[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
    float mx[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        for (int c = 0; c < 3; c++) {
            float v = x[i * 3 + c];
            if (v < mn[c])
                mn[c] = v;
            if (v > mx[c])
                mx[c] = v;
        }
    for (int c = 0; c < 3; c++) {
        smin[c][threadIdx.x] = mn[c];
        smax[c][threadIdx.x] = mx[c];
    }
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int c = 0; c < 3; c++) {
                float a = smin[c][threadIdx.x + o], b = smax[c][threadIdx.x + o];
                if (a < smin[c][threadIdx.x])
                    smin[c][threadIdx.x] = a;
                if (b > smax[c][threadIdx.x])
                    smax[c][threadIdx.x] = b;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++) {
            out[c] = smin[c][0];
            out[3 + c] = smax[c][0];
        }
}

__device__ __forceinline__ uint64_t part1by2(uint64_t v)
{
    // part1by2, src/pcct.rs:10-18
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void morton_codes_kernel(const float *__restrict__ x, int64_t n,
                                    const float *__restrict__ mm, uint64_t *codes, int64_t *idx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint64_t u[3];
    for (int c = 0; c < 3; c++) {
        float denom = __builtin_fabsf(mm[3 + c] - mm[c]);
        if (!(denom >= 1e-9f))
            denom = 1e-9f;
        float nv = (x[i * 3 + c] - mm[c]) / denom;
        nv = nv < 0.0f ? 0.0f : nv;
        nv = nv > 1.0f ? 1.0f : nv;
        float sv = nv * 2097151.0f;
        u[c] = (sv != sv) ? 0ull : (uint64_t)(uint32_t)sv;
    }
    codes[i] = part1by2(u[0]) | (part1by2(u[1]) << 1) | (part1by2(u[2]) << 2);
    idx[i] = i;
}

__global__ void iota_kernel(int64_t *out, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = i;
}

// Morton permutation on the device; order_dev receives n entries.
static void morton_order_on_stream(const float *coords_dev, int64_t n, int32_t d,
                                   int64_t *order_dev, hipStream_t st)
{
    const int threads = 256;
    const unsigned blocks = (unsigned)((n + threads - 1) / threads);
    if (n <= 0)
        return;
    if (d != 3) {
        iota_kernel<<<blocks, threads, 0, st>>>(order_dev, n);
        CX_HIP(hipGetLastError());
        return;
    }
    DevBuf<float> mm(6);
    DevBuf<uint64_t> codes(n), codes_sorted(n);
    DevBuf<int64_t> idx(n);
    minmax3_kernel<<<1, 256, 0, st>>>(coords_dev, n, mm.ptr);
    morton_codes_kernel<<<blocks, threads, 0, st>>>(coords_dev, n, mm.ptr, codes.ptr, idx.ptr);
    CX_HIP(hipGetLastError());
    // radix sort is stable: ties keep index order, as slice::sort_by_key does
    size_t tmp_bytes = 0;
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, codes.ptr, codes_sorted.ptr,
                                              idx.ptr, order_dev, (int)n, 0, 64, st));
    DevBuf<uint8_t> tmp(tmp_bytes);
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, codes.ptr, codes_sorted.ptr,
                                              idx.ptr, order_dev, (int)n, 0, 64, st));
    CX_HIP(hipStreamSynchronize(st));
}

// ---- reorder + metric precompute -------------------------------------------

__global__ void gather_rows_kernel(const float *__restrict__ src, const int64_t *__restrict__ order,
                                   int64_t n, int32_t w, float *__restrict__ dst)
{
    int64_t i = blockIdx.x;
    if (i >= n)
        return;
    const float *s = src + order[i] * (int64_t)w;
    float *t = dst + i * (int64_t)w;
    for (int32_t c = threadIdx.x; c < w; c += blockDim.x)
        t[c] = s[c];
}

__global__ void scale_rows_kernel(const float *__restrict__ raw, int64_t n, int32_t d,
                                  const float *__restrict__ inv_ls, float *__restrict__ scaled,
                                  float *__restrict__ norms)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float acc = 0.0f; // iterator sum, src/metric.rs:110-114
    for (int32_t c = 0; c < d; c++) {
        float sv = raw[i * d + c] * inv_ls[c];
        scaled[i * d + c] = sv;
        acc = acc + sv * sv;
    }
    norms[i] = acc;
}

__global__ void inverse_perm_kernel(const int64_t *__restrict__ order, int64_t n, int64_t *inv)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        inv[order[i]] = i;
}

// ---- depth, si_cache, subtree bounds -----------------------------------------

__global__ void depth_kernel(const int64_t *__restrict__ parents, int64_t n, int32_t *depth,
                             int32_t *top, int32_t *max_depth)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int32_t dep = 0;
    int64_t cur = i;
    while (parents[cur] >= 0 && dep <= (int32_t)n) {
        cur = parents[cur];
        dep++;
    }
    depth[i] = dep;
    top[i] = (int32_t)cur;
    atomicMax(max_depth, dep);
}

template <int KIND>
__global__ void si_level_kernel(MetricView mv, const int64_t *__restrict__ parents,
                                const int32_t *__restrict__ depth, const int32_t *__restrict__ top,
                                int64_t n, int32_t level, float *si)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || depth[i] != level || top[i] != 0)
        return;
    const int64_t par = parents[i];
    float dist = __builtin_sqrtf(build_dist_sq<KIND>(mv, par, i));
    float bound = dist + si[i];
    // r(node) = max(0, max bound) with `bound > max_radius` (src/algo.rs:81-84):
    // non-negative floats order like their bit patterns; NaN never wins.
    if (bound > 0.0f)
        atomicMax((int32_t *)&si[par], __float_as_int(bound));
}

__global__ void bounds_init_kernel(const int64_t *__restrict__ n2d, int64_t n, int64_t *mn,
                                   int64_t *mx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        mn[i] = n2d[i];
        mx[i] = n2d[i];
    }
}

__global__ void bounds_level_kernel(const int64_t *__restrict__ parents,
                                    const int32_t *__restrict__ depth, int64_t n, int32_t level,
                                    int64_t *mn, int64_t *mx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || depth[i] != level)
        return;
    const int64_t par = parents[i];
    atomicMin((long long *)&mn[par], (long long)mn[i]);
    atomicMax((long long *)&mx[par], (long long)mx[i]);
}

__global__ void root_flags_kernel(const int64_t *__restrict__ parents, int64_t n, int32_t *flags)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        flags[i] = parents[i] == -1 ? 1 : 0;
}

static unsigned nblocks(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

__global__ void child_count_kernel(const int64_t *__restrict__ parents, int64_t n, int32_t *cnt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && parents[i] >= 0)
        atomicAdd(&cnt[parents[i]], 1);
}

// Copy each node's linked child list (prepend order, src/tree.rs:98-103) into
// CSR, walking with the reference's loop guards.
__global__ void child_fill_kernel(const int64_t *__restrict__ children,
                                  const int64_t *__restrict__ next, const int32_t *__restrict__ off,
                                  int64_t n, int32_t *list)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t first = children[i];
    int64_t c = first;
    int32_t w = off[i];
    const int32_t end = off[i + 1];
    while (c != -1 && w < end) {
        list[w++] = (int32_t)c;
        int64_t nx = next[c];
        if (nx == c)
            break;
        c = nx;
        if (c == first)
            break;
    }
}

static void build_child_csr(TreeState &t, hipStream_t st)
{
    const int64_t n = t.n;
    DevBuf<int32_t> cnt(n + 1);
    CX_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(int32_t) * (n + 1), st));
    child_count_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.parents.ptr, n, cnt.ptr);
    t.child_off.alloc(n + 1);
    size_t tmp_bytes = 0;
    CX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt.ptr, t.child_off.ptr,
                                            (int)(n + 1), st));
    DevBuf<uint8_t> tmp(tmp_bytes);
    CX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, tmp_bytes, cnt.ptr, t.child_off.ptr,
                                            (int)(n + 1), st));
    t.child_list.alloc(n > 0 ? n : 1);
    child_fill_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.children.ptr, t.next.ptr,
                                                       t.child_off.ptr, n, t.child_list.ptr);
    CX_HIP(hipGetLastError());
    std::vector<int32_t> h(n);
    CX_HIP(hipMemcpyAsync(h.data(), cnt.ptr, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    CX_HIP(hipStreamSynchronize(st));
    int64_t internal = 0;
    for (int64_t i = 0; i < n; i++)
        internal += h[i] > 0;
    t.n_internal = internal;
}

static void subtree_bounds_on_stream(const int64_t *parents, const int64_t *n2d,
                                     const int32_t *depth, int32_t max_depth, int64_t n,
                                     int64_t *mn, int64_t *mx, hipStream_t st)
{
    bounds_init_kernel<<<nblocks(n, 256), 256, 0, st>>>(n2d, n, mn, mx);
    for (int32_t lv = max_depth; lv >= 1; lv--)
        bounds_level_kernel<<<nblocks(n, 256), 256, 0, st>>>(parents, depth, n, lv, mn, mx);
    CX_HIP(hipGetLastError());
}

static int32_t compute_depths(const int64_t *parents, int64_t n, int32_t *depth, int32_t *top,
                              hipStream_t st)
{
    DevBuf<int32_t> md(1);
    CX_HIP(hipMemsetAsync(md.ptr, 0, sizeof(int32_t), st));
    depth_kernel<<<nblocks(n, 256), 256, 0, st>>>(parents, n, depth, top, md.ptr);
    CX_HIP(hipGetLastError());
    int32_t h = 0;
    CX_HIP(hipMemcpyAsync(&h, md.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CX_HIP(hipStreamSynchronize(st));
    return h;
}

// Runs the batch insert of rows [s_begin, s_end) (default: to t.n) and
// appends the per-batch telemetry to t.batch_stats.
template <int KIND>
static void run_insert(TreeState &t, const MetricView &mv, int32_t use_grid, float grid_scale,
                       int64_t s_begin, int32_t chunk, hipStream_t st, int64_t s_end = -1)
{
    const int64_t n = s_end < 0 ? t.n : s_end;
    if (s_begin >= n)
        return;
    const int64_t nb = (n - s_begin + chunk - 1) / chunk;
    DevBuf<int64_t> bst((size_t)nb * kBatchFields);
    CX_HIP(hipMemsetAsync(bst.ptr, 0, bst.bytes(), st));
    TreeLinks links{t.parents.ptr, t.children.ptr, t.next.ptr, t.levels.ptr};
    batch_insert_kernel<KIND><<<1, kBiThreads, 0, st>>>(mv, links, t.raw.ptr, t.d, grid_scale,
                                                        use_grid, s_begin, n, chunk, bst.ptr);
    CX_HIP(hipGetLastError());
    for (int64_t b = 0; b < nb; b++) {
        const int64_t b0 = s_begin + b * chunk;
        t.live_batches.push_back(b0);
        t.live_batches.push_back(n - b0 < chunk ? n - b0 : chunk);
    }
    const size_t old = t.batch_stats.size();
    t.batch_stats.resize(old + (size_t)nb * kBatchFields);
    CX_HIP(hipMemcpyAsync(t.batch_stats.data() + old, bst.ptr, bst.bytes(), hipMemcpyDeviceToHost,
                          st));
    CX_HIP(hipStreamSynchronize(st));
}

// Arrays derived from the links after inserts: child CSR, depths, si_cache
// (bottom-up one depth at a time), subtree index bounds and the roots in
// index order.
template <int KIND> static void finalize_tree(TreeState &t, const MetricView &mv, hipStream_t st)
{
    const int64_t n = t.n;
    t.min_level = n > 0 ? -10 : -20;
    t.max_level = n > 0 ? 10 : 20;
    build_child_csr(t, st);
    t.depth.alloc(n);
    DevBuf<int32_t> top(n);
    t.max_depth = compute_depths(t.parents.ptr, n, t.depth.ptr, top.ptr, st);
    t.si.alloc(n);
    CX_HIP(hipMemsetAsync(t.si.ptr, 0, sizeof(float) * n, st));
    for (int32_t lv = t.max_depth; lv >= 1; lv--)
        si_level_kernel<KIND><<<nblocks(n, 256), 256, 0, st>>>(mv, t.parents.ptr, t.depth.ptr,
                                                               top.ptr, n, lv, t.si.ptr);
    CX_HIP(hipGetLastError());
    t.smin.alloc(n);
    t.smax.alloc(n);
    subtree_bounds_on_stream(t.parents.ptr, t.order.ptr, t.depth.ptr, t.max_depth, n, t.smin.ptr,
                             t.smax.ptr, st);
    DevBuf<int32_t> flags(n), nsel(1);
    t.roots.alloc(n);
    root_flags_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.parents.ptr, n, flags.ptr);
    hipcub::CountingInputIterator<int32_t> it(0);
    size_t tmp_bytes = 0;
    CX_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, it, flags.ptr, t.roots.ptr, nsel.ptr,
                                         (int)n, st));
    DevBuf<uint8_t> tmp(tmp_bytes);
    CX_HIP(hipcub::DeviceSelect::Flagged(tmp.ptr, tmp_bytes, it, flags.ptr, t.roots.ptr, nsel.ptr,
                                         (int)n, st));
    int32_t nr = 0;
    CX_HIP(hipMemcpyAsync(&nr, nsel.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CX_HIP(hipStreamSynchronize(st));
    t.n_roots = nr;
}

static MetricView residual_view(const TreeState &t)
{
    return MetricView{t.scaled.ptr, t.norms.ptr, t.v.ptr, t.p.ptr, t.d, t.m, t.var, t.kernel};
}

// Rows [n0, n0 + nn) of the tree arrays from host V, p, coords (nn rows):
// tree row n0 + i takes staged row sel[i]; coordinates are also scaled by
// 1/lengthscale.
static void load_residual_rows(TreeState &t, const float *V, const float *p, const float *coords,
                               int64_t n0, int64_t nn, const int64_t *sel, hipStream_t st)
{
    const int32_t m = t.m, d = t.d;
    DevBuf<float> v_in((size_t)nn * m), c_in((size_t)nn * d), p_in(nn);
    CX_HIP(hipMemcpyAsync(v_in.ptr, V, sizeof(float) * nn * m, hipMemcpyHostToDevice, st));
    CX_HIP(hipMemcpyAsync(c_in.ptr, coords, sizeof(float) * nn * d, hipMemcpyHostToDevice, st));
    CX_HIP(hipMemcpyAsync(p_in.ptr, p, sizeof(float) * nn, hipMemcpyHostToDevice, st));
    gather_rows_kernel<<<(unsigned)nn, 64, 0, st>>>(v_in.ptr, sel, nn, m, t.v.ptr + n0 * m);
    gather_rows_kernel<<<(unsigned)nn, 64, 0, st>>>(c_in.ptr, sel, nn, d, t.raw.ptr + n0 * d);
    gather_rows_kernel<<<(unsigned)nn, 64, 0, st>>>(p_in.ptr, sel, nn, 1, t.p.ptr + n0);
    DevBuf<float> inv_ls(d);
    CX_HIP(hipMemcpyAsync(inv_ls.ptr, t.inv_ls.data(), sizeof(float) * d, hipMemcpyHostToDevice,
                          st));
    scale_rows_kernel<<<nblocks(nn, 256), 256, 0, st>>>(t.raw.ptr + n0 * d, nn, d, inv_ls.ptr,
                                                        t.scaled.ptr + n0 * d, t.norms.ptr + n0);
    CX_HIP(hipGetLastError());
    CX_HIP(hipStreamSynchronize(st)); // the staging buffers go out of scope
}

void build_residual_tree(TreeState &t, const float *V, const float *p, const float *coords,
                         int32_t chunk, int32_t order_mode, float grid_scale)
{
    const int64_t n = t.n;
    const int32_t m = t.m, d = t.d;
    hipStream_t st = t.stream;
    CX_REQUIRE(chunk > 0 && chunk <= kBiMaxChunk, "chunk must be in [1, 1024]");
    t.chunk = chunk;
    t.grid_scale = grid_scale;

    t.order.alloc(n);
    t.inv_order.alloc(n);
    if (order_mode == CX_ORDER_NATURAL || d != 3) {
        iota_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.order.ptr, n);
        CX_HIP(hipGetLastError());
    } else {
        DevBuf<float> c_in((size_t)n * d);
        CX_HIP(hipMemcpyAsync(c_in.ptr, coords, sizeof(float) * n * d, hipMemcpyHostToDevice,
                              st));
        morton_order_on_stream(c_in.ptr, n, d, t.order.ptr, st);
        CX_HIP(hipStreamSynchronize(st));
    }
    inverse_perm_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.order.ptr, n, t.inv_order.ptr);

    t.v.alloc((size_t)n * m);
    t.raw.alloc((size_t)n * d);
    t.scaled.alloc((size_t)n * d);
    t.norms.alloc(n);
    t.p.alloc(n);
    load_residual_rows(t, V, p, coords, 0, n, t.order.ptr, st);

    t.parents.alloc(n);
    t.children.alloc(n);
    t.next.alloc(n);
    t.levels.alloc(n);
    const MetricView mv = residual_view(t);
    run_insert<kMetricResidual>(t, mv, 1, grid_scale, 0, chunk, st);
    finalize_tree<kMetricResidual>(t, mv, st);
}

__global__ void iota_from_kernel(int64_t *out, int64_t n0, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[n0 + i] = n0 + i;
}

// Appends rows to the tree's arrays: the new rows keep their given order
// (node n0 + i is dataset row n0 + i) and are batch-inserted with the grid
// scale of the build, as CoverTreeWrapper.insert_residual does with its
// batch (src/lib.rs:254-300).
void insert_residual_rows(TreeState &t, const float *V, const float *p, const float *coords,
                          int64_t n_new, int32_t chunk)
{
    hipStream_t st = t.stream;
    CX_REQUIRE(chunk > 0 && chunk <= kBiMaxChunk, "chunk must be in [1, 1024]");
    const int64_t n0 = t.n, n1 = t.n + n_new;
    const int32_t m = t.m, d = t.d;
    CX_REQUIRE(n1 < (int64_t)1 << 31, "tree size must stay below 2^31");
    t.v.grow_keep((size_t)n1 * m, st);
    t.raw.grow_keep((size_t)n1 * d, st);
    t.scaled.grow_keep((size_t)n1 * d, st);
    t.norms.grow_keep(n1, st);
    t.p.grow_keep(n1, st);
    t.order.grow_keep(n1, st);
    t.inv_order.grow_keep(n1, st);
    t.parents.grow_keep(n1, st);
    t.children.grow_keep(n1, st);
    t.next.grow_keep(n1, st);
    t.levels.grow_keep(n1, st);
    iota_from_kernel<<<nblocks(n_new, 256), 256, 0, st>>>(t.order.ptr, n0, n_new);
    iota_from_kernel<<<nblocks(n_new, 256), 256, 0, st>>>(t.inv_order.ptr, n0, n_new);
    // the staged rows are taken in their given order
    DevBuf<int64_t> ident(n_new);
    iota_kernel<<<nblocks(n_new, 256), 256, 0, st>>>(ident.ptr, n_new);
    CX_HIP(hipGetLastError());
    load_residual_rows(t, V, p, coords, n0, n_new, ident.ptr, st);
    t.n = n1;
    const MetricView mv = residual_view(t);
    run_insert<kMetricResidual>(t, mv, 1, t.grid_scale, n0, chunk, st);
    finalize_tree<kMetricResidual>(t, mv, st);
    t.mixed_ready = false;
}

// Euclidean tree: chunked batch_insert_collect with the Euclidean metric in
// natural order and no grid edges (src/lib.rs:207-223, src/algo/batch/mod.rs:
// 402-409), plus the same cover-radius recurrence with Euclidean distances.
void build_euclidean_tree(TreeState &t, const float *points, int32_t chunk)
{
    const int64_t n = t.n;
    const int32_t d = t.d;
    hipStream_t st = t.stream;
    CX_REQUIRE(chunk > 0 && chunk <= kBiMaxChunk, "chunk must be in [1, 1024]");
    t.chunk = chunk;
    t.raw.alloc((size_t)n * d);
    CX_HIP(hipMemcpyAsync(t.raw.ptr, points, sizeof(float) * n * d, hipMemcpyHostToDevice, st));
    t.order.alloc(n);
    t.inv_order.alloc(n);
    iota_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.order.ptr, n);
    iota_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.inv_order.ptr, n);
    t.parents.alloc(n);
    t.children.alloc(n);
    t.next.alloc(n);
    t.levels.alloc(n);
    const MetricView mv{t.raw.ptr, nullptr, nullptr, nullptr, d, 0, 1.0f, 0};
    run_insert<kMetricEuclidean>(t, mv, 0, 1.0f, 0, chunk, st);
    finalize_tree<kMetricEuclidean>(t, mv, st);
}

// CoverTreeWrapper.insert (src/lib.rs:207-223) on a device-resident
// Euclidean tree: the new points become nodes n0.. in their given order.
void insert_euclidean_rows(TreeState &t, const float *points, int64_t n_new, int32_t chunk)
{
    hipStream_t st = t.stream;
    CX_REQUIRE(chunk > 0 && chunk <= kBiMaxChunk, "chunk must be in [1, 1024]");
    const int64_t n0 = t.n, n1 = t.n + n_new;
    const int32_t d = t.d;
    CX_REQUIRE(n1 < (int64_t)1 << 31, "tree size must stay below 2^31");
    t.raw.grow_keep((size_t)n1 * d, st);
    t.order.grow_keep(n1, st);
    t.inv_order.grow_keep(n1, st);
    t.parents.grow_keep(n1, st);
    t.children.grow_keep(n1, st);
    t.next.grow_keep(n1, st);
    t.levels.grow_keep(n1, st);
    CX_HIP(hipMemcpyAsync(t.raw.ptr + n0 * d, points, sizeof(float) * n_new * d,
                          hipMemcpyHostToDevice, st));
    iota_from_kernel<<<nblocks(n_new, 256), 256, 0, st>>>(t.order.ptr, n0, n_new);
    iota_from_kernel<<<nblocks(n_new, 256), 256, 0, st>>>(t.inv_order.ptr, n0, n_new);
    CX_HIP(hipGetLastError());
    t.n = n1;
    const MetricView mv{t.raw.ptr, nullptr, nullptr, nullptr, d, 0, 1.0f, 0};
    run_insert<kMetricEuclidean>(t, mv, 0, 1.0f, n0, chunk, st);
    finalize_tree<kMetricEuclidean>(t, mv, st);
}

// ---- delete -------------------------------------------------------------------

// Removes dataset rows from a device-resident tree, following the reference's
// batch_delete (covertreex/algo/batch_delete.py:105-525): every descendant of
// a removed node is detached, the remaining tree keeps its links (each kept
// node's ancestors are kept), and the detached nodes are inserted again in
// groups of equal max(level, 0), lowest level first, each group in
// batches of t.chunk.  Here the re-insertion is this tree's batch insert.
// Remaining dataset rows are renumbered in their order (removed rows close
// up); tree order becomes [kept nodes in tree order, detached nodes by group].
// `remove` holds dataset rows, or tree nodes when by_node.
void delete_rows(TreeState &t, const int64_t *remove, int64_t n_remove, bool by_node)
{
    hipStream_t st = t.stream;
    const int64_t n = t.n;
    std::vector<int64_t> par(n), chi(n), nxt(n), ord(n), inv(n);
    std::vector<int32_t> lev(n);
    CX_HIP(hipMemcpyAsync(par.data(), t.parents.ptr, 8 * n, hipMemcpyDeviceToHost, st));
    CX_HIP(hipMemcpyAsync(chi.data(), t.children.ptr, 8 * n, hipMemcpyDeviceToHost, st));
    CX_HIP(hipMemcpyAsync(nxt.data(), t.next.ptr, 8 * n, hipMemcpyDeviceToHost, st));
    CX_HIP(hipMemcpyAsync(ord.data(), t.order.ptr, 8 * n, hipMemcpyDeviceToHost, st));
    CX_HIP(hipMemcpyAsync(inv.data(), t.inv_order.ptr, 8 * n, hipMemcpyDeviceToHost, st));
    CX_HIP(hipMemcpyAsync(lev.data(), t.levels.ptr, 4 * n, hipMemcpyDeviceToHost, st));
    CX_HIP(hipStreamSynchronize(st));

    // 1 removed, 2 detached (a descendant of a removed node)
    std::vector<uint8_t> state(n, 0), removed_ds(n, 0);
    std::vector<int64_t> stack;
    for (int64_t i = 0; i < n_remove; i++) {
        const int64_t x = remove[i];
        CX_REQUIRE(x >= 0 && x < n, "Deletion index out of range.");
        const int64_t r = by_node ? ord[x] : x;
        CX_REQUIRE(!removed_ds[r], "Duplicate deletion indices are not allowed.");
        removed_ds[r] = 1;
        state[inv[r]] = 1;
        stack.push_back(inv[r]);
    }
    auto for_children = [&](int64_t u, auto &&fn) {
        const int64_t first = chi[u];
        for (int64_t c = first, steps = 0; c != -1 && steps < n; steps++) {
            fn(c);
            const int64_t nx = nxt[c];
            if (nx == c || nx == first)
                break;
            c = nx;
        }
    };
    while (!stack.empty()) {
        const int64_t u = stack.back();
        stack.pop_back();
        for_children(u, [&](int64_t c) {
            if (state[c] == 0) {
                state[c] = 2;
                stack.push_back(c);
            }
        });
    }
    // new tree positions: kept nodes in tree order, then detached nodes by
    // (max(level, 0), tree order)
    std::vector<int64_t> newpos(n, -1), src;
    src.reserve(n);
    for (int64_t u = 0; u < n; u++)
        if (state[u] == 0) {
            newpos[u] = (int64_t)src.size();
            src.push_back(u);
        }
    const int64_t kept = (int64_t)src.size();
    std::vector<int64_t> det;
    for (int64_t u = 0; u < n; u++)
        if (state[u] == 2)
            det.push_back(u);
    auto glevel = [&](int64_t u) { return lev[u] > 0 ? lev[u] : 0; };
    std::stable_sort(det.begin(), det.end(),
                     [&](int64_t a, int64_t b) { return glevel(a) < glevel(b); });
    std::vector<int64_t> group_end;
    for (size_t i = 0; i < det.size(); i++) {
        newpos[det[i]] = (int64_t)src.size();
        src.push_back(det[i]);
        if (i + 1 == det.size() || glevel(det[i + 1]) != glevel(det[i]))
            group_end.push_back((int64_t)src.size());
    }
    const int64_t n1 = (int64_t)src.size();
    // dataset rows close up over the removed ones
    std::vector<int64_t> below(n + 1, 0);
    for (int64_t r = 0; r < n; r++)
        below[r + 1] = below[r] + removed_ds[r];
    std::vector<int64_t> ord1(n1), par1(n1, -1), chi1(n1, -1), nxt1(n1, -1);
    std::vector<int32_t> lev1(n1, INT32_MIN);
    for (int64_t i = 0; i < n1; i++)
        ord1[i] = ord[src[i]] - below[ord[src[i]]];
    for (int64_t i = 0; i < kept; i++) {
        const int64_t u = src[i];
        par1[i] = par[u] >= 0 ? newpos[par[u]] : -1;
        lev1[i] = lev[u];
        int64_t last = -1;
        for_children(u, [&](int64_t c) {
            if (state[c] != 0)
                return;
            if (last < 0)
                chi1[i] = newpos[c];
            else
                nxt1[last] = newpos[c];
            last = newpos[c];
        });
    }

    // device arrays in the new tree order
    DevBuf<int64_t> sel(n1 > 0 ? n1 : 1);
    CX_HIP(hipMemcpyAsync(sel.ptr, src.data(), 8 * n1, hipMemcpyHostToDevice, st));
    auto regather = [&](DevBuf<float> &b, int32_t w) {
        if (!b.ptr)
            return;
        DevBuf<float> nb((size_t)(n1 > 0 ? n1 : 1) * w);
        if (n1 > 0)
            gather_rows_kernel<<<(unsigned)n1, 64, 0, st>>>(b.ptr, sel.ptr, n1, w, nb.ptr);
        CX_HIP(hipGetLastError());
        CX_HIP(hipStreamSynchronize(st));
        b = std::move(nb);
    };
    regather(t.raw, t.d);
    if (t.metric == 0) {
        regather(t.v, t.m);
        regather(t.scaled, t.d);
        regather(t.norms, 1);
        regather(t.p, 1);
    }
    auto upload = [&](auto &b, const auto &h) {
        b.alloc(n1 > 0 ? n1 : 1);
        CX_HIP(hipMemcpyAsync(b.ptr, h.data(), sizeof(h[0]) * n1, hipMemcpyHostToDevice, st));
    };
    upload(t.order, ord1);
    upload(t.parents, par1);
    upload(t.children, chi1);
    upload(t.next, nxt1);
    upload(t.levels, lev1);
    t.inv_order.alloc(n1 > 0 ? n1 : 1);
    if (n1 > 0)
        inverse_perm_kernel<<<nblocks(n1, 256), 256, 0, st>>>(t.order.ptr, n1, t.inv_order.ptr);
    CX_HIP(hipGetLastError());
    CX_HIP(hipStreamSynchronize(st));
    t.n = n1;
    t.mixed_ready = false;

    // re-insert the detached groups, then the derived arrays; the kept nodes
    // moved, so only the re-inserted batches stay separation groups
    t.live_batches.clear();
    if (t.metric == 0) {
        const MetricView mv = residual_view(t);
        int64_t g0 = kept;
        for (int64_t g1 : group_end) {
            run_insert<kMetricResidual>(t, mv, 1, t.grid_scale, g0, t.chunk, st, g1);
            g0 = g1;
        }
        finalize_tree<kMetricResidual>(t, mv, st);
    } else {
        const MetricView mv{t.raw.ptr, nullptr, nullptr, nullptr, t.d, 0, 1.0f, 0};
        int64_t g0 = kept;
        for (int64_t g1 : group_end) {
            run_insert<kMetricEuclidean>(t, mv, 0, 1.0f, g0, t.chunk, st, g1);
            g0 = g1;
        }
        finalize_tree<kMetricEuclidean>(t, mv, st);
    }
}

void morton_order_device(const float *coords_host, int64_t n, int32_t d, int64_t *order_host)
{
    hipStream_t st = nullptr;
    DevBuf<float> c((size_t)n * d);
    DevBuf<int64_t> o(n);
    CX_HIP(hipMemcpy(c.ptr, coords_host, sizeof(float) * n * d, hipMemcpyHostToDevice));
    morton_order_on_stream(c.ptr, n, d, o.ptr, st);
    CX_HIP(hipMemcpy(order_host, o.ptr, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
}

void subtree_bounds_device(const int64_t *parents_host, const int64_t *n2d_host, int64_t n,
                           int64_t *min_host, int64_t *max_host)
{
    hipStream_t st = nullptr;
    DevBuf<int64_t> par(n), n2d(n), mn(n), mx(n);
    DevBuf<int32_t> depth(n), top(n);
    CX_HIP(hipMemcpy(par.ptr, parents_host, sizeof(int64_t) * n, hipMemcpyHostToDevice));
    CX_HIP(hipMemcpy(n2d.ptr, n2d_host, sizeof(int64_t) * n, hipMemcpyHostToDevice));
    int32_t md = compute_depths(par.ptr, n, depth.ptr, top.ptr, st);
    subtree_bounds_on_stream(par.ptr, n2d.ptr, depth.ptr, md, n, mn.ptr, mx.ptr, st);
    CX_HIP(hipMemcpy(min_host, mn.ptr, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    CX_HIP(hipMemcpy(max_host, mx.ptr, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
}


// ---- invariant checker ---------------------------------------------------------
// The cover-tree invariants batch insert establishes (the properties of the
// reference's construction, src/algo/batch/mod.rs:447-800, checked the same
// way by the oracle's tests):
// - covering: a node inserted at level L lies within 2^(L+1) of its parent
//   (d^2 <= 4^(L+1), the build's own FP32 distance);
// - nesting: the cover sets only grow — a node's parent was in the tree when
//   the node was inserted (the parent's batch is the node's or an earlier
//   one) and the parent chain reaches a root;
// - separation: nodes one batch inserted at the same level L, strictly between
//   the bottom and top levels, are pairwise farther than 2^L (d^2 > 4^L) and,
//   when the grid conflicts are on, not grid neighbours;
// - forest: the child lists and the parent pointers describe the same forest,
//   and a dropped node (no level) has neither parent nor children.
// A batch's links, levels and distances are final once it is inserted (later
// batches only prepend to child lists), so checking every live batch after an
// insert call is the check after each batch.

// index of the live batch holding position i (batches sorted by start), -1
// when none does
__device__ __forceinline__ int64_t batch_of(const int64_t *__restrict__ batches, int64_t nb,
                                            int64_t i)
{
    int64_t lo = 0, hi = nb;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (batches[2 * mid] <= i)
            lo = mid + 1;
        else
            hi = mid;
    }
    const int64_t b = lo - 1;
    return b >= 0 && i < batches[2 * b] + batches[2 * b + 1] ? b : -1;
}

template <int KIND>
__global__ void check_links_kernel(MetricView mv, const int64_t *__restrict__ parents,
                                   const int64_t *__restrict__ children,
                                   const int64_t *__restrict__ next,
                                   const int32_t *__restrict__ levels, int64_t n,
                                   const int64_t *__restrict__ batches, int64_t nb,
                                   int32_t max_depth, int64_t *cnt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t p = parents[i];
    const int32_t L = levels[i];
    unsigned long long cov = 0, nest = 0, forest = 0, kids = 0;
    if (p >= 0) {
        if (L == kLevelUnset) {
            forest++;
        } else {
            const float d2 = build_dist_sq<KIND>(mv, i, p);
            if (!(d2 <= __builtin_ldexpf(1.0f, 2 * (L + 1))))
                cov++;
            const int64_t bi = batch_of(batches, nb, i), bp = batch_of(batches, nb, p);
            if (bi >= 0 && bp >= 0 && bp > bi)
                nest++;
        }
        // the parent chain reaches a root within the tree's depth
        int64_t a = p;
        for (int32_t steps = 0; a >= 0 && steps <= max_depth; steps++)
            a = parents[a];
        if (a >= 0)
            nest++;
    }
    // the child list: each entry's parent is i (bounded walk)
    int64_t c = children[i];
    for (int64_t steps = 0; c != -1 && steps <= n; steps++) {
        kids++;
        if (c < 0 || c >= n || parents[c] != i) {
            forest++;
            break;
        }
        c = next[c];
    }
    if (c != -1)
        forest++; // a list longer than n entries is a cycle
    if (L == kLevelUnset && kids > 0)
        forest++;
    atomicAdd((unsigned long long *)&cnt[CX_INV_COVERING], cov);
    atomicAdd((unsigned long long *)&cnt[CX_INV_NESTING], nest);
    atomicAdd((unsigned long long *)&cnt[CX_INV_FOREST], forest);
    atomicAdd((unsigned long long *)&cnt[CX_INV_NODES], p >= 0 ? 1ull : 0ull);
    // list entries minus linked nodes: zero when the two describe one forest
    atomicAdd((unsigned long long *)&cnt[CX_INV_FIELDS], kids);
}

// one workgroup per live batch: every pair of its nodes at one level
template <int KIND>
__global__ void check_separation_kernel(MetricView mv, const int32_t *__restrict__ levels,
                                        const float *__restrict__ raw, int32_t gd, float gs,
                                        int32_t use_grid, const int64_t *__restrict__ batches,
                                        int32_t lo, int32_t hi, int64_t *cnt)
{
    const int64_t start = batches[2 * blockIdx.x], size = batches[2 * blockIdx.x + 1];
    unsigned long long viol = 0, pairs = 0;
    for (int64_t a = start + threadIdx.x; a < start + size; a += blockDim.x) {
        const int32_t La = levels[a];
        if (La <= lo || La >= hi)
            continue;
        const float lim = __builtin_ldexpf(1.0f, 2 * La);
        for (int64_t b = a + 1; b < start + size; b++) {
            if (levels[b] != La)
                continue;
            pairs++;
            if (!(build_dist_sq<KIND>(mv, a, b) > lim))
                viol++;
            else if (use_grid && grid_adjacent(raw, gd, gs, a, b))
                viol++;
        }
    }
    atomicAdd((unsigned long long *)&cnt[CX_INV_SEPARATION], viol);
    atomicAdd((unsigned long long *)&cnt[CX_INV_PAIRS], pairs);
}

void check_invariants(TreeState &t, int64_t *out)
{
    hipStream_t st = t.stream;
    const int64_t n = t.n;
    DevBuf<int64_t> cnt(CX_INV_FIELDS + 1);
    CX_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(int64_t) * (CX_INV_FIELDS + 1), st));
    const int64_t nb = (int64_t)t.live_batches.size() / 2;
    DevBuf<int64_t> bat(nb > 0 ? 2 * nb : 2);
    if (nb > 0)
        CX_HIP(hipMemcpyAsync(bat.ptr, t.live_batches.data(), sizeof(int64_t) * 2 * nb,
                              hipMemcpyHostToDevice, st));
    auto run = [&](const MetricView &mv, auto kind) {
        constexpr int K = decltype(kind)::value;
        if (n > 0)
            check_links_kernel<K><<<nblocks(n, 256), 256, 0, st>>>(
                mv, t.parents.ptr, t.children.ptr, t.next.ptr, t.levels.ptr, n, bat.ptr, nb,
                t.max_depth, cnt.ptr);
        if (nb > 0)
            check_separation_kernel<K><<<(unsigned)nb, 256, 0, st>>>(
                mv, t.levels.ptr, t.raw.ptr, t.d, t.grid_scale, t.metric == 0 ? 1 : 0, bat.ptr,
                t.min_level, t.max_level, cnt.ptr);
    };
    if (t.metric == 0)
        run(residual_view(t), std::integral_constant<int, kMetricResidual>{});
    else
        run(MetricView{t.raw.ptr, nullptr, nullptr, nullptr, t.d, 0, 1.0f, 0},
            std::integral_constant<int, kMetricEuclidean>{});
    CX_HIP(hipGetLastError());
    int64_t h[CX_INV_FIELDS + 1];
    CX_HIP(hipMemcpyAsync(h, cnt.ptr, sizeof(h), hipMemcpyDeviceToHost, st));
    CX_HIP(hipStreamSynchronize(st));
    const int64_t list_minus_linked = h[CX_INV_FIELDS] - h[CX_INV_NODES];
    h[CX_INV_FOREST] += list_minus_linked < 0 ? -list_minus_linked : list_minus_linked;
    h[CX_INV_BATCHES] = nb;
    for (int f = 0; f < CX_INV_FIELDS; f++)
        out[f] = h[f];
}

} // namespace cx
