// Device tree build for the residual-correlation (and Euclidean) cover tree.
//
//   morton order  -> hilbert_like_order, src/pcct.rs:27-67
//   gather        -> build_pcct2_residual_tree reorder, src/lib.rs:1404-1443
//                    ResidualMetric::new scaled coords / norms, src/metric.rs:84-177
//   batch insert  -> batch_insert_collect, src/algo/batch/mod.rs:447-800
//   si_cache      -> compute_si_cache_residual, src/algo.rs:52-106
//   subtree bounds-> compute_subtree_index_bounds, src/tree.rs:144-185
//
// Insertion is sequential over batches and levels: each level of each batch
// is a scan over all CUs, a conflict pass over all CUs and one resolving
// workgroup, while the tree stays in HBM.  Distances use the reference f32
// arithmetic (device_common.hpp) so every comparison matches the reference
// bit for bit.
#include <hipcub/hipcub.hpp>

#include <atomic>
#include <string>

#include <algorithm>
#include <cstdlib>

#include "cx_internal.hpp"
#include "device_common.hpp"

namespace cx {

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

// LDS rows of per-chunk dot sums, padded by a word: 64 chunks of eight
// (m <= 512) in the 64-row areas, 32 (m <= 256) in the eight-row areas of
// the fused kernel, whose LDS is nearly full
template <int CR> constexpr int coop_row() { return CR == 64 ? 65 : 33; }
template <int CR> __device__ __forceinline__ bool coop_rank(int32_t m)
{
    return (m & 7) == 0 && m >= 64 && m <= 8 * (coop_row<CR>() - 1);
}

// orders the wavefront's own LDS accesses (no workgroup barrier)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Residual distances d(q, node) of a wavefront's lanes (q the same on every
// lane, node per lane, wanted where live; every lane calls it) with the V-row
// dots eight lanes per pair: pass p gives lane group gr the pair of lane
// 8p + gr, whose chunk sums (reduce8 of eight products each) go to an LDS
// row of coop_row<CR>() floats; each lane then adds its own row in chunk order
// from 0.0 -- dot_slice's operations, so the result has build_dist_sq's
// bits.  CR = 64: cs holds a row per lane (64 rows) and every lane folds
// after the eight passes; CR = 8: cs holds the eight rows of one pass, which
// its owner lanes (8p .. 8p + 7) fold before the next pass.  Needs
// coop_rank<CR>(m).
template <int CR>
__device__ __forceinline__ float coop_wave_dist_sq(const MetricView &mv, int64_t q,
                                                   int64_t node, bool live, float *cs)
{
    static_assert(CR == 64 || CR == 8, "64 rows, or the eight of one pass");
    constexpr int kRow = coop_row<CR>(), kU = (kRow - 1) / 8; // chunks per lane
    const int lane = threadIdx.x & 63, sub = lane & 7, gr = lane >> 3, nch = mv.m >> 3;
    const float *va = mv.v + q * mv.m;
    const uint64_t ball = __ballot(live);
    float acc = 0.0f;
    for (int p = 0; p < 8; p++) {
        const int jj = 8 * p + gr;
        const int64_t nd = __shfl(node, jj, 64);
        float *row = cs + (CR == 64 ? jj : gr) * kRow;
        if ((ball >> jj) & 1ull) {
            const float *vb = mv.v + nd * mv.m;
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int c = sub + 8 * u;
                if (c < nch) {
                    const float4 a0 = *(const float4 *)(va + 8 * c);
                    const float4 a1 = *(const float4 *)(va + 8 * c + 4);
                    const float4 b0 = *(const float4 *)(vb + 8 * c);
                    const float4 b1 = *(const float4 *)(vb + 8 * c + 4);
                    const float pr[8] = {a0.x * b0.x, a0.y * b0.y, a0.z * b0.z, a0.w * b0.w,
                                         a1.x * b1.x, a1.y * b1.y, a1.z * b1.z, a1.w * b1.w};
                    row[c] = reduce8(pr);
                }
            }
        }
        if (CR == 8) {
            wave_lds_sync();
            if (gr == p && live) {
                const float *own = cs + sub * kRow;
                for (int c = 0; c < nch; c++)
                    acc += own[c];
            }
            wave_lds_sync(); // the next pass rewrites the rows
        }
    }
    if (CR == 64) {
        wave_lds_sync();
        if (live)
            for (int c = 0; c < nch; c++)
                acc += cs[lane * kRow + c];
        wave_lds_sync(); // the next call rewrites the rows
    }
    return live ? dist_sq_ref_dot(mv, q, node, acc) : 0.0f;
}

// build_dist_sq(q, node) where live (0 elsewhere), q the same on every lane;
// every lane calls it.  With an LDS area (cs: CR rows of coop_row<CR>()
// floats) and a rank coop_rank<CR> takes, the dots go eight lanes per pair
// (coop_wave_dist_sq).
template <int KIND, int CR = 64>
__device__ __forceinline__ float wave_dist_sq(const MetricView &mv, int64_t q, int64_t node,
                                              bool live, float *cs)
{
    if (KIND == kMetricResidual && cs != nullptr && coop_rank<CR>(mv.m))
        return coop_wave_dist_sq<CR>(mv, q, node, live, cs);
    return live ? build_dist_sq<KIND>(mv, q, node) : 0.0f;
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

// Ordered compaction over a workgroup of WAVES wavefronts: returns the
// position of this thread's flagged item among flagged items of lower thread
// id; *total receives the count.
template <int WAVES> __device__ __forceinline__ int wg_prefix(bool flag, int *s_cnt, int *total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t mask = __ballot(flag);
    const int pre = __popcll(mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (lane == 0)
        s_cnt[wid] = __popcll(mask);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        const int c = s_cnt[w];
        off += (w < wid) ? c : 0;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + pre;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t u = __shfl_xor(v, o, 64);
        v = u < v ? u : v;
    }
    return v;
}

// ---- batch insert ------------------------------------------------------------
//
// batch_insert_collect (src/algo/batch/mod.rs:447-800) is sequential over
// batches and, inside a batch, over levels 9..-10.  Each level is three
// launches on the tree's stream:
//
//   bi_scan_kernel     (all CUs)  near/far scan: for every candidate, the
//                                 nearest of its cover and the cover's
//                                 children within 2^level (580-627); the
//                                 root's children are kept as an array, so
//                                 the long root list is scored by many
//                                 wavefronts at once
//   bi_pairs_kernel    (all CUs)  d^2 of every pair of far candidates and the
//                                 conflict bits: grid-adjacent or d^2 <= r^2
//                                 (370-400, 662-674)
//   bi_resolve_kernel  (one WG)   greedy MIS in far order over the conflict
//                                 bits (compute_mis_greedy, mis.rs:9-59),
//                                 leader links (694-714), non-leader
//                                 attachment from the stored distances
//                                 (716-769), the next level's candidates,
//                                 and at level -10 the bottom survivors
//                                 (791-800) and the batch's counters
//
// Every distance is the reference's f32 arithmetic (device_common.hpp) and
// every choice keeps the reference's order (strict <, first in list / far /
// leader order wins, links prepended in the reference's link order), so the
// tree equals the sequential build's bit for bit.  The batch state between
// launches lives in a BiState in device memory.

constexpr int kBiMaxChunk = 1024;
constexpr int kBiWords = kBiMaxChunk / 64;   // conflict words per far row
constexpr int kBiResThreads = 1024;          // resolve workgroup
constexpr int kBiResWaves = kBiResThreads / 64;
constexpr int kBiParThreads = 256;           // scan / pair workgroups
constexpr int kBiParWaves = kBiParThreads / 64;
constexpr int kBiParBlocks = 512;
// upper-triangle conflict words of kBiMaxChunk far rows: row i keeps words
// (i / 64) .. W-1, i.e. 64 * W * (W + 1) / 2 words for W = 16
constexpr int kBiAdjWords = 64 * kBiWords * (kBiWords + 1) / 2;
constexpr uint64_t kBiFar = ~0ull; // scan key of a candidate with no cover within r

struct BiState {
    int64_t s, nb;  // first row and size of the batch
    int64_t bidx;   // batch index within the insert run (advanced at level -10)
    // scan evaluations made inside bi_levels_kernel over the insert run (the
    // host reads it to choose between the fused and the wide levels)
    int64_t ev_fused, ev_fused_pairs;
    int32_t nc;     // candidates at the current level
    int32_t nrk;    // entries of the root's child array (oldest link first)
    int64_t ev_scan, ev_mis, ev_nl, orphans, e_grid, e_dist, leaders, far_levels;
    uint64_t tk_begin, tk_mark, tk_pairs, tk_scan, tk_conf, tk_mis, tk_nl;
    int32_t cand[kBiMaxChunk], cover[kBiMaxChunk];
    // scan result: (d^2 bits << 32) | position in [cover, children...], or kBiFar
    uint64_t key[kBiMaxChunk];
    // winner node and children scanned, for candidates whose cover is not the root
    int32_t node[kBiMaxChunk], len[kBiMaxChunk];
    // distance evaluations of levels 9..0 per candidate when one launch ran them
    int64_t ev_pre[kBiMaxChunk];
    // Incremental scans.  stay[i] = the candidate's previous scan chose its
    // cover itself (position 0), so the cover is unchanged and every child it
    // scanned then had d^2 >= d0[i] = d^2(candidate, cover): none of them can
    // win now (position 0 wins ties; if d0 > r^2 they are beyond r^2 too), and
    // only the children linked since (the list before phead[i], or root
    // positions 1 .. nrk - plen[i]) are scored.  plen[i] = the list length at
    // that scan, which the evaluation count still includes.
    int32_t stay[kBiMaxChunk], plen[kBiMaxChunk], phead[kBiMaxChunk];
    float d0[kBiMaxChunk];
    // what this level's scan saw: d^2(candidate, cover) and the list head
    float sd0[kBiMaxChunk];
    int32_t shead[kBiMaxChunk];
};

// Offset of far row i's first stored word (word i / 64) in the packed upper
// triangle of W words per row.
__device__ __forceinline__ int adj_off(int i, int W)
{
    const int b = i >> 6;
    return 64 * (b * W - (b * (b - 1)) / 2) + (i & 63) * (W - b);
}

// Node 0's child list into rk, oldest link first (list order is newest
// first: link_child prepends, src/tree.rs:98-103).  A list holds at most
// n_nodes entries; the walk stops there.
__global__ void bi_root_kids_kernel(TreeLinks t, BiState *S, int32_t *__restrict__ rk,
                                    int64_t n_nodes)
{
    const int64_t first = t.children[0];
    int32_t cnt = 0;
    for (int64_t c = first; c != -1 && cnt < n_nodes;) {
        cnt++;
        const int64_t nx = t.next[c];
        if (nx == c || nx == first)
            break;
        c = nx;
    }
    int64_t c = first;
    for (int32_t i = 0; i < cnt; i++) {
        rk[cnt - 1 - i] = (int32_t)c;
        c = t.next[c];
    }
    S->nrk = cnt;
}

// add_point(row, i32::MIN, -1) for the batch's rows (511-513); the root
// bootstrap at level 10 when the batch starts at row 0 (521-531); every
// candidate starts at cover 0 (539-540).
// The batch's first row is s_base + bidx * chunk, so one recorded launch
// sequence serves every full batch of the run.
__global__ void __launch_bounds__(kBiResThreads)
    bi_begin_kernel(TreeLinks t, BiState *S, int64_t s_base, int64_t chunk, int64_t nb)
{
    const int tid = threadIdx.x;
    const int64_t s = s_base + S->bidx * chunk;
    for (int i = tid; i < nb; i += kBiResThreads) {
        st_rx(&t.parents[s + i], (int64_t)-1);
        st_rx(&t.children[s + i], (int64_t)-1);
        st_rx(&t.next[s + i], (int64_t)-1);
        st_rx(&t.levels[s + i], kLevelUnset);
    }
    const int first = s == 0 ? 1 : 0;
    const int nc = (int)nb - first;
    for (int i = tid; i < nc; i += kBiResThreads) {
        S->cand[i] = (int32_t)(s + first + i);
        S->cover[i] = 0;
        S->key[i] = kBiFar;
        S->ev_pre[i] = 0;
        S->stay[i] = 0;
    }
    if (tid == 0) {
        if (s == 0) {
            st_rx(&t.levels[0], (int32_t)10);
            S->nrk = 0;
        }
        S->s = s;
        S->nb = nb;
        S->nc = nc;
        S->ev_scan = S->ev_mis = S->ev_nl = S->orphans = 0;
        S->e_grid = S->e_dist = S->leaders = S->far_levels = 0;
        const uint64_t now = wall_clock64();
        S->tk_begin = S->tk_mark = S->tk_pairs = now;
        S->tk_scan = S->tk_conf = S->tk_mis = S->tk_nl = 0;
    }
}

// Near/far scan (580-627).  Work items are (candidate i, group g): when the
// cover is the root, group g scores positions 64g .. 64g+63 of [root,
// children...] (one per lane) and folds its best into key[i] with a 64-bit
// atomic min, which picks the smallest d^2 and, among equal ones, the first
// position, as the sequential strict-< scan does; otherwise group 0 walks
// the cover's child list, 64 positions per pass, and stores the result.  A
// candidate that stayed at its cover scores only the children linked since
// its previous scan (BiState::stay).
// wave0 / nw: this wavefront's first work item and the stride (the number of
// wavefronts sharing the scan).
template <int KIND, int CR = 64>
__device__ __forceinline__ void bi_scan_body(const MetricView &mv, const TreeLinks &t, BiState *S,
                                             const int32_t *__restrict__ rk, float r2,
                                             int64_t n_nodes, int64_t wave0, int64_t nw,
                                             float *cs = nullptr)
{
    const int nc = S->nc;
    if (nc == 0)
        return;
    const int lane = threadIdx.x & 63;
    const int nrk = S->nrk;
    const int R = (nrk + 64) / 64;
    const int64_t items = (int64_t)nc * R;
    for (int64_t it = wave0; it < items; it += nw) {
        const int i = (int)(it / R), g = (int)(it % R);
        const int64_t q = S->cand[i], p = S->cover[i];
        const bool st = S->stay[i] != 0;
        if (p == 0) {
            // positions past kn are children an earlier scan already lost to d0
            const int kn = st ? nrk - S->plen[i] : nrk;
            if (g * 64 > kn)
                continue;
            const int pos = g * 64 + lane;
            uint64_t key = kBiFar;
            const bool cached = pos == 0 && st;
            const int64_t node = (pos == 0 || pos > kn) ? 0 : rk[nrk - pos];
            float dc = wave_dist_sq<KIND, CR>(mv, q, node, pos <= kn && !cached, cs);
            if (pos <= kn) {
                if (pos == 0) {
                    if (cached)
                        dc = S->d0[i];
                    S->sd0[i] = dc;
                }
                if (dc <= r2)
                    key = ((uint64_t)__float_as_uint(dc) << 32) | (uint32_t)pos;
            }
            key = wave_min_u64(key);
            if (lane == 0 && key != kBiFar)
                atomicMin((unsigned long long *)&S->key[i], (unsigned long long)key);
            continue;
        }
        if (g != 0)
            continue;
        const int64_t first = t.children[p];
        // a stayed candidate walks only up to the head its previous scan saw
        const int64_t stop = st ? (int64_t)S->phead[i] : (int64_t)-2;
        int64_t c = first;
        bool more = c != -1 && c != stop;
        uint64_t best = kBiFar;
        int64_t bnode = -1;
        int32_t count = 0;
        float d0 = 0.0f;
        for (int base = 0;; base += 64) {
            int64_t mine = (base == 0 && lane == 0) ? p : -1;
            for (int l = base == 0 ? 1 : 0; l < 64 && more; l++) {
                if (lane == l)
                    mine = c;
                count++;
                const int64_t nx = t.next[c];
                if (nx == c || nx == first || nx == -1 || nx == stop || count >= n_nodes)
                    more = false;
                c = nx;
            }
            uint64_t key = kBiFar;
            const bool cached = base == 0 && lane == 0 && st;
            float dc =
                wave_dist_sq<KIND, CR>(mv, q, mine >= 0 ? mine : q, mine >= 0 && !cached, cs);
            if (mine >= 0) {
                if (base == 0 && lane == 0) {
                    if (cached)
                        dc = S->d0[i];
                    d0 = dc;
                }
                if (dc <= r2)
                    key = ((uint64_t)__float_as_uint(dc) << 32) | (uint32_t)(base + lane);
            }
            const uint64_t wmin = wave_min_u64(key);
            if (wmin < best) {
                const int src = __ffsll((unsigned long long)__ballot(key == wmin)) - 1;
                best = wmin;
                bnode = __shfl(mine, src, 64);
            }
            if (!more)
                break;
        }
        if (lane == 0) {
            S->key[i] = best;
            S->node[i] = (int32_t)bnode;
            S->len[i] = count + (st ? S->plen[i] : 0);
            S->sd0[i] = d0;
            S->shead[i] = (int32_t)first;
        }
    }
}

template <int KIND>
__global__ void __launch_bounds__(kBiParThreads)
    bi_scan_kernel(MetricView mv, TreeLinks t, BiState *S, const int32_t *__restrict__ rk,
                   float r2, int64_t n_nodes)
{
    __shared__ float s_cs[kBiParWaves * 64 * coop_row<64>()];
    bi_scan_body<KIND>(mv, t, S, rk, r2, n_nodes,
                       (int64_t)blockIdx.x * kBiParWaves + (threadIdx.x >> 6),
                       (int64_t)gridDim.x * kBiParWaves,
                       s_cs + (threadIdx.x >> 6) * 64 * coop_row<64>());
}

// Level 9 of a residual batch: every candidate starts at the root, so the
// scan is the dense block (candidates) x (root, root's children).  A
// workgroup takes 64 candidates x 64 positions (position tiles grid-strided),
// a thread 4 x 4 pairs: the V dots in the reference's f32 order (dot_slice:
// chunks of eight, each summed as two in-order quadruples and added in order,
// then the tail) from LDS-staged chunks, so each row is read once per tile
// rather than once per pair; then the rest of dist_sq_ref.  Each candidate's
// best (d^2 bits, position) is folded into its key with an atomic min (first
// minimum wins, as in the sequential scan); d^2 to the root itself goes to
// sd0.
constexpr int kRtTile = 64;
// below this many root children the descent scans the root's list itself
// (one workgroup per candidate is then quicker than the tiled block)
constexpr int kRtMin = 2048;
__global__ void __launch_bounds__(256)
    bi_rootscan_kernel(MetricView mv, BiState *S, const int32_t *__restrict__ rk, float r2)
{
    __shared__ float sC[kRtTile][9], sP[kRtTile][9]; // a chunk of eight (+1 pad)
    __shared__ int64_t sCand[kRtTile], sNode[kRtTile];
    const int nc = S->nc;
    const int c0 = (int)blockIdx.y * kRtTile;
    if (c0 >= nc || S->nrk < kRtMin)
        return;
    const int nrk = S->nrk, np = nrk + 1;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int32_t m = mv.m;
    const int32_t tail8 = (m / 8) * 8;
    if (tid < kRtTile)
        sCand[tid] = S->cand[c0 + tid < nc ? c0 + tid : c0];
    for (int p0 = (int)blockIdx.x * kRtTile; p0 < np; p0 += (int)gridDim.x * kRtTile) {
        __syncthreads(); // the previous tile is done with sNode
        if (tid < kRtTile) {
            const int pos = p0 + tid;
            sNode[tid] = pos < np ? (pos == 0 ? 0 : rk[nrk - pos]) : 0;
        }
        __syncthreads();
        float acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++)
                acc[a][b] = 0.0f;
        for (int32_t k0 = 0; k0 < tail8; k0 += 8) {
            {
                // threads 0..127 stage the candidate rows, 128..255 the
                // position rows: four floats each
                const int r = (tid & 127) >> 1, h = tid & 1;
                const int64_t row = tid < 128 ? sCand[r] : sNode[r];
                const float *src = mv.v + row * m + k0 + 4 * h;
                float *dst = tid < 128 ? &sC[r][4 * h] : &sP[r][4 * h];
                if ((m & 3) == 0) {
                    const float4 q = *(const float4 *)src;
                    dst[0] = q.x;
                    dst[1] = q.y;
                    dst[2] = q.z;
                    dst[3] = q.w;
                } else {
                    dst[0] = src[0];
                    dst[1] = src[1];
                    dst[2] = src[2];
                    dst[3] = src[3];
                }
            }
            __syncthreads();
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    float pr[8];
#pragma unroll
                    for (int l = 0; l < 8; l++)
                        pr[l] = sC[4 * ty + a][l] * sP[4 * tx + b][l];
                    acc[a][b] += reduce8(pr);
                }
            __syncthreads();
        }
        for (int32_t jt = tail8; jt < m; jt++)
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[a][b] += mv.v[sCand[4 * ty + a] * m + jt] * mv.v[sNode[4 * tx + b] * m + jt];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int i = c0 + 4 * ty + a;
            uint64_t best = kBiFar;
            if (i < nc) {
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const int pos = p0 + 4 * tx + b;
                    if (pos >= np)
                        continue;
                    const float ds = dist_sq_ref_dot(mv, sCand[4 * ty + a], sNode[4 * tx + b],
                                                     acc[a][b]);
                    if (pos == 0)
                        S->sd0[i] = ds;
                    if (ds <= r2) {
                        const uint64_t key = ((uint64_t)__float_as_uint(ds) << 32) | (uint32_t)pos;
                        best = key < best ? key : best;
                    }
                }
            }
            // the 16 lanes of this candidate row (lanes 16 ty' .. 16 ty' + 15)
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) {
                const uint64_t u = __shfl_xor(best, o, 16);
                best = u < best ? u : best;
            }
            if (tx == 0 && best != kBiFar)
                atomicMin((unsigned long long *)&S->key[i], (unsigned long long)best);
        }
    }
}

// Levels 9..bottom where every candidate is near: under the residual metric
// d^2 = 1 - |clamp(rho)| lies in [0, 1] <= 4^level for level >= 0 (bottom 0);
// under the Euclidean metric the FP32 d^2 of any two points is at most the
// FP32 fold of the squared bounding-box edges (subtraction, squaring and
// addition of non-negative terms are monotone under rounding), so every
// level whose r^2 covers that bound qualifies.  At these levels no leader is
// linked and each candidate's descent depends only on the tree as the batch
// found it: one workgroup per candidate runs the scans back to back (the
// same positions, keys and first-minimum choice as bi_scan_kernel), taking
// level 9 from bi_rootscan_kernel when the root's list is long (residual).
// Once a scan chooses the cover itself, every later scan down to bottom
// would repeat it (same list, every d^2 within r^2), so they are only
// counted.  The candidate is left at its cover for level bottom - 1 with its
// evaluations in ev_pre and its incremental-scan state.
template <int KIND>
__global__ void __launch_bounds__(kBiParThreads)
    bi_descend_kernel(MetricView mv, TreeLinks t, BiState *S, const int32_t *__restrict__ rk,
                      int64_t n_nodes, int32_t bottom)
{
    __shared__ unsigned long long s_wmin[kBiParWaves];
    __shared__ int64_t s_cover, s_len, s_head;
    __shared__ int32_t s_pos;
    __shared__ float s_d0;
    __shared__ float s_cs[kBiParWaves * 64 * coop_row<64>()];
    const int i = (int)blockIdx.x;
    if (i >= S->nc)
        return;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t q = S->cand[i];
    const int nrk = S->nrk;
    int64_t p = 0, ev = 0;
    int32_t top = 9;
    if (KIND == kMetricResidual && nrk >= kRtMin) {
        // level 9 from bi_rootscan_kernel: a root child, or the root itself
        const int pos = (int)(uint32_t)S->key[i];
        ev = 1 + nrk;
        if (pos == 0) { // every scan down to level 0 keeps the root
            if (tid == 0) {
                S->cover[i] = 0;
                S->key[i] = kBiFar;
                S->ev_pre[i] = (int64_t)(10 - bottom) * (1 + nrk);
                S->stay[i] = 1;
                S->d0[i] = S->sd0[i];
                S->plen[i] = nrk;
                S->phead[i] = -1;
            }
            return;
        }
        p = rk[nrk - pos];
        top = 8;
    }
    for (int32_t level = top; level >= bottom; --level) {
        const float r = ldexpf(1.0f, level), r2 = r * r;
        if (p == 0) {
            uint64_t best = kBiFar;
            // wave-uniform trips: every lane calls wave_dist_sq
            for (int base = wid * 64; base <= nrk; base += kBiParThreads) {
                const int pos = base + lane;
                const bool in = pos <= nrk;
                const int64_t c = (pos == 0 || !in) ? 0 : rk[nrk - pos];
                const float dc =
                    wave_dist_sq<KIND>(mv, q, c, in, s_cs + wid * 64 * coop_row<64>());
                if (pos == 0)
                    s_d0 = dc;
                if (in && dc <= r2) {
                    const uint64_t key = ((uint64_t)__float_as_uint(dc) << 32) | (uint32_t)pos;
                    best = key < best ? key : best;
                }
            }
            best = wave_min_u64(best);
            if (lane == 0)
                s_wmin[wid] = best;
            __syncthreads();
            if (tid == 0) {
                uint64_t b = s_wmin[0];
                for (int w = 1; w < kBiParWaves; w++)
                    b = s_wmin[w] < b ? s_wmin[w] : b;
                const int pos = (int)(uint32_t)b;
                s_cover = pos == 0 ? 0 : rk[nrk - pos];
                s_pos = pos;
                s_len = nrk;
                s_head = -1;
            }
        } else if (wid == 0) {
            const int64_t first = t.children[p];
            int64_t c = first;
            bool more = c != -1;
            uint64_t best = kBiFar;
            int64_t bnode = -1;
            int32_t count = 0;
            for (int base = 0;; base += 64) {
                int64_t mine = (base == 0 && lane == 0) ? p : -1;
                for (int l = base == 0 ? 1 : 0; l < 64 && more; l++) {
                    if (lane == l)
                        mine = c;
                    count++;
                    const int64_t nx = t.next[c];
                    if (nx == c || nx == first || nx == -1 || count >= n_nodes)
                        more = false;
                    c = nx;
                }
                uint64_t key = kBiFar;
                const float dc = wave_dist_sq<KIND>(mv, q, mine >= 0 ? mine : q, mine >= 0, s_cs);
                if (mine >= 0) {
                    if (base == 0 && lane == 0)
                        s_d0 = dc;
                    if (dc <= r2)
                        key = ((uint64_t)__float_as_uint(dc) << 32) | (uint32_t)(base + lane);
                }
                const uint64_t wmin = wave_min_u64(key);
                if (wmin < best) {
                    const int src = __ffsll((unsigned long long)__ballot(key == wmin)) - 1;
                    best = wmin;
                    bnode = __shfl(mine, src, 64);
                }
                if (!more)
                    break;
            }
            if (lane == 0) {
                s_cover = bnode;
                s_pos = (int32_t)(uint32_t)best;
                s_len = count;
                s_head = first;
            }
        }
        __syncthreads();
        const int64_t len = s_len;
        p = s_cover;
        const bool stayed = s_pos == 0;
        __syncthreads();
        // this level and, when the cover stayed, levels level-1 .. bottom as well
        ev += (stayed ? (int64_t)(level - bottom) + 1 : 1) * (1 + len);
        if (stayed || level == bottom) {
            if (tid == 0) {
                S->cover[i] = (int32_t)p;
                S->key[i] = kBiFar;
                S->ev_pre[i] = ev;
                S->stay[i] = stayed ? 1 : 0;
                S->d0[i] = s_d0;
                S->plen[i] = (int32_t)len;
                S->phead[i] = (int32_t)s_head;
            }
            return;
        }
    }
}

// The far candidates, in candidate order, into s_far; returns their count
// (a workgroup of WAVES wavefronts).
template <int WAVES>
__device__ __forceinline__ int far_list(const BiState *S, int nc, int32_t *s_far, int *s_cnt)
{
    int nf = 0;
    for (int base = 0; base < nc; base += WAVES * 64) {
        const int i = base + (int)threadIdx.x;
        const bool f = i < nc && S->key[i] == kBiFar;
        int tot;
        const int pos = wg_prefix<WAVES>(f, s_cnt, &tot);
        if (f)
            s_far[nf + pos] = S->cand[i];
        nf += tot;
    }
    __syncthreads();
    return nf;
}

// Conflicts among far candidates: for far indices i < j, dist[i][j] = d^2
// and bit j of row i's conflict word = grid-adjacent || d^2 <= r^2, with the
// grid bits alone in gadj.  Work items are (i, word w >= i / 64).
// A workgroup of WAVES wavefronts; wave0 / nw as in bi_scan_body; stamp: this
// workgroup records the pass's start tick.
template <int KIND, int WAVES, int CR = 64>
__device__ __forceinline__ void bi_pairs_body(const MetricView &mv, BiState *S,
                                              const float *__restrict__ raw, int32_t gd,
                                              float gs, int32_t use_grid, float r2,
                                              float *__restrict__ dist,
                                              uint64_t *__restrict__ adj,
                                              uint64_t *__restrict__ gadj, int32_t *s_far,
                                              int *s_cnt, int wave0, int nw, bool stamp,
                                              float *cs = nullptr)
{
    if (stamp && threadIdx.x == 0)
        S->tk_pairs = wall_clock64();
    const int nc = S->nc;
    if (nc < 2)
        return;
    const int nf = far_list<WAVES>(S, nc, s_far, s_cnt);
    if (nf < 2)
        return;
    const int lane = threadIdx.x & 63;
    const int W = (nf + 63) >> 6;
    const int items = nf * W;
    for (int it = wave0; it < items; it += nw) {
        const int i = it / W, w = it - i * W;
        if (w < (i >> 6))
            continue;
        const int j = w * 64 + lane;
        bool a = false, g = false;
        const bool live = j > i && j < nf;
        const int64_t x = s_far[i], y = live ? s_far[j] : x;
        const float d = wave_dist_sq<KIND, CR>(mv, x, y, live, cs);
        if (live) {
            if (use_grid)
                g = grid_adjacent(raw, gd, gs, x, y);
            a = g || d <= r2;
            dist[(size_t)i * kBiMaxChunk + j] = d;
        }
        const uint64_t ab = __ballot(a), gb = __ballot(g);
        if (lane == 0) {
            adj[i * kBiWords + w] = ab;
            gadj[i * kBiWords + w] = gb;
        }
    }
}

template <int KIND>
__global__ void __launch_bounds__(kBiParThreads)
    bi_pairs_kernel(MetricView mv, BiState *S, const float *__restrict__ raw, int32_t gd,
                    float gs, int32_t use_grid, float r2, float *__restrict__ dist,
                    uint64_t *__restrict__ adj, uint64_t *__restrict__ gadj)
{
    __shared__ int32_t s_far[kBiMaxChunk];
    __shared__ int s_cnt[kBiParWaves];
    __shared__ float s_cs[kBiParWaves * 64 * coop_row<64>()];
    bi_pairs_body<KIND, kBiParWaves>(mv, S, raw, gd, gs, use_grid, r2, dist, adj, gadj, s_far,
                                     s_cnt, (int)blockIdx.x * kBiParWaves + (int)(threadIdx.x >> 6),
                                     (int)gridDim.x * kBiParWaves, blockIdx.x == 0,
                                     s_cs + (threadIdx.x >> 6) * 64 * coop_row<64>());
}

// *acc (LDS) += the sum of v over the workgroup.
__device__ __forceinline__ void wg_add_1024(int64_t v, int64_t *s_red, int64_t *acc)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    if (lane == 0)
        s_red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t r = 0;
        for (int w = 0; w < kBiResWaves; w++)
            r += s_red[w];
        *acc += r;
    }
    __syncthreads();
}

struct LinkScratch {
    int32_t prev[kBiMaxChunk];
    uint8_t last[kBiMaxChunk];
    int cnt[kBiResWaves];
    int nrk;
};

// set_level + set_parent for nodes[0..cnt) in this order (src/tree.rs:
// 105-120): each link prepends to the parent's list, so a parent's run of
// nodes ends up newest first ahead of its old head.  In parallel: each node
// finds the previous node of its run (a wavefront compares 64 earlier
// parents per step), old heads are read before any write, and node 0's new
// children are appended to rk in order.
__device__ void link_in_order(TreeLinks t, int32_t *__restrict__ rk, const int32_t *nodes,
                              const int32_t *pars, int cnt, int32_t level, LinkScratch &L)
{
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid < cnt)
        L.last[tid] = 1;
    __syncthreads();
    for (int k = wid; k < cnt; k += kBiResWaves) {
        const int32_t P = pars[k];
        int prev = -1;
        for (int base = k - 1; base >= 0; base -= 64) {
            const int idx = base - lane;
            const uint64_t b = __ballot(idx >= 0 && pars[idx] == P);
            if (b) {
                prev = base - (__ffsll((unsigned long long)b) - 1);
                break;
            }
        }
        if (lane == 0) {
            L.prev[k] = prev;
            if (prev >= 0)
                L.last[prev] = 0;
        }
    }
    __syncthreads();
    int64_t old = -1;
    if (tid < cnt && L.prev[tid] < 0)
        old = ld_rx(&t.children[pars[tid]]);
    __syncthreads();
    if (tid < cnt) {
        const int32_t v = nodes[tid], P = pars[tid];
        st_rx(&t.parents[v], (int64_t)P);
        st_rx(&t.levels[v], level);
        st_rx(&t.next[v], L.prev[tid] >= 0 ? (int64_t)nodes[L.prev[tid]] : old);
        if (L.last[tid])
            st_rx(&t.children[P], (int64_t)v);
    }
    const bool r0 = tid < cnt && pars[tid] == 0;
    int nr;
    const int pr = wg_prefix<kBiResWaves>(r0, L.cnt, &nr);
    const int base = L.nrk;
    if (r0)
        rk[base + pr] = nodes[tid];
    __syncthreads();
    if (tid == 0)
        L.nrk = base + nr;
    __syncthreads();
}

// Optional per-level record of a batch's conflict graph, for the reference's
// BatchInsertTelemetry payload (src/algo/batch/mod.rs:802-956).  buf[0..1]
// is a 64-bit cursor counting the words used after that header; entries are
//   level:  1, batch, level, nf, far node ids [nf], leader bits [W],
//           covered-non-leader bits [W], per far row i: distance-conflict
//           bits [W] (d^2 <= r^2, columns j > i) for every row, then grid
//           bits [W] for every row (W = ceil(nf / 32) words; bit j of row i
//           in word j / 32)
//   bottom: 2, batch, count, survivor node ids [count]
// A run that needs more than cap words leaves the cursor past cap and
// writes nothing past it (the host reports the overflow).
struct ConflictLog {
    int32_t *buf; // null: no record
    int64_t cap;  // words after the header
    int32_t batch0;
};

// The rest of a level after the scan and the conflict pass: near/far split,
// greedy MIS, leaders, non-leaders, the next level's candidates; at level
// -10 also the bottom survivors and the batch's counters (bs: its row of
// kBatchFields, or null).
template <int KIND>
__device__ __forceinline__ void bi_resolve_body(const TreeLinks &t, BiState *S,
                                                int32_t *__restrict__ rk,
                                                const float *__restrict__ dist,
                                                const uint64_t *__restrict__ adj,
                                                const uint64_t *__restrict__ gadj, int32_t level,
                                                float r2, int64_t *__restrict__ bs,
                                                const ConflictLog &CL)
{
    __shared__ int32_t s_nxc[kBiMaxChunk], s_nxv[kBiMaxChunk]; // next level
    __shared__ int32_t s_nxst[kBiMaxChunk], s_nxpl[kBiMaxChunk], s_nxph[kBiMaxChunk];
    __shared__ float s_nxd0[kBiMaxChunk];
    __shared__ int32_t s_far[kBiMaxChunk], s_fcov[kBiMaxChunk];
    __shared__ int32_t s_lead[kBiMaxChunk], s_lfar[kBiMaxChunk], s_lpar[kBiMaxChunk];
    __shared__ int32_t s_nlf[kBiMaxChunk];
    __shared__ unsigned long long s_nlkey[kBiMaxChunk];
    __shared__ uint64_t s_adj[kBiAdjWords];
    __shared__ uint64_t s_sel[kBiWords];
    __shared__ int64_t s_red[kBiResWaves];
    __shared__ int64_t s_acc[3]; // scan evaluations, grid edges, distance-only edges
    __shared__ LinkScratch L;
    __shared__ int64_t s_log; // this level's entry in the conflict record (-1: none)

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint64_t tk0 = wall_clock64();
    const int nc = S->nc;
    const int nrk0 = S->nrk;
    if (tid == 0) {
        L.nrk = nrk0;
        s_acc[0] = s_acc[1] = s_acc[2] = 0;
    }
    int64_t ev = 0;
    bool near = false;
    int32_t cand = 0, cover = 0, node = -1;
    // the next level's incremental-scan state of a near candidate
    int32_t nst = 0, npl = 0, nph = -1;
    float nd0 = 0.0f;
    if (tid < nc) {
        cand = S->cand[tid];
        cover = S->cover[tid];
        const uint64_t key = S->key[tid];
        near = key != kBiFar;
        const int pos = (int)(uint32_t)key;
        if (cover == 0) {
            ev = 1 + nrk0;
            npl = nrk0;
            if (near)
                node = pos == 0 ? 0 : rk[nrk0 - pos];
        } else {
            npl = S->len[tid];
            ev = 1 + npl;
            node = S->node[tid];
            nph = S->shead[tid];
        }
        nst = near && pos == 0;
        nd0 = S->sd0[tid];
        ev += S->ev_pre[tid];
        S->ev_pre[tid] = 0;
    }
    wg_add_1024(ev, s_red, &s_acc[0]);
    int nn;
    const int pn = wg_prefix<kBiResWaves>(near, L.cnt, &nn);
    if (near) {
        s_nxc[pn] = cand;
        s_nxv[pn] = node;
        s_nxst[pn] = nst;
        s_nxd0[pn] = nd0;
        s_nxpl[pn] = npl;
        s_nxph[pn] = nph;
    } else if (tid < nc) {
        s_far[tid - pn] = cand;
        s_fcov[tid - pn] = cover;
    }
    const int nf = nc - nn;
    __syncthreads();
    int64_t leaders = 0, orphans = 0, ev_mis = 0, ev_nl = 0;
    uint64_t tk_mis = 0, tk_nl = 0;
    if (nf > 0) {
        const int W = (nf + 63) >> 6;
        int64_t eg = 0, ed = 0;
        for (int idx = tid; idx < nf * W; idx += kBiResThreads) {
            const int i = idx / W, w = idx - i * W;
            if (w < (i >> 6))
                continue;
            uint64_t a = 0, g = 0;
            if (nf > 1) {
                a = adj[i * kBiWords + w];
                g = gadj[i * kBiWords + w];
            }
            s_adj[adj_off(i, W) + w - (i >> 6)] = a;
            eg += __popcll(g);
            ed += __popcll(a & ~g);
        }
        wg_add_1024(eg, s_red, &s_acc[1]);
        wg_add_1024(ed, s_red, &s_acc[2]);
        ev_mis = (int64_t)nf * (nf - 1) / 2;
        // greedy MIS in far order: take the first undecided candidate, drop
        // every later candidate in conflict with it
        // Word by word (64 candidates): the block's own choices run in
        // registers (lane b holds candidate 64 f + b's conflict word f, read
        // by the chosen lane), then every later word drops the conflicts of
        // the block's leaders; the same leaders as one candidate at a time.
        if (wid == 0) {
            uint64_t U = 0, sel = 0;
            if (lane < W)
                U = (lane < W - 1 || (nf & 63) == 0) ? ~0ull : ((1ull << (nf & 63)) - 1);
            for (int f = 0; f < W; f++) {
                uint64_t u = __shfl(U, f, 64); // undecided after the earlier words
                const int i = f * 64 + lane;
                const uint64_t A = i < nf ? s_adj[adj_off(i, W)] : 0ull;
                const uint32_t a_lo = (uint32_t)A, a_hi = (uint32_t)(A >> 32);
                uint64_t sf = 0;
                while (u != 0) {
                    const int b = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)u) - 1);
                    sf |= 1ull << b;
                    const uint64_t row =
                        (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)a_lo, b) |
                        ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)a_hi, b) << 32);
                    u &= ~((1ull << b) | row);
                }
                if (lane == f) {
                    sel = sf;
                    U = 0;
                } else if (lane > f && lane < W) {
                    uint64_t drop = 0;
                    for (uint64_t t = sf; t != 0; t &= t - 1) {
                        const int b = __ffsll((unsigned long long)t) - 1;
                        drop |= s_adj[adj_off(f * 64 + b, W) + (lane - f)];
                    }
                    U &= ~drop;
                }
            }
            if (lane < W)
                s_sel[lane] = sel;
        }
        __syncthreads();
        // the level's conflict record: far set, leaders, conflict bits
        const int W32 = (nf + 31) >> 5;
        if (CL.buf) {
            if (tid == 0) {
                const int64_t words = 4 + nf + 2 * W32 + 2 * (int64_t)nf * W32;
                const int64_t base =
                    (int64_t)atomicAdd((unsigned long long *)CL.buf, (unsigned long long)words);
                s_log = base + words <= CL.cap ? base : -1;
            }
            __syncthreads();
            if (s_log >= 0) {
                int32_t *e = CL.buf + 2 + s_log;
                if (tid == 0) {
                    e[0] = 1;
                    e[1] = CL.batch0 + (int32_t)S->bidx;
                    e[2] = level;
                    e[3] = nf;
                }
                for (int i = tid; i < nf; i += kBiResThreads)
                    e[4 + i] = s_far[i];
                for (int w = tid; w < W32; w += kBiResThreads) {
                    e[4 + nf + w] = (int32_t)(uint32_t)(s_sel[w >> 1] >> (32 * (w & 1)));
                    e[4 + nf + W32 + w] = 0; // covered bits, set below
                }
                int32_t *db = e + 4 + nf + 2 * W32, *gb = db + (int64_t)nf * W32;
                for (int it = tid; it < nf * W32; it += kBiResThreads) {
                    const int i = it / W32, w = it - i * W32;
                    uint32_t dw = 0u, gw = 0u;
                    if (nf > 1) {
#pragma unroll 4
                        for (int b = 0; b < 32; b++) {
                            const int j = 32 * w + b;
                            if (j > i && j < nf) {
                                if (dist[(size_t)i * kBiMaxChunk + j] <= r2)
                                    dw |= 1u << b;
                                if ((gadj[i * kBiWords + (j >> 6)] >> (j & 63)) & 1)
                                    gw |= 1u << b;
                            }
                        }
                    }
                    db[it] = (int32_t)dw;
                    gb[it] = (int32_t)gw;
                }
            }
        }
        const bool isl = tid < nf && ((s_sel[tid >> 6] >> (tid & 63)) & 1);
        int nl;
        const int pl = wg_prefix<kBiResWaves>(isl, L.cnt, &nl);
        if (isl) {
            s_lead[pl] = s_far[tid];
            s_lfar[pl] = tid;
            s_lpar[pl] = s_fcov[tid];
        }
        const bool isn = tid < nf && !isl;
        int nnl;
        const int pu = wg_prefix<kBiResWaves>(isn, L.cnt, &nnl);
        if (isn)
            s_nlf[pu] = tid;
        if (tid < nnl)
            s_nlkey[tid] = kBiFar;
        __syncthreads();
        link_in_order(t, rk, s_lead, s_lpar, nl, level, L);
        leaders = nl;
        const uint64_t tk1 = wall_clock64();
        tk_mis = tk1 - tk0;
        // non-leaders: nearest leader within r, strict <, first leader wins;
        // items are (non-leader, run of 32 leaders)
        const int nch = (nl + 31) >> 5;
        for (int it = tid; it < nnl * nch; it += kBiResThreads) {
            const int u = it / nch, k0 = (it - u * nch) * 32;
            const int k1 = k0 + 32 < nl ? k0 + 32 : nl;
            const int i = s_nlf[u];
            uint64_t best = kBiFar;
#pragma unroll 4
            for (int k = k0; k < k1; k++) {
                const int lf = s_lfar[k];
                const float d = i < lf ? dist[(size_t)i * kBiMaxChunk + lf]
                                       : dist[(size_t)lf * kBiMaxChunk + i];
                if (d <= r2) {
                    const uint64_t key = ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)k;
                    best = key < best ? key : best;
                }
            }
            if (best != kBiFar)
                atomicMin(&s_nlkey[u], (unsigned long long)best);
        }
        __syncthreads();
        const bool keep = tid < nnl && s_nlkey[tid] != kBiFar;
        if (CL.buf && s_log >= 0 && keep) {
            const int f = s_nlf[tid];
            atomicOr((unsigned int *)(CL.buf + 2 + s_log + 4 + nf + W32 + (f >> 5)),
                     1u << (f & 31));
        }
        int nk;
        const int pk = wg_prefix<kBiResWaves>(keep, L.cnt, &nk);
        if (keep) {
            s_nxc[nn + pk] = s_far[s_nlf[tid]];
            s_nxv[nn + pk] = s_lead[(int)(uint32_t)s_nlkey[tid]];
            s_nxst[nn + pk] = 0;
        }
        orphans = nnl - nk;
        ev_nl = (int64_t)nnl * nl;
        nn += nk;
        __syncthreads();
        tk_nl = wall_clock64() - tk1;
    }
    for (int i = tid; i < nn; i += kBiResThreads) {
        S->cand[i] = s_nxc[i];
        S->cover[i] = s_nxv[i];
        S->key[i] = kBiFar;
        S->stay[i] = s_nxst[i];
        S->d0[i] = s_nxd0[i];
        S->plen[i] = s_nxpl[i];
        S->phead[i] = s_nxph[i];
    }
    if (level == -10)
        link_in_order(t, rk, s_nxc, s_nxv, nn, level, L);
    if (level == -10 && CL.buf) {
        if (tid == 0) {
            const int64_t words = 3 + nn;
            const int64_t base =
                (int64_t)atomicAdd((unsigned long long *)CL.buf, (unsigned long long)words);
            s_log = base + words <= CL.cap ? base : -1;
        }
        __syncthreads();
        if (s_log >= 0) {
            int32_t *e = CL.buf + 2 + s_log;
            if (tid == 0) {
                e[0] = 2;
                e[1] = CL.batch0 + (int32_t)S->bidx;
                e[2] = nn;
            }
            for (int i = tid; i < nn; i += kBiResThreads)
                e[3 + i] = s_nxc[i];
        }
    }
    if (tid == 0) {
        S->nc = nn;
        S->nrk = L.nrk;
        S->ev_scan += s_acc[0];
        S->ev_mis += ev_mis;
        S->ev_nl += ev_nl;
        S->orphans += orphans;
        S->e_grid += s_acc[1];
        S->e_dist += s_acc[2];
        S->leaders += leaders;
        S->far_levels += nf > 0 ? 1 : 0;
        S->tk_scan += S->tk_pairs > S->tk_mark ? S->tk_pairs - S->tk_mark : 0;
        S->tk_conf += tk0 > S->tk_pairs ? tk0 - S->tk_pairs : 0;
        S->tk_mis += tk_mis;
        S->tk_nl += tk_nl;
        const uint64_t now = wall_clock64();
        S->tk_mark = now;
        if (level == -10 && bs) {
            bs += S->bidx * kBatchFields;
            bs[0] = S->s;
            bs[1] = S->nb;
            bs[2] = S->leaders;
            bs[3] = nn;
            bs[4] = S->orphans;
            bs[5] = S->far_levels;
            bs[6] = S->ev_scan;
            bs[7] = S->ev_mis;
            bs[8] = S->ev_nl;
            bs[9] = S->e_grid;
            bs[10] = S->e_dist;
            bs[11] = (int64_t)S->tk_scan;
            bs[12] = (int64_t)S->tk_mis;
            bs[13] = (int64_t)S->tk_nl;
            bs[14] = (int64_t)(now - S->tk_begin);
            bs[15] = (int64_t)S->tk_conf;
        }
        if (level == -10)
            S->bidx += 1;
    }
}

template <int KIND>
__global__ void __launch_bounds__(kBiResThreads)
    bi_resolve_kernel(TreeLinks t, BiState *S, int32_t *__restrict__ rk,
                      const float *__restrict__ dist, const uint64_t *__restrict__ adj,
                      const uint64_t *__restrict__ gadj, int32_t level, float r2,
                      int64_t *__restrict__ bs, int32_t track_from, ConflictLog CL)
{
    // levels at or below track_from count towards the work the fused
    // kernel would have had (thread 0 makes every update of these counters)
    const bool track = level <= track_from && threadIdx.x == 0;
    const int64_t ev0 = track ? S->ev_scan : 0;
    const int64_t evp0 = track ? S->ev_mis + S->ev_nl : 0;
    bi_resolve_body<KIND>(t, S, rk, dist, adj, gadj, level, r2, bs, CL);
    if (track) {
        S->ev_fused += S->ev_scan - ev0;
        S->ev_fused_pairs += S->ev_mis + S->ev_nl - evp0;
    }
}

// Levels top .. -10 of a batch in one workgroup: the scan, the conflict pass
// and the resolve of each level back to back, separated by workgroup
// barriers.  Launching three kernels per level costs more than the work of a
// level once few candidates remain (and for the Euclidean metric at every
// level).  The scan's atomic minima are made at the L2 cache, so after the
// scan the vector L1 is invalidated (an agent-scope acquire) before the keys
// are read; everything else stays within this workgroup's CU.  Levels with
// no candidates left are skipped, except level -10, whose resolve records
// the batch's counters.
template <int KIND>
__global__ void __launch_bounds__(kBiResThreads)
    bi_levels_kernel(MetricView mv, TreeLinks t, BiState *S, int32_t *__restrict__ rk,
                     const float *__restrict__ raw, int32_t gd, float gs, int32_t use_grid,
                     float *__restrict__ dist, uint64_t *__restrict__ adj,
                     uint64_t *__restrict__ gadj, int32_t top, int64_t *__restrict__ bs,
                     int64_t n_nodes, ConflictLog CL)
{
    __shared__ int32_t s_pfar[kBiMaxChunk];
    __shared__ int s_pcnt[kBiResWaves];
    // per wavefront the eight chunk-sum rows of one cooperative pass (the
    // resolve's arrays leave no room for 64)
    __shared__ float s_cs8[kBiResWaves * 8 * coop_row<8>()];
    const int wave = (int)(threadIdx.x >> 6);
    float *cs = s_cs8 + wave * 8 * coop_row<8>();
    // thread 0 updates ev_scan in every resolve, so its own read is current
    const int64_t ev0 = threadIdx.x == 0 ? S->ev_scan : 0;
    const int64_t evp0 = threadIdx.x == 0 ? S->ev_mis + S->ev_nl : 0;
    for (int32_t level = top; level >= -10; level--) {
        if (level > -10 && S->nc == 0)
            continue;
        const float r = ldexpf(1.0f, level), r2 = r * r;
        bi_scan_body<KIND, 8>(mv, t, S, rk, r2, n_nodes, wave, kBiResWaves, cs);
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        bi_pairs_body<KIND, kBiResWaves, 8>(mv, S, raw, gd, gs, use_grid, r2, dist, adj, gadj,
                                            s_pfar, s_pcnt, wave, kBiResWaves, true, cs);
        __syncthreads();
        bi_resolve_body<KIND>(t, S, rk, dist, adj, gadj, level, r2, bs, CL);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        S->ev_fused += S->ev_scan - ev0;
        S->ev_fused_pairs += S->ev_mis + S->ev_nl - evp0;
    }
}

// ---- Morton order ---------------------------------------------------------

constexpr int kMinmaxThreads = 1024;
__global__ void __launch_bounds__(kMinmaxThreads)
    minmax3_kernel(const float *__restrict__ x, int64_t n, float *out)
{
    // per-axis min/max; NaN never updates (matches `if v < mins[i]`).  The
    // threads read consecutive elements (element e is axis e % 3), eight
    // loads in flight per thread; min and max of the values taken do not
    // depend on the order they are seen in.
    __shared__ float smin[3][kMinmaxThreads], smax[3][kMinmaxThreads];
    float mn[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
    float mx[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    const int64_t total = n * 3;
    constexpr int kDepth = 8;
    for (int64_t base = threadIdx.x; base < total; base += kDepth * kMinmaxThreads) {
        float v[kDepth];
#pragma unroll
        for (int u = 0; u < kDepth; u++) {
            const int64_t e = base + (int64_t)u * kMinmaxThreads;
            v[u] = e < total ? x[e] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kDepth; u++) {
            const int64_t e = base + (int64_t)u * kMinmaxThreads;
            if (e >= total)
                break;
            const int c = (int)(e % 3);
#pragma unroll
            for (int a = 0; a < 3; a++)
                if (a == c) {
                    if (v[u] < mn[a])
                        mn[a] = v[u];
                    if (v[u] > mx[a])
                        mx[a] = v[u];
                }
        }
    }
    for (int a = 0; a < 3; a++) {
        smin[a][threadIdx.x] = mn[a];
        smax[a][threadIdx.x] = mx[a];
    }
    __syncthreads();
    for (int o = kMinmaxThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int a = 0; a < 3; a++) {
                const float lo = smin[a][threadIdx.x + o], hi = smax[a][threadIdx.x + o];
                if (lo < smin[a][threadIdx.x])
                    smin[a][threadIdx.x] = lo;
                if (hi > smax[a][threadIdx.x])
                    smax[a][threadIdx.x] = hi;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; a++) {
            out[a] = smin[a][0];
            out[3 + a] = smax[a][0];
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
// (sync: wait for the sort before returning; with stream-ordered work
// buffers, tls_alloc_stream == st, the caller may leave it queued)
static void morton_order_on_stream(const float *coords_dev, int64_t n, int32_t d,
                                   int64_t *order_dev, hipStream_t st, bool sync = true)
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
    minmax3_kernel<<<1, kMinmaxThreads, 0, st>>>(coords_dev, n, mm.ptr);
    morton_codes_kernel<<<blocks, threads, 0, st>>>(coords_dev, n, mm.ptr, codes.ptr, idx.ptr);
    CX_HIP(hipGetLastError());
    // radix sort is stable: ties keep index order, as slice::sort_by_key does
    size_t tmp_bytes = 0;
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, codes.ptr, codes_sorted.ptr,
                                              idx.ptr, order_dev, (int)n, 0, 64, st));
    DevBuf<uint8_t> tmp(tmp_bytes);
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, codes.ptr, codes_sorted.ptr,
                                              idx.ptr, order_dev, (int)n, 0, 64, st));
    if (sync || tls_alloc_stream != st)
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

// Conflict records of batch inserts (off by default): set_conflict_log
// turns them on for every later insert in the process.
static std::atomic<int> g_conflict_log{0};

void set_conflict_log(bool enable) { g_conflict_log.store(enable ? 1 : 0); }

// Words of conflict record one insert call may write (4-byte words; 32M by
// default, COVERTREEX_HIP_CONFLICT_LOG_WORDS overrides it).
static int64_t conflict_log_words()
{
    if (const char *e = getenv("COVERTREEX_HIP_CONFLICT_LOG_WORDS")) {
        const long long v = atoll(e);
        if (v > 0)
            return (int64_t)v;
    }
    return (int64_t)1 << 25;
}

// Runs the batch insert of rows [s_begin, s_end) (default: to t.n) and
// appends the per-batch telemetry to t.batch_stats.
// Runs of at least this many full batches replay a recorded graph.
constexpr int64_t kGraphMinBatches = 3;
// Batches per segment between the host's reads of the fused levels' work
// (the first segment is shorter and read at once).
constexpr int64_t kSegBatches = 8, kFirstSegBatches = 2;

// COVERTREEX_HIP_BUILD_GRAPHS=0 launches every batch directly.
static bool build_graphs_enabled()
{
    const char *e = getenv("COVERTREEX_HIP_BUILD_GRAPHS");
    return !(e && e[0] == '0');
}

// Highest level that a batch runs in one workgroup (bi_levels_kernel).  Its
// scan walks the candidates' child lists with 16 wavefronts instead of one
// wavefront per candidate, so it pays only where few candidates remain.
// Under the residual metric levels 9..0 run in the descent kernels and most
// far candidates (hundreds per batch) appear at levels -1 and -2, whose
// conflict passes of rank-m dot products need the whole GPU; deeper levels
// hold a few candidates each.  A Euclidean batch of clustered points keeps
// most candidates down to the cluster scale and settles them there (levels
// -4 .. -6 for clusters of width 0.02); a handful reach level -7.
// COVERTREEX_HIP_FUSED_LEVEL overrides the choice (any level in [-11, 9];
// -11 runs no level fused).
static int32_t fused_build_level(bool euclidean)
{
    if (const char *e = getenv("COVERTREEX_HIP_FUSED_LEVEL")) {
        const int v = atoi(e);
        if (v >= -11 && v <= 9)
            return v;
    }
    return euclidean ? -7 : -3;
}

// Work of the fused kernel per batch, in row elements read (evaluations of
// the scans, conflict pairs and non-leader checks, times the row length),
// above which the rest of an insert runs every level wide.
// COVERTREEX_HIP_FUSED_MAX_WORK overrides it (0: wide after the first
// segments whose fused levels did any work).
static int64_t fused_max_work()
{
    if (const char *e = getenv("COVERTREEX_HIP_FUSED_MAX_WORK")) {
        const long long v = atoll(e);
        if (v >= 0)
            return (int64_t)v;
    }
    return 65536;
}

static bool fused_too_slow(int64_t d_scan, int64_t d_pairs, int64_t d_batches, int32_t dim,
                           int64_t max_work)
{
    return (d_scan + d_pairs) * (int64_t)(dim + 8) > max_work * d_batches;
}

// Per-axis minimum and maximum of the rows [blockIdx.x * rows_per, ...) of
// pts (n x d): part[(c * 2 + 0) * gridDim.x + blockIdx.x] = min, +1 = max.
__global__ void __launch_bounds__(256)
    box_partial_kernel(const float *__restrict__ pts, int64_t n, int32_t d, int64_t rows_per,
                       float *__restrict__ part)
{
    __shared__ float s_lo[256], s_hi[256];
    const int64_t r0 = (int64_t)blockIdx.x * rows_per;
    const int64_t r1 = r0 + rows_per < n ? r0 + rows_per : n;
    for (int32_t c = 0; c < d; c++) {
        float lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
        for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
            const float v = pts[r * d + c];
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
        }
        s_lo[threadIdx.x] = lo;
        s_hi[threadIdx.x] = hi;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                s_lo[threadIdx.x] = fminf(s_lo[threadIdx.x], s_lo[threadIdx.x + o]);
                s_hi[threadIdx.x] = fmaxf(s_hi[threadIdx.x], s_hi[threadIdx.x + o]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            part[(2 * (int64_t)c) * gridDim.x + blockIdx.x] = s_lo[0];
            part[(2 * (int64_t)c + 1) * gridDim.x + blockIdx.x] = s_hi[0];
        }
        __syncthreads();
    }
}

// The FP32 fold over the axes of (hi - lo)^2 of the box of all points: no
// FP32 squared distance between two of the points exceeds it.
__global__ void box_diag_kernel(const float *__restrict__ part, int32_t nparts, int32_t d,
                                float *__restrict__ out)
{
    float acc = 0.0f;
    for (int32_t c = 0; c < d; c++) {
        float lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
        for (int32_t b = 0; b < nparts; b++) {
            lo = fminf(lo, part[(2 * (int64_t)c) * nparts + b]);
            hi = fmaxf(hi, part[(2 * (int64_t)c + 1) * nparts + b]);
        }
        const float g = hi - lo;
        acc = acc + g * g;
    }
    *out = acc;
}

// Lowest level in [-9, 9] whose r^2 = (2^level)^2 covers the FP32 squared
// diameter bound of the tree's points, with every level above it covering it
// too; 10 when level 9 does not.  Level -10 always runs the full level (its
// resolve links the bottom survivors).
static int32_t euclid_near_bottom(TreeState &t, hipStream_t st)
{
    const int64_t n = t.n;
    if (n <= 0)
        return 10;
    const int64_t rows_per = 4096;
    const int32_t nparts = (int32_t)((n + rows_per - 1) / rows_per);
    DevBuf<float> part((size_t)nparts * 2 * t.d + 1);
    box_partial_kernel<<<(unsigned)nparts, 256, 0, st>>>(t.raw.ptr, n, t.d, rows_per, part.ptr);
    box_diag_kernel<<<1, 1, 0, st>>>(part.ptr, nparts, t.d, part.ptr + (size_t)nparts * 2 * t.d);
    CX_HIP(hipGetLastError());
    float d2 = 0.0f;
    CX_HIP(hipMemcpyAsync(&d2, part.ptr + (size_t)nparts * 2 * t.d, sizeof(float),
                          hipMemcpyDeviceToHost, st));
    CX_HIP(hipStreamSynchronize(st));
    int32_t bottom = 10;
    for (int32_t level = 9; level >= -9; level--) {
        const float r = ldexpf(1.0f, level), r2 = r * r;
        if (!(d2 <= r2))
            break;
        bottom = level;
    }
    return bottom;
}

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
    CX_REQUIRE(chunk > 0 && chunk <= kBiMaxChunk, "chunk must be in [1, 1024]");
    TreeLinks links{t.parents.ptr, t.children.ptr, t.next.ptr, t.levels.ptr};
    DevBuf<BiState> state(1);
    DevBuf<int32_t> rk(t.n);
    DevBuf<float> dist((size_t)kBiMaxChunk * kBiMaxChunk);
    DevBuf<uint64_t> adj((size_t)kBiMaxChunk * kBiWords), gadj((size_t)kBiMaxChunk * kBiWords);
    if (s_begin > 0)
        bi_root_kids_kernel<<<1, 1, 0, st>>>(links, state.ptr, rk.ptr, t.n);
    // the conflict record of this run's batches (when turned on)
    ConflictLog clog{nullptr, 0, (int32_t)(t.batch_stats.size() / kBatchFields)};
    DevBuf<int32_t> logbuf;
    if (g_conflict_log.load() != 0) {
        clog.cap = conflict_log_words();
        logbuf.alloc((size_t)clog.cap + 2);
        clog.buf = logbuf.ptr;
        CX_HIP(hipMemsetAsync(logbuf.ptr, 0, 2 * sizeof(int32_t), st));
    }
    const int32_t fused_from = fused_build_level(KIND == kMetricEuclidean);
    const int32_t near_bottom = KIND == kMetricEuclidean ? euclid_near_bottom(t, st) : 10;
    CX_HIP(hipMemsetAsync(&state.ptr->bidx, 0, sizeof(int64_t), st));
    CX_HIP(hipMemsetAsync(&state.ptr->ev_fused, 0, 2 * sizeof(int64_t), st));
    // one batch of bn rows: every launch of levels 9 .. -10; with `fused`
    // false every level below the descent runs as three launches
    auto launch_batch = [&](int64_t bn, bool fused_mode) {
        bi_begin_kernel<<<1, kBiResThreads, 0, st>>>(links, state.ptr, s_begin, chunk, bn);
        int32_t top = 9;
        if constexpr (KIND == kMetricResidual) {
            const float r9 = ldexpf(1.0f, 9), r9sq = r9 * r9;
            bi_rootscan_kernel<<<dim3(128, (unsigned)((bn + kRtTile - 1) / kRtTile)), 256, 0,
                                 st>>>(mv, state.ptr, rk.ptr, r9sq);
            bi_descend_kernel<KIND><<<(unsigned)bn, kBiParThreads, 0, st>>>(
                mv, links, state.ptr, rk.ptr, t.n, 0);
            top = -1;
        } else if (near_bottom <= 9) {
            bi_descend_kernel<KIND><<<(unsigned)bn, kBiParThreads, 0, st>>>(
                mv, links, state.ptr, rk.ptr, t.n, near_bottom);
            top = near_bottom - 1;
        }
        // levels above `fused` run as three launches each (all CUs), the
        // rest in one workgroup
        const int32_t fused = !fused_mode ? -11 : fused_from < top ? fused_from : top;
        for (int32_t level = top; level > fused; level--) {
            const float r = ldexpf(1.0f, level), r2 = r * r;
            bi_scan_kernel<KIND>
                <<<kBiParBlocks, kBiParThreads, 0, st>>>(mv, links, state.ptr, rk.ptr, r2, t.n);
            bi_pairs_kernel<KIND><<<kBiParBlocks, kBiParThreads, 0, st>>>(
                mv, state.ptr, t.raw.ptr, t.d, grid_scale, use_grid, r2, dist.ptr, adj.ptr,
                gadj.ptr);
            bi_resolve_kernel<KIND><<<1, kBiResThreads, 0, st>>>(
                links, state.ptr, rk.ptr, dist.ptr, adj.ptr, gadj.ptr, level, r2, bst.ptr,
                fused_from, clog);
        }
        if (fused >= -10)
            bi_levels_kernel<KIND><<<1, kBiResThreads, 0, st>>>(
                mv, links, state.ptr, rk.ptr, t.raw.ptr, t.d, grid_scale, use_grid, dist.ptr,
                adj.ptr, gadj.ptr, fused, bst.ptr, t.n, clog);
    };
    // Full batches replay one recorded graph of the batch's launches per
    // mode (the host then submits a batch at a time instead of tens of
    // kernels); the graphs hold kernel nodes only.  The last, partial batch
    // is launched directly.
    const int64_t n_full = (n - s_begin) / chunk;
    const bool graphs = n_full >= kGraphMinBatches && build_graphs_enabled();
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    auto drop_graphs = [&]() {
        for (hipGraphExec_t &e : exec)
            if (e) {
                (void)hipGraphExecDestroy(e);
                e = nullptr;
            }
    };
    auto launch_range = [&](int64_t b0, int64_t b1, bool fused_mode) {
        for (int64_t b = b0; b < b1; b++) {
            const int64_t r0 = s_begin + b * chunk;
            if (!graphs || b >= n_full) {
                launch_batch(n - r0 < chunk ? n - r0 : chunk, fused_mode);
                CX_HIP(hipGetLastError());
                continue;
            }
            hipGraphExec_t &e = exec[fused_mode ? 1 : 0];
            if (!e) {
                hipGraph_t graph = nullptr;
                CX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                launch_batch(chunk, fused_mode);
                const hipError_t launch_err = hipGetLastError();
                const hipError_t cap_err = hipStreamEndCapture(st, &graph);
                CX_HIP(launch_err);
                CX_HIP(cap_err);
                const hipError_t inst_err = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                CX_HIP(inst_err);
            }
            CX_HIP(hipGraphLaunch(e, st));
        }
    };
    // The fused levels pay while the candidates that reach them scan short
    // lists.  Where they stay under covers with long child lists (smooth
    // fields: correlations at rounding level keep most candidates at the
    // root, whose list grows to tens of thousands), one workgroup scanning
    // them is many times slower than the wide levels.  Batches therefore
    // run in segments, each in the mode the work measured so far asks for:
    // both modes count the work of the levels at or below fused_from, and
    // after each segment the host reads it (one segment behind, so the
    // device stays busy; the first segment of two batches is read at once).
    // The tree is the same in either mode.
    int64_t b = 0;
    bool fused_mode = fused_from >= -10;
    try {
        if (!fused_mode) {
            launch_range(0, nb, false);
        } else {
            const int64_t max_work = fused_max_work();
            const int32_t dim = mv.m > 0 ? mv.m : mv.d;
            int64_t *word[2] = {(int64_t *)acquire_host_word(), (int64_t *)acquire_host_word()};
            hipEvent_t ev[2] = {acquire_event(), acquire_event()};
            // A pooled event stays tied to the stream it was last recorded
            // on, which may by now be another thread's, under graph capture:
            // waiting on it there invalidates that capture.  Only events
            // recorded here are waited on.
            bool recorded[2] = {false, false};
            int64_t seg_end[2] = {-1, -1};
            int64_t seen_scan = 0, seen_pairs = 0, seen_b = 0;
            int slot = 0;
            auto give_back = [&]() {
                for (int i = 0; i < 2; i++) {
                    if (recorded[i])
                        (void)hipEventSynchronize(ev[i]);
                    release_event(ev[i]);
                    release_host_word((int32_t *)word[i]);
                }
            };
            auto decide = [&](int s) {
                CX_HIP(hipEventSynchronize(ev[s]));
                const int64_t d_scan = word[s][0] - seen_scan, d_pairs = word[s][1] - seen_pairs;
                const int64_t d_b = seg_end[s] - seen_b;
                seen_scan = word[s][0];
                seen_pairs = word[s][1];
                seen_b = seg_end[s];
                seg_end[s] = -1;
                fused_mode = !fused_too_slow(d_scan, d_pairs, d_b, dim, max_work);
            };
            try {
                while (b < nb) {
                    const int64_t len = b == 0 ? kFirstSegBatches : kSegBatches;
                    const int64_t e = b + len < nb ? b + len : nb;
                    launch_range(b, e, fused_mode);
                    b = e;
                    if (b >= nb)
                        break;
                    CX_HIP(hipMemcpyAsync(word[slot], &state.ptr->ev_fused, 2 * sizeof(int64_t),
                                          hipMemcpyDeviceToHost, st));
                    CX_HIP(hipEventRecord(ev[slot], st));
                    recorded[slot] = true;
                    seg_end[slot] = b;
                    if (b == len) {
                        decide(slot);
                    } else if (seg_end[slot ^ 1] >= 0) {
                        decide(slot ^ 1); // the segment before: finished or about to
                    }
                    slot ^= 1;
                }
            } catch (...) {
                give_back();
                throw;
            }
            give_back();
        }
    } catch (...) {
        (void)hipStreamSynchronize(st);
        drop_graphs();
        throw;
    }
    CX_HIP(hipStreamSynchronize(st)); // the graphs' last replays have finished
    drop_graphs();
    for (b = 0; b < nb; b++) {
        const int64_t b0 = s_begin + b * chunk;
        t.live_batches.push_back(b0);
        t.live_batches.push_back(n - b0 < chunk ? n - b0 : chunk);
    }
    const size_t old = t.batch_stats.size();
    t.batch_stats.resize(old + (size_t)nb * kBatchFields);
    CX_HIP(hipMemcpyAsync(t.batch_stats.data() + old, bst.ptr, bst.bytes(), hipMemcpyDeviceToHost,
                          st));
    CX_HIP(hipStreamSynchronize(st));
    if (clog.buf) {
        int64_t used = 0;
        CX_HIP(hipMemcpyAsync(&used, logbuf.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        CX_HIP(hipStreamSynchronize(st));
        CX_REQUIRE(used <= clog.cap, "conflict record overflow: the insert needs " +
                                         std::to_string(used) +
                                         " words; raise COVERTREEX_HIP_CONFLICT_LOG_WORDS");
        const size_t at = t.conflict_log.size();
        t.conflict_log.resize(at + (size_t)used);
        if (used)
            CX_HIP(hipMemcpyAsync(t.conflict_log.data() + at, logbuf.ptr + 2,
                                  (size_t)used * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        CX_HIP(hipStreamSynchronize(st));
    }
}

// Arrays derived from the links after inserts: child CSR, depths, si_cache
// (bottom-up one depth at a time), subtree index bounds and the roots in
// index order.
template <int KIND> static void finalize_tree(TreeState &t, const MetricView &mv, hipStream_t st)
{
    const int64_t n = t.n;
    t.eu_ready = false; // the exact Euclidean blocks follow the point set
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
// Rows [n0, n0 + nn) of the tree arrays from device staging buffers of V,
// coords and p (nn rows each, ordered on st): tree row n0 + i takes staged
// row sel[i]; coordinates are also scaled by 1/lengthscale.
static void gather_staged_rows(TreeState &t, const float *v_in, const float *c_in,
                               const float *p_in, int64_t n0, int64_t nn, const int64_t *sel,
                               hipStream_t st)
{
    const int32_t m = t.m, d = t.d;
    gather_rows_kernel<<<(unsigned)nn, 64, 0, st>>>(v_in, sel, nn, m, t.v.ptr + n0 * m);
    gather_rows_kernel<<<(unsigned)nn, 64, 0, st>>>(c_in, sel, nn, d, t.raw.ptr + n0 * d);
    gather_rows_kernel<<<(unsigned)nn, 64, 0, st>>>(p_in, sel, nn, 1, t.p.ptr + n0);
    DevBuf<float> inv_ls(d);
    CX_HIP(hipMemcpyAsync(inv_ls.ptr, t.inv_ls.data(), sizeof(float) * d, hipMemcpyHostToDevice,
                          st));
    scale_rows_kernel<<<nblocks(nn, 256), 256, 0, st>>>(t.raw.ptr + n0 * d, nn, d, inv_ls.ptr,
                                                        t.scaled.ptr + n0 * d, t.norms.ptr + n0);
    CX_HIP(hipGetLastError());
    CX_HIP(hipStreamSynchronize(st)); // the staging buffers go out of scope
}

// The same rows from host V, p, coords (nn rows).
static void load_residual_rows(TreeState &t, const float *V, const float *p, const float *coords,
                               int64_t n0, int64_t nn, const int64_t *sel, hipStream_t st)
{
    const int32_t m = t.m, d = t.d;
    DevBuf<float> v_in((size_t)nn * m), c_in((size_t)nn * d), p_in(nn);
    CX_HIP(hipMemcpyAsync(v_in.ptr, V, sizeof(float) * nn * m, hipMemcpyHostToDevice, st));
    CX_HIP(hipMemcpyAsync(c_in.ptr, coords, sizeof(float) * nn * d, hipMemcpyHostToDevice, st));
    CX_HIP(hipMemcpyAsync(p_in.ptr, p, sizeof(float) * nn, hipMemcpyHostToDevice, st));
    gather_staged_rows(t, v_in.ptr, c_in.ptr, p_in.ptr, n0, nn, sel, st);
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
    t.v.alloc((size_t)n * m);
    t.raw.alloc((size_t)n * d);
    t.scaled.alloc((size_t)n * d);
    t.norms.alloc(n);
    t.p.alloc(n);
    if (order_mode == CX_ORDER_NATURAL || d != 3) {
        iota_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.order.ptr, n);
        CX_HIP(hipGetLastError());
        inverse_perm_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.order.ptr, n, t.inv_order.ptr);
        load_residual_rows(t, V, p, coords, 0, n, t.order.ptr, st);
    } else {
        // The staged V and p go up on a second stream while the Morton
        // order is computed: the host copies from pageable memory (blocking)
        // only after the sort is queued, and the second stream waits only
        // for the staging buffers' allocation.
        DevBuf<float> v_in((size_t)n * m), c_in((size_t)n * d), p_in(n);
        struct Upload {
            int dev;
            hipStream_t s = nullptr;
            hipEvent_t alloc_done = nullptr, copied = nullptr;
            ~Upload()
            {
                if (s)
                    release_stream(dev, s); // synchronizes it
                if (alloc_done)
                    release_event(alloc_done);
                if (copied)
                    release_event(copied);
            }
        } up{t.device};
        up.s = acquire_stream(t.device);
        up.alloc_done = acquire_event();
        up.copied = acquire_event();
        CX_HIP(hipEventRecord(up.alloc_done, st));
        CX_HIP(hipMemcpyAsync(c_in.ptr, coords, sizeof(float) * n * d, hipMemcpyHostToDevice,
                              st));
        morton_order_on_stream(c_in.ptr, n, d, t.order.ptr, st, false);
        inverse_perm_kernel<<<nblocks(n, 256), 256, 0, st>>>(t.order.ptr, n, t.inv_order.ptr);
        CX_HIP(hipGetLastError());
        CX_HIP(hipStreamWaitEvent(up.s, up.alloc_done, 0));
        CX_HIP(hipMemcpyAsync(v_in.ptr, V, sizeof(float) * n * m, hipMemcpyHostToDevice, up.s));
        CX_HIP(hipMemcpyAsync(p_in.ptr, p, sizeof(float) * n, hipMemcpyHostToDevice, up.s));
        CX_HIP(hipEventRecord(up.copied, up.s));
        CX_HIP(hipStreamWaitEvent(st, up.copied, 0));
        gather_staged_rows(t, v_in.ptr, c_in.ptr, p_in.ptr, 0, n, t.order.ptr, st);
    }

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

// expf_cr of the floats whose bit patterns follow first_bits (the exp of
// the build and the reference-form metric, exposed for its exhaustive check)
__global__ void expf_cr_range_kernel(uint32_t first_bits, int64_t count, float *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        out[i] = expf_cr(__builtin_bit_cast(float, first_bits + (uint32_t)i));
}

void expf_cr_range(uint32_t first_bits, int64_t count, float *out)
{
    DevBuf<float> d((size_t)count);
    expf_cr_range_kernel<<<nblocks(count, 256), 256>>>(first_bits, count, d.ptr);
    CX_HIP(hipGetLastError());
    CX_HIP(hipMemcpy(out, d.ptr, sizeof(float) * count, hipMemcpyDeviceToHost));
}

} // namespace cx
