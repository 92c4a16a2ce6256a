// Reference-form residual kNN: the reference's search semantics with its f32
// arithmetic, one wavefront per query (CX_KNN_REFERENCE_F32).
//
// Restates single_residual_knn_query (src/algo.rs:612-1242) with an empty
// budget schedule (exact search), stream tile 64, dynamic tile sizes
// (1001-1022), radius floor 1e-3 and level cap 2.0:
//   - every root (parent == -1) is scored and seeded; valid roots enter the
//     result heap without pops, kth = smallest root distance (688-734);
//   - each level gathers the children of the frontier, pruning a parent when
//     the heap holds exactly k and d(parent) - r(parent) > kth (862-995);
//   - children are scored in tiles, sorted (stable) by distance and pushed in
//     order with the cover-radius test (997-1189).
// The result heap is kept as its k smallest entries plus the reference's heap
// length (which the root seeding can take above k, disabling every prune);
// ties at equal f32 distance may resolve differently from the reference's
// BinaryHeap.  Distances: roots and predecessor-failed children use
// sqrt(distance_sq_idx_f32), tiles use distances_f32_batch_idx_into_with_kth,
// exactly as the reference does.
#include "cx_internal.hpp"
#include "device_common.hpp"

namespace cx {

constexpr int kRefMaxK = 256;

// Budget ladder (src/algo.rs:648-683, 1191-1219); n == 0 disables it.
struct Budget {
    int32_t n;
    int64_t sched[4];
    double up, down;
};

struct RefTree {
    const int64_t *__restrict__ order;     // node -> dataset
    const int64_t *__restrict__ inv_order; // dataset -> node
    const int32_t *__restrict__ roots;
    int64_t n_roots;
    const int32_t *__restrict__ child_off;
    const int32_t *__restrict__ child_list;
    const float *__restrict__ si;
    const int64_t *__restrict__ smin;
    int64_t n;
};

__device__ __forceinline__ float radius_of(const RefTree &t, int64_t node)
{
    float r = t.si[node];
    if (2.0f < r) // apply_level_cap with cap_default 2.0 (src/metric.rs:202-213)
        r = 2.0f;
    if (r < 1e-3f) // COVERTREEX_RESIDUAL_RADIUS_FLOOR default (src/algo.rs:814-818)
        r = 1e-3f;
    return r;
}

__device__ __forceinline__ bool has_children(const RefTree &t, int64_t node)
{
    return t.child_off[node + 1] > t.child_off[node];
}

// k smallest (distance, node) pairs in LDS with a tracked maximum.
struct TopK {
    float *d;
    int32_t *nd;
    int count;
    int k;
    float maxv;
    int maxpos;
};

__device__ __forceinline__ void topk_refresh_max(TopK &tk)
{
    const int lane = threadIdx.x;
    float best = -__builtin_huge_valf();
    int bpos = 0x7fffffff;
    for (int s = lane; s < tk.count; s += 64) {
        float v = tk.d[s];
        if (v > best || (v == best && s < bpos)) {
            best = v;
            bpos = s;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        float ov = __shfl_xor(best, o, 64);
        int op = __shfl_xor(bpos, o, 64);
        if (ov > best || (ov == best && op < bpos)) {
            best = ov;
            bpos = op;
        }
    }
    tk.maxv = best;
    tk.maxpos = bpos;
}

// Wave-uniform insertion of (x, node).
__device__ __forceinline__ void topk_insert(TopK &tk, float x, int32_t node)
{
    const int lane = threadIdx.x;
    if (tk.count < tk.k) {
        if (lane == 0) {
            tk.d[tk.count] = x;
            tk.nd[tk.count] = node;
        }
        tk.count++;
        __builtin_amdgcn_wave_barrier();
        __syncthreads();
        topk_refresh_max(tk);
    } else if (x < tk.maxv) {
        if (lane == 0) {
            tk.d[tk.maxpos] = x;
            tk.nd[tk.maxpos] = node;
        }
        __syncthreads();
        topk_refresh_max(tk);
    }
}

__global__ void __launch_bounds__(64)
    knn_ref_kernel(MetricView mv, RefTree tr, const int64_t *__restrict__ queries, int64_t nq,
                   int32_t k, int32_t pred, int64_t *out_idx, float *out_dist, float *out_err,
                   int32_t *counter, int32_t *ws_node, float *ws_val, int64_t fr_cap,
                   int64_t ch_cap, Budget bud, int64_t *telem)
{
    __shared__ float s_td[kRefMaxK];
    __shared__ int32_t s_tn[kRefMaxK];
    __shared__ int s_q;
    const int lane = threadIdx.x;
    const int64_t per_wave = 2 * fr_cap + ch_cap;
    int32_t *fa_n = ws_node + (int64_t)blockIdx.x * per_wave;
    float *fa_d = ws_val + (int64_t)blockIdx.x * per_wave;
    int32_t *fb_n = fa_n + fr_cap;
    float *fb_d = fa_d + fr_cap;
    int32_t *ch_n = fb_n + fr_cap;
    float *ch_l = fb_d + fr_cap;
    const int64_t n = tr.n;

    for (;;) {
        if (lane == 0)
            s_q = atomicAdd(counter, 1);
        __syncthreads();
        const int64_t qi = s_q;
        __syncthreads();
        if (qi >= nq)
            break;
        const int64_t q_ds = queries[qi];
        const int64_t q = (q_ds >= 0 && q_ds < n) ? tr.inv_order[q_ds] : 0;
        // predecessor bound: the original dataset index as usize (src/lib.rs:444-448),
        // so a negative index admits every node
        const int64_t max_idx = q_ds < 0 ? INT64_MAX : q_ds;
        TopK tk{s_td, s_tn, 0, k, -__builtin_huge_valf(), 0};
        for (int j = lane; j < k; j += 64) {
            out_idx[qi * k + j] = -1;
            out_dist[qi * k + j] = 3.40282347e38f;
            if (out_err)
                out_err[qi * k + j] = __builtin_nanf("");
        }
        if (n == 0 || (pred && q_ds == 0))
            continue;

        // telemetry counters (wave-uniform), added to telem at the end
        int64_t t_evals = 0, t_lb = 0, t_lbc = 0, t_pred = 0, t_sub = 0, t_esc = 0, t_early = 0;
        int32_t t_level = 0;

        // ---- roots (688-734) ------------------------------------------------
        int32_t budget_idx = 0;
        int64_t budget_limit = bud.n > 0 ? bud.sched[0] : INT64_MAX;
        int64_t survivors = 0;
        int32_t low_yield = 0;
        int64_t heap_len = 0;
        float kth = 3.40282347e38f;
        bool all_prunable = true;
        int64_t fr_count = 0, fr_stored = 0;
        for (int64_t base = 0; base < tr.n_roots; base += 64) {
            const int64_t ri = base + lane;
            const bool live = ri < tr.n_roots;
            int32_t r = live ? tr.roots[ri] : 0;
            bool valid = false;
            float rd = 0.0f;
            bool store = false;
            if (live) {
                valid = pred ? tr.order[r] < max_idx : true;
                rd = __builtin_sqrtf(dist_sq_ref(mv, q, r));
                store = has_children(tr, r);
                if (!(pred && tr.smin[r] >= max_idx))
                    all_prunable = false;
            }
            uint64_t smask = __ballot(store);
            int spos = __popcll(smask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
            if (store) {
                fa_n[fr_stored + spos] = r;
                fa_d[fr_stored + spos] = rd;
            }
            fr_stored += __popcll(smask);
            fr_count += __popcll(__ballot(live));
            t_evals += __popcll(__ballot(live));
            t_pred += __popcll(__ballot(live && !valid));
            uint64_t vmask = __ballot(valid);
            heap_len += __popcll(vmask);
            survivors += __popcll(vmask);
            float vmin = valid ? rd : 3.40282347e38f;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                float u = __shfl_xor(vmin, o, 64);
                vmin = u < vmin ? u : vmin;
            }
            if (vmin < kth)
                kth = vmin;
            // candidates that can change the k smallest, in index order
            uint64_t imask = __ballot(valid && (tk.count < k || rd < tk.maxv));
            while (imask) {
                int l = __ffsll((long long)imask) - 1;
                imask &= imask - 1;
                float x = __shfl(rd, l, 64);
                int32_t nd = __shfl(r, l, 64);
                topk_insert(tk, x, nd);
            }
        }
        all_prunable = __ballot(!all_prunable) == 0;
        bool early = pred && all_prunable; // 737-749

        // ---- level loop (828-1226) -------------------------------------------
        while (!early && fr_count > 0 &&
               (survivors < budget_limit || (pred && heap_len < k))) {
            int64_t ch_count = 0, nfr_count = 0, nfr_stored = 0;
            if (telem && lane == 0 && t_level < CX_QT_LEVELS)
                atomicAdd((unsigned long long *)&telem[CX_QT_FRONTIER + t_level],
                          (unsigned long long)fr_count);
            // gather (862-995)
            for (int64_t f = 0; f < fr_stored; f++) {
                const int32_t parent = fa_n[f];
                const float pdist = fa_d[f];
                const float plb = pdist - radius_of(tr, parent);
                if (heap_len == k && plb > kth) {
                    t_lb++;
                    continue;
                }
                const int32_t c0 = tr.child_off[parent], c1 = tr.child_off[parent + 1];
                for (int32_t cb = c0; cb < c1; cb += 64) {
                    const int32_t ci = cb + lane;
                    const bool live = ci < c1;
                    int32_t ch = live ? tr.child_list[ci] : 0;
                    bool skip = !live || (pred && tr.smin[ch] >= max_idx);
                    bool ok = !skip && (pred ? tr.order[ch] < max_idx : true);
                    bool failed = !skip && !ok; // subtree still has valid nodes
                    float dd = 0.0f;
                    if (failed)
                        dd = __builtin_sqrtf(dist_sq_ref(mv, q, ch));
                    bool fstore = failed && has_children(tr, ch);
                    uint64_t fm = __ballot(fstore);
                    int fp = __popcll(fm & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
                    if (fstore) {
                        fb_n[nfr_stored + fp] = ch;
                        fb_d[nfr_stored + fp] = dd;
                    }
                    nfr_stored += __popcll(fm);
                    nfr_count += __popcll(__ballot(failed));
                    t_evals += __popcll(__ballot(failed));
                    t_pred += __popcll(__ballot(failed));
                    t_sub += __popcll(__ballot(live && pred && tr.smin[ch] >= max_idx));
                    uint64_t om = __ballot(ok);
                    int op = __popcll(om & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
                    if (ok) {
                        ch_n[ch_count + op] = ch;
                        ch_l[ch_count + op] = plb;
                    }
                    ch_count += __popcll(om);
                }
            }
            if (telem && lane == 0 && t_level < CX_QT_LEVELS)
                atomicAdd((unsigned long long *)&telem[CX_QT_EXPANDED + t_level],
                          (unsigned long long)ch_count);
            t_level++;
            // tiles (997-1222)
            int64_t start = 0;
            while (start < ch_count && (survivors < budget_limit || (pred && heap_len < k))) {
                const int64_t active = ch_count - start;
                const int64_t wave = fr_count > 1 ? fr_count : 1;
                const int64_t comb = active + wave;
                int64_t tile = comb <= 32 ? 16 : comb <= 128 ? 32 : comb <= 512 ? 64 : 128;
                tile = tile < 64 ? tile : 64;
                const int64_t end = start + tile < ch_count ? start + tile : ch_count;
                const int len = (int)(end - start);
                const bool live = lane < len;
                int32_t node = live ? ch_n[start + lane] : 0;
                float lb = live ? ch_l[start + lane] : 3.40282347e38f;
                if (heap_len == k) { // chunk prune (1032-1045)
                    float mn = lb;
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) {
                        float u = __shfl_xor(mn, o, 64);
                        mn = u < mn ? u : mn;
                    }
                    if (mn > kth) {
                        t_lbc++;
                        start = end;
                        continue;
                    }
                }
                if (telem && lane == 0)
                    atomicAdd((unsigned long long *)&telem[CX_QT_BLOCKS +
                                                           (tile == 16 ? 0 : tile == 32 ? 1 : 2)],
                              1ull);
                bool eval = live && !(heap_len == k && lb > kth); // cached_lb prune (1053-1066)
                float dist = eval ? dist_batch_ref(mv, q, node) : __builtin_huge_valf();
                int pos = lane;
                const int n_ev = __popcll(__ballot(eval));
                t_evals += n_ev;
                t_lb += __popcll(__ballot(live && !eval));
                // stable sort by distance (1103-1109): bitonic on (dist, position)
                for (int size = 2; size <= 64; size <<= 1) {
                    for (int j = size >> 1; j > 0; j >>= 1) {
                        float od = __shfl_xor(dist, j, 64);
                        int opos = __shfl_xor(pos, j, 64);
                        int32_t onode = __shfl_xor(node, j, 64);
                        float olb = __shfl_xor(lb, j, 64);
                        const bool up = (lane & size) == 0;
                        const bool lower = (lane & j) == 0;
                        const bool other_less = od < dist || (od == dist && opos < pos);
                        const bool take = up ? (lower == other_less) : (lower != other_less);
                        if (take) {
                            dist = od;
                            pos = opos;
                            node = onode;
                            lb = olb;
                        }
                    }
                }
                int64_t added = 0;
                for (int e = 0; e < n_ev; e++) { // 1111-1189
                    const float dv = __shfl(dist, e, 64);
                    const int32_t child = __shfl(node, e, 64);
                    const float plbc = __shfl(lb, e, 64);
                    const float crad = radius_of(tr, child);
                    if (heap_len == k) {
                        float lbv = fminf(plbc, dv - crad);
                        if (lbv > kth) {
                            t_lb++;
                            continue;
                        }
                    }
                    topk_insert(tk, dv, child);
                    if (heap_len < k)
                        heap_len++;
                    kth = tk.maxv; // heap peek after push (+pop)
                    survivors++;
                    added++;
                    if (!(survivors < budget_limit || (pred && heap_len < k)))
                        break;
                    bool sub_valid = pred ? tr.smin[child] < max_idx : true;
                    if (sub_valid) {
                        if (has_children(tr, child)) {
                            if (lane == 0) {
                                fb_n[nfr_stored] = child;
                                fb_d[nfr_stored] = dv;
                            }
                            nfr_stored++;
                        }
                        nfr_count++;
                    }
                }
                bool stop = false;
                if (bud.n > 0 && len > 0) { // budget ladder (1191-1219)
                    const double ratio = (double)added / (double)len;
                    if (ratio >= bud.up) {
                        if (budget_idx + 1 < bud.n) {
                            budget_idx++;
                            budget_limit = bud.sched[budget_idx];
                            t_esc++;
                        }
                        low_yield = 0;
                    } else if (ratio < bud.down) {
                        low_yield++;
                        const bool need_more = pred && heap_len < k;
                        if (low_yield >= 2 && !need_more) {
                            stop = true;
                            t_early++;
                        }
                    } else {
                        low_yield = 0;
                    }
                }
                if (stop)
                    break;
                start = end;
            }
            int32_t *tn = fa_n;
            float *td = fa_d;
            fa_n = fb_n;
            fa_d = fb_d;
            fb_n = tn;
            fb_d = td;
            fr_count = nfr_count;
            fr_stored = nfr_stored;
            __syncthreads();
        }

        if (telem && lane == 0) {
            atomicAdd((unsigned long long *)&telem[CX_QT_DISTANCE_EVALS], (unsigned long long)t_evals);
            atomicAdd((unsigned long long *)&telem[CX_QT_PRUNES_LB], (unsigned long long)t_lb);
            atomicAdd((unsigned long long *)&telem[CX_QT_PRUNES_LB_CHUNKS],
                      (unsigned long long)t_lbc);
            atomicAdd((unsigned long long *)&telem[CX_QT_PRED_FILTERED], (unsigned long long)t_pred);
            atomicAdd((unsigned long long *)&telem[CX_QT_SUBTREES_PRUNED],
                      (unsigned long long)t_sub);
            atomicAdd((unsigned long long *)&telem[CX_QT_BUDGET_ESC], (unsigned long long)t_esc);
            atomicAdd((unsigned long long *)&telem[CX_QT_BUDGET_EARLY],
                      (unsigned long long)t_early);
        }

        // ---- output: ascending (into_sorted_vec, 1232-1240), truncated to k
        __syncthreads();
        for (int s = lane; s < tk.count; s += 64) {
            const float dv = s_td[s];
            const int32_t nd = s_tn[s];
            int rank = 0;
            for (int u = 0; u < tk.count; u++) {
                const float du = s_td[u];
                rank += (du < dv) || (du == dv && s_tn[u] < nd);
            }
            out_idx[qi * k + rank] = tr.order[nd];
            out_dist[qi * k + rank] = dv;
        }
        __syncthreads();
    }
}

void knn_reference_f32(TreeState &t, const int64_t *d_queries, int64_t nq, int32_t k,
                       int32_t predecessor, int32_t budgeted, int64_t *d_idx, float *d_dist,
                       float *d_err)
{
    CX_REQUIRE(k >= 1 && k <= kRefMaxK, "reference-form kNN supports 1 <= k <= 256");
    if (nq <= 0)
        return;
    hipStream_t st = t.stream;
    const int64_t waves = nq < 2048 ? nq : 2048;
    const int64_t fr_cap = t.n_internal + 1, ch_cap = t.n + 1;
    const int64_t per_wave = 2 * fr_cap + ch_cap;
    const size_t bytes = (size_t)waves * per_wave * (sizeof(int32_t) + sizeof(float)) + 1024;
    uint8_t *ws = workspace(t, bytes);
    int32_t *counter = (int32_t *)ws;
    int32_t *ws_node = (int32_t *)(ws + 256);
    float *ws_val = (float *)(ws + 256 + (size_t)waves * per_wave * sizeof(int32_t));
    CX_HIP(hipMemsetAsync(counter, 0, sizeof(int32_t), st));
    prof_mark(t, 0);
    MetricView mv{t.scaled.ptr, t.norms.ptr, t.v.ptr, t.p.ptr, t.d, t.m, t.var, t.kernel};
    RefTree tr{t.order.ptr, t.inv_order.ptr,  t.roots.ptr, t.n_roots, t.child_off.ptr,
               t.child_list.ptr, t.si.ptr, t.smin.ptr, t.n};
    Budget bud{};
    if (budgeted) { // reference defaults (src/algo.rs:653-676)
        bud.n = 3;
        bud.sched[0] = 32;
        bud.sched[1] = 64;
        bud.sched[2] = 96;
        bud.up = 0.015;
        bud.down = 0.002;
    }
    knn_ref_kernel<<<(unsigned)waves, 64, 0, st>>>(mv, tr, d_queries, nq, k, predecessor, d_idx,
                                                   d_dist, d_err, counter, ws_node, ws_val,
                                                   fr_cap, ch_cap, bud, t.telem);
    CX_HIP(hipGetLastError());
    prof_mark(t, 1);
}

} // namespace cx
