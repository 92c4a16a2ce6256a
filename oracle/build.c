/* Batch insert, conflict graph, MIS, cover radii, subtree bounds -- oracle
 * restatement.  TEST INFRASTRUCTURE ONLY (see oracle.h). */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* CoverTreeData, src/tree.rs:6-125 (flat arrays; payload = tree index). */
ox_tree *ox_tree_new(int64_t cap)
{
    ox_tree *t = (ox_tree *)calloc(1, sizeof(*t));
    t->cap = cap;
    t->parents = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cap + 1));
    t->children = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cap + 1));
    t->next = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cap + 1));
    t->levels = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cap + 1));
    t->si = (float *)calloc((size_t)(cap + 1), sizeof(float));
    /* build_pcct2_residual_tree starts from an empty tree with levels -20..20
     * (src/lib.rs:1472-1490) */
    t->min_level = -20;
    t->max_level = 20;
    return t;
}

void ox_tree_free(ox_tree *t)
{
    if (!t)
        return;
    free(t->parents);
    free(t->children);
    free(t->next);
    free(t->levels);
    free(t->si);
    free(t);
}

/* link_child, src/tree.rs:98-103: prepend. */
static void link_child(ox_tree *t, int64_t parent, int64_t child)
{
    int64_t old_first = t->children[parent];
    t->children[parent] = child;
    t->next[child] = old_first;
}

static void set_level(ox_tree *t, int64_t idx, int32_t level) /* tree.rs:105-113 */
{
    t->levels[idx] = level;
    if (level > t->max_level)
        t->max_level = level;
    if (level < t->min_level)
        t->min_level = level;
}

static void set_parent(ox_tree *t, int64_t idx, int64_t parent) /* tree.rs:115-120 */
{
    t->parents[idx] = parent;
    if (parent >= 0)
        link_child(t, parent, idx);
}

static int64_t grid_key(float v, float scale)
{
    /* (v * scale).floor().to_i64().unwrap_or(0), src/algo/batch/mod.rs:232-235 */
    float s = floorf(v * scale);
    if (isnan(s) || s >= 9.2233720368547758e18f || s < -9.2233720368547758e18f)
        return 0;
    return (int64_t)s;
}

/* Grid conflict predicate, grid_conflicts_for_candidates,
 * src/algo/batch/mod.rs:212-327: two far candidates are adjacent when their
 * floor(coord * scale) cell keys are equal, or (D <= 3 only, since neighbour
 * keys are built over the first min(D,3) axes and only match full-length keys
 * when D <= 3) when the keys are Chebyshev-1 neighbours. */
static int grid_adjacent(const int64_t *ka, const int64_t *kb, int32_t gd)
{
    if (gd <= 3) {
        for (int32_t c = 0; c < gd; c++) {
            int64_t dlt = ka[c] - kb[c];
            if (dlt > 1 || dlt < -1)
                return 0;
        }
        return 1;
    }
    for (int32_t c = 0; c < gd; c++)
        if (ka[c] != kb[c])
            return 0;
    return 1;
}

void ox_grid_edges_dense(int64_t n, int32_t d, const float *coords, float scale, uint8_t *adj)
{
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n * d + 1));
    for (int64_t i = 0; i < n * d; i++)
        keys[i] = grid_key(coords[i], scale);
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++)
            adj[i * n + j] = (i != j) && grid_adjacent(keys + i * d, keys + j * d, d);
    free(keys);
}

/* Conflict record of the batches inserted while a buffer is set (the
 * reference's BatchInsertTelemetry conflict CSR, selected and dominated
 * marks, src/algo/batch/mod.rs:696-800, 802-956), in the layout the device
 * library writes: per level with a far set
 *   1, batch, level, nf, far ids [nf], leader bits [W], covered non-leader
 *   bits [W], distance-conflict bits [nf][W] (d^2 <= r^2, columns j > i),
 *   grid-conflict bits [nf][W]                     (W = ceil(nf / 32))
 * and per batch 2, batch, count, bottom survivors [count].  Words past cap
 * are counted in *used but not written. */
static int32_t *g_log = NULL;
static int64_t g_log_cap = 0;
static int64_t *g_log_used = NULL;
static int32_t g_log_batch = 0;

void ox_set_conflict_log(int32_t *buf, int64_t cap, int64_t *used, int32_t batch)
{
    g_log = buf;
    g_log_cap = cap;
    g_log_used = used;
    g_log_batch = batch;
}

static int32_t *log_reserve(int64_t words)
{
    if (!g_log)
        return NULL;
    int64_t at = *g_log_used;
    *g_log_used += words;
    return at + words <= g_log_cap ? g_log + at : NULL;
}

static void log_set_bit(int32_t *words, int64_t j)
{
    words[j >> 5] = (int32_t)((uint32_t)words[j >> 5] | (1u << (j & 31)));
}

/* stats: [0] filter distance evals, [1] conflict pairs (reference count
 * n_far*(n_far-1)/2), [2] non-leader distance evals, [3] leaders,
 * [4] orphans dropped, [5] levels with a far set */
int64_t ox_batch_insert(ox_tree *t, const ox_metric *mt, int64_t start, int64_t end,
                        const float *grid_coords, int32_t gd, float gscale, int32_t use_grid,
                        int64_t *stats)
{
    /* batch_insert_collect, src/algo/batch/mod.rs:447-800 */
    for (int64_t i = start; i < end; i++) { /* add_point(row, i32::MIN, -1), 511-513 */
        t->parents[i] = -1;
        t->children[i] = -1;
        t->next[i] = -1;
        t->levels[i] = OX_LEVEL_UNSET;
        t->si[i] = 0.0f;
    }
    t->n = end;
    int64_t nb = end - start;
    if (nb <= 0)
        return 0;
    int64_t *cands = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
    int64_t *covers = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
    int64_t *ncands = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
    int64_t *ncovers = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
    int64_t *best = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
    int64_t *far = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
    int64_t *far_orig = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
    int8_t *state = (int8_t *)malloc((size_t)nb);
    int64_t *leaders = (int64_t *)malloc(sizeof(int64_t) * (size_t)nb);
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nb * (gd > 0 ? gd : 1)));
    int64_t *evals = (int64_t *)calloc((size_t)nb, sizeof(int64_t));
    int64_t nc = 0;
    for (int64_t i = start; i < end; i++)
        cands[nc++] = i;
    if (start == 0 && nc > 0) { /* bootstrap root, 521-531 */
        t->min_level = -10;
        t->max_level = 10;
        set_level(t, cands[0], 10);
        memmove(cands, cands + 1, sizeof(int64_t) * (size_t)(nc - 1));
        nc--;
    }
    for (int64_t i = 0; i < nc; i++)
        covers[i] = 0; /* active sets start as [root_idx = 0], 539-540 */
    int32_t level = t->max_level - 1;
    const int32_t min_level = t->min_level;
    while (nc > 0 && level >= min_level) {
        const float r = ldexpf(1.0f, level);
        const float r2 = r * r;
        /* near/far filter over each cover and its children, 580-627 */
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < nc; i++) {
            int64_t q = cands[i], p = covers[i];
            float bd = FLT_MAX;
            int64_t bn = -1, ev = 0;
            float dp = ox_dist_sq(mt, q, p);
            ev++;
            if (dp <= r2 && dp < bd) {
                bd = dp;
                bn = p;
            }
            int64_t c = t->children[p];
            while (c != -1) {
                float dc = ox_dist_sq(mt, q, c);
                ev++;
                if (dc <= r2 && dc < bd) {
                    bd = dc;
                    bn = c;
                }
                int64_t nx = t->next[c];
                if (nx == c)
                    break;
                c = nx;
                if (c == t->children[p])
                    break;
            }
            best[i] = bn;
            evals[i] += ev;
        }
        int64_t nn = 0, nf = 0;
        for (int64_t i = 0; i < nc; i++) {
            if (best[i] >= 0) {
                ncands[nn] = cands[i];
                ncovers[nn] = best[i];
                nn++;
            } else {
                far[nf] = cands[i];
                far_orig[nf] = i;
                nf++;
            }
        }
        if (nf > 0) {
            if (stats) {
                stats[1] += nf * (nf - 1) / 2;
                stats[5] += 1;
            }
            if (use_grid && grid_coords)
                for (int64_t i = 0; i < nf; i++)
                    for (int32_t c = 0; c < gd; c++)
                        keys[i * gd + c] = grid_key(grid_coords[far[i] * gd + c], gscale);
            /* greedy MIS in far-list order over pairwise (d_sq <= r^2, 370-400)
             * and grid edges (662-674); compute_mis_greedy, mis.rs:9-59 */
            memset(state, 0, (size_t)nf);
            for (int64_t i = 0; i < nf; i++) {
                if (state[i] != 0)
                    continue;
                state[i] = 1;
#pragma omp parallel for schedule(static)
                for (int64_t j = i + 1; j < nf; j++) {
                    if (state[j] != 0)
                        continue;
                    int adj = 0;
                    if (use_grid && grid_coords && gd > 0)
                        adj = grid_adjacent(keys + i * gd, keys + j * gd, gd);
                    if (!adj)
                        adj = ox_dist_sq(mt, far[i], far[j]) <= r2;
                    if (adj)
                        state[j] = -1;
                }
            }
            /* leaders: level and parent = previous cover, 694-714 */
            int64_t nl = 0;
            for (int64_t i = 0; i < nf; i++) {
                if (state[i] != 1)
                    continue;
                leaders[nl++] = far[i];
                set_level(t, far[i], level);
                set_parent(t, far[i], covers[far_orig[i]]);
            }
            if (stats)
                stats[3] += nl;
            /* non-leaders: nearest leader within r (strict <, first wins) or
             * dropped, 716-769 */
            int64_t *nl_best = (int64_t *)malloc(sizeof(int64_t) * (size_t)nf);
#pragma omp parallel for schedule(dynamic, 8)
            for (int64_t i = 0; i < nf; i++) {
                nl_best[i] = -1;
                if (state[i] == 1)
                    continue;
                float bd = FLT_MAX;
                for (int64_t l = 0; l < nl; l++) {
                    float dl = ox_dist_sq(mt, far[i], leaders[l]);
                    if (dl <= r2 && dl < bd) {
                        bd = dl;
                        nl_best[i] = leaders[l];
                    }
                }
            }
            if (g_log) {
                const int64_t W = (nf + 31) / 32;
                int32_t *e = log_reserve(4 + nf + 2 * W + 2 * nf * W);
                if (e) {
                    memset(e, 0, sizeof(int32_t) * (size_t)(4 + nf + 2 * W + 2 * nf * W));
                    e[0] = 1;
                    e[1] = g_log_batch;
                    e[2] = level;
                    e[3] = (int32_t)nf;
                    int32_t *sel = e + 4 + nf, *cov = sel + W, *db = cov + W, *gb = db + nf * W;
                    for (int64_t i = 0; i < nf; i++) {
                        e[4 + i] = (int32_t)far[i];
                        if (state[i] == 1)
                            log_set_bit(sel, i);
                        else if (nl_best[i] >= 0)
                            log_set_bit(cov, i);
                    }
                    /* every far pair, as compute_candidate_conflicts (370-400)
                     * and grid_conflicts_for_candidates (212-327) list them */
#pragma omp parallel for schedule(dynamic, 8)
                    for (int64_t i = 0; i < nf; i++)
                        for (int64_t j = i + 1; j < nf; j++) {
                            if (ox_dist_sq(mt, far[i], far[j]) <= r2)
                                log_set_bit(db + i * W, j);
                            if (use_grid && grid_coords && gd > 0 &&
                                grid_adjacent(keys + i * gd, keys + j * gd, gd))
                                log_set_bit(gb + i * W, j);
                        }
                }
            }
            for (int64_t i = 0; i < nf; i++) {
                if (state[i] == 1)
                    continue;
                if (stats)
                    stats[2] += nl;
                if (nl_best[i] >= 0) {
                    ncands[nn] = far[i];
                    ncovers[nn] = nl_best[i];
                    nn++;
                } else if (stats) {
                    stats[4] += 1;
                }
            }
            free(nl_best);
        }
        int64_t *tmp = cands;
        cands = ncands;
        ncands = tmp;
        tmp = covers;
        covers = ncovers;
        ncovers = tmp;
        nc = nn;
        level -= 1;
    }
    /* survivors at the bottom, 791-800 */
    for (int64_t i = 0; i < nc; i++) {
        set_level(t, cands[i], min_level);
        set_parent(t, cands[i], covers[i]);
    }
    if (g_log) {
        int32_t *e = log_reserve(3 + nc);
        if (e) {
            e[0] = 2;
            e[1] = g_log_batch;
            e[2] = (int32_t)nc;
            for (int64_t i = 0; i < nc; i++)
                e[3 + i] = (int32_t)cands[i];
        }
    }
    if (stats)
        for (int64_t i = 0; i < nb; i++)
            stats[0] += evals[i];
    free(cands);
    free(covers);
    free(ncands);
    free(ncovers);
    free(best);
    free(far);
    free(far_orig);
    free(state);
    free(leaders);
    free(keys);
    free(evals);
    return nb;
}

/* build_pcct2_residual_tree's chunk loop, src/lib.rs:1502-1539. */
void ox_build(ox_tree *t, const ox_metric *mt, int64_t chunk, const float *grid_coords,
              int32_t gd, float gscale, int32_t use_grid, int64_t *stats)
{
    int64_t n = mt->n;
    if (chunk <= 0)
        chunk = n;
    for (int64_t s = 0; s < n; s += chunk) {
        int64_t e = s + chunk < n ? s + chunk : n;
        ox_batch_insert(t, mt, s, e, grid_coords, gd, gscale, use_grid, stats);
    }
}

/* compute_si_cache_residual, src/algo.rs:52-106: post-order DFS from node 0
 * only; r(node) = max over children of dist(node, child) + r(child), with
 * dist = sqrt(distance_sq_idx).  Iterative to survive deep chains. */
void ox_si_cache(const ox_tree *t, const ox_metric *mt, float *out)
{
    int64_t n = t->n;
    for (int64_t i = 0; i < n; i++)
        out[i] = 0.0f;
    if (n == 0)
        return;
    int64_t *stack = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    float *acc = (float *)calloc((size_t)(n + 1), sizeof(float));
    int64_t sp = 0;
    stack[sp] = 0;
    cur[0] = t->children[0];
    acc[0] = 0.0f;
    sp = 1;
    while (sp > 0) {
        int64_t node = stack[sp - 1];
        int64_t c = cur[node];
        if (c != -1) {
            /* descend into child c first (its radius feeds the bound) */
            stack[sp++] = c;
            cur[c] = t->children[c];
            acc[c] = 0.0f;
            continue;
        }
        out[node] = acc[node];
        sp--;
        if (sp == 0)
            break;
        int64_t parent = stack[sp - 1];
        float dist = sqrtf(ox_dist_sq(mt, parent, node));
        float bound = dist + out[node];
        if (bound > acc[parent])
            acc[parent] = bound;
        /* advance the parent's sibling cursor with the reference's guards */
        int64_t nx = t->next[node];
        if (nx == -1 || nx == node)
            cur[parent] = -1;
        else if (nx == t->children[parent])
            cur[parent] = -1;
        else
            cur[parent] = nx;
    }
    free(stack);
    free(cur);
    free(acc);
}

/* compute_subtree_index_bounds, src/tree.rs:144-185: child->parent relaxation
 * to a fixpoint (at most n sweeps). */
void ox_subtree_bounds(int64_t n, const int64_t *parents, const int64_t *n2d, int64_t *mn,
                       int64_t *mx)
{
    for (int64_t i = 0; i < n; i++) {
        mn[i] = n2d[i];
        mx[i] = n2d[i];
    }
    int changed = 1;
    int64_t iters = 0;
    while (changed && iters < n) {
        changed = 0;
        iters++;
        for (int64_t node = 0; node < n; node++) {
            int64_t p = parents[node];
            if (p >= 0) {
                if (mn[node] < mn[p]) {
                    mn[p] = mn[node];
                    changed = 1;
                }
                if (mx[node] > mx[p]) {
                    mx[p] = mx[node];
                    changed = 1;
                }
            }
        }
    }
}

/* compute_mis_greedy over a CSR graph, src/algo/batch/mis.rs:9-59 with equal
 * priorities (a stable sort keeps index order). */
void ox_greedy_mis(int64_t n, const int64_t *indptr, const int64_t *indices, uint8_t *out)
{
    int8_t *st = (int8_t *)calloc((size_t)(n + 1), 1);
    for (int64_t i = 0; i < n; i++) {
        if (st[i] != 0)
            continue;
        st[i] = 1;
        for (int64_t e = indptr[i]; e < indptr[i + 1]; e++)
            if (st[indices[e]] == 0)
                st[indices[e]] = -1;
    }
    for (int64_t i = 0; i < n; i++)
        out[i] = st[i] == 1;
    free(st);
}

/* _hash_priority, covertreex/algo/_mis_numba.py:19-28, with the 32-bit
 * wrapping of the JAX path (covertreex/algo/mis.py:146-168). */
static double luby_priority(int64_t node, uint32_t seed, int64_t iteration)
{
    uint32_t x = (uint32_t)node ^ seed;
    x ^= (uint32_t)((uint64_t)iteration * 0x9E3779B9ULL);
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return ((double)x + 1.0) / 4294967296.0;
}

/* _run_mis_numba_impl, covertreex/algo/_mis_numba.py:64-131: greedy when the
 * max degree is <= 8, else Luby rounds.  Returns the iteration count. */
int64_t ox_luby_mis(int64_t n, const int64_t *indptr, const int64_t *indices, uint32_t seed,
                    uint8_t *out)
{
    int64_t maxdeg = 0;
    for (int64_t i = 0; i < n; i++)
        if (indptr[i + 1] - indptr[i] > maxdeg)
            maxdeg = indptr[i + 1] - indptr[i];
    if (maxdeg <= 8) {
        uint8_t *active = (uint8_t *)malloc((size_t)(n + 1));
        memset(active, 1, (size_t)(n + 1));
        for (int64_t i = 0; i < n; i++) {
            out[i] = 0;
        }
        for (int64_t i = 0; i < n; i++) {
            if (!active[i])
                continue;
            out[i] = 1;
            for (int64_t e = indptr[i]; e < indptr[i + 1]; e++)
                active[indices[e]] = 0;
            active[i] = 0;
        }
        free(active);
        return 1;
    }
    uint8_t *active = (uint8_t *)malloc((size_t)(n + 1));
    uint8_t *win = (uint8_t *)malloc((size_t)(n + 1));
    uint8_t *dom = (uint8_t *)malloc((size_t)(n + 1));
    double *prio = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    memset(active, 1, (size_t)n);
    memset(out, 0, (size_t)n);
    int64_t iters = 0;
    for (;;) {
        int any = 0;
        for (int64_t i = 0; i < n; i++)
            any |= active[i];
        if (!any)
            break;
        iters++;
        for (int64_t i = 0; i < n; i++)
            prio[i] = luby_priority(i, seed, iters);
        int anyw = 0;
        for (int64_t i = 0; i < n; i++) {
            win[i] = 0;
            if (!active[i])
                continue;
            int best = 1;
            for (int64_t e = indptr[i]; e < indptr[i + 1]; e++) {
                int64_t nb = indices[e];
                if (!active[nb])
                    continue;
                if (prio[nb] > prio[i] || (prio[nb] == prio[i] && nb < i)) {
                    best = 0;
                    break;
                }
            }
            if (best) {
                win[i] = 1;
                anyw = 1;
            }
        }
        if (!anyw)
            for (int64_t i = 0; i < n; i++)
                if (active[i]) {
                    win[i] = 1;
                    break;
                }
        memset(dom, 0, (size_t)n);
        for (int64_t i = 0; i < n; i++)
            if (win[i])
                for (int64_t e = indptr[i]; e < indptr[i + 1]; e++)
                    dom[indices[e]] = 1;
        for (int64_t i = 0; i < n; i++) {
            if (win[i])
                out[i] = 1;
            if (win[i] || dom[i])
                active[i] = 0;
        }
    }
    free(active);
    free(win);
    free(dom);
    free(prio);
    return iters;
}
