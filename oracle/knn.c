/* Residual and Euclidean kNN over the cover tree -- oracle restatement.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void ox_query_opts_default(ox_query_opts *o)
{
    memset(o, 0, sizeof(*o));
    /* default budget ladder [32,64,96], up 0.015, down 0.002, src/algo.rs:653-676 */
    o->n_budget = 3;
    o->budget[0] = 32;
    o->budget[1] = 64;
    o->budget[2] = 96;
    o->budget_up = 0.015;
    o->budget_down = 0.002;
    o->stream_tile = 64;     /* 752-760 */
    o->dynamic_block = 1;    /* 796-802 */
    o->radius_floor = 1e-3f; /* 814-818 */
    o->cap_default = 2.0f;   /* src/metric.rs:155 */
    o->threads = 0;
}

/* ---- std::collections::BinaryHeap<Neighbor> (max-heap by distance) -------
 * Restated so that ties resolve as the reference's heap does: sift_up moves
 * while strictly greater than the parent; pop sifts the hole to the bottom
 * choosing the right child when left <= right, then sifts up; into_sorted_vec
 * swaps the max to the end and sifts down.  NaN compares Equal. */
typedef struct {
    float d;
    int64_t node;
} nb_t;

static inline int gt(float a, float b) { return a > b; }
static inline int lt(float a, float b) { return a < b; }
static inline int le(float a, float b) { return !(a > b); }
static inline int ge(float a, float b) { return !(a < b); }

typedef struct {
    nb_t *a;
    int64_t len, cap;
} heap_t;

static void heap_reserve(heap_t *h, int64_t need)
{
    if (need <= h->cap)
        return;
    int64_t nc = h->cap ? h->cap * 2 : 64;
    while (nc < need)
        nc *= 2;
    h->a = (nb_t *)realloc(h->a, sizeof(nb_t) * (size_t)nc);
    h->cap = nc;
}

static int64_t sift_up(heap_t *h, int64_t start, int64_t pos)
{
    nb_t hole = h->a[pos];
    while (pos > start) {
        int64_t parent = (pos - 1) / 2;
        if (le(hole.d, h->a[parent].d))
            break;
        h->a[pos] = h->a[parent];
        pos = parent;
    }
    h->a[pos] = hole;
    return pos;
}

static void heap_push(heap_t *h, nb_t x)
{
    heap_reserve(h, h->len + 1);
    h->a[h->len++] = x;
    sift_up(h, 0, h->len - 1);
}

static void sift_down_to_bottom(heap_t *h, int64_t pos)
{
    int64_t end = h->len, start = pos;
    nb_t hole = h->a[pos];
    int64_t child = 2 * pos + 1;
    int64_t lim = end >= 2 ? end - 2 : 0;
    while (end >= 2 && child <= lim) {
        child += le(h->a[child].d, h->a[child + 1].d) ? 1 : 0;
        h->a[pos] = h->a[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (child == end - 1) {
        h->a[pos] = h->a[child];
        pos = child;
    }
    h->a[pos] = hole;
    sift_up(h, start, pos);
}

static void heap_pop(heap_t *h)
{
    if (h->len == 0)
        return;
    nb_t last = h->a[--h->len];
    if (h->len > 0) {
        h->a[0] = last;
        sift_down_to_bottom(h, 0);
    }
}

static void sift_down_range(heap_t *h, int64_t pos, int64_t end)
{
    nb_t hole = h->a[pos];
    int64_t child = 2 * pos + 1;
    int64_t lim = end >= 2 ? end - 2 : 0;
    while (end >= 2 && child <= lim) {
        child += le(h->a[child].d, h->a[child + 1].d) ? 1 : 0;
        if (ge(hole.d, h->a[child].d)) {
            h->a[pos] = hole;
            return;
        }
        h->a[pos] = h->a[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (child == end - 1 && lt(hole.d, h->a[child].d)) {
        h->a[pos] = h->a[child];
        pos = child;
    }
    h->a[pos] = hole;
}

static void heap_into_sorted(heap_t *h)
{
    int64_t end = h->len;
    while (end > 1) {
        end--;
        nb_t t = h->a[0];
        h->a[0] = h->a[end];
        h->a[end] = t;
        sift_down_range(h, 0, end);
    }
}

/* ---- per-thread search context (SearchContext, src/algo.rs:329-360) ---- */
typedef struct {
    int64_t node;
    float dist;
} fr_t;

typedef struct {
    float *cached_lb;
    int64_t *touched;
    int64_t n_touched;
    fr_t *frontier, *next;
    int64_t nf, nn, capf, capn;
    int64_t *ch_nodes;
    float *ch_lb;
    int64_t nch, capch;
    heap_t heap;
} ctx_t;

static void fr_push(fr_t **arr, int64_t *len, int64_t *cap, int64_t node, float d)
{
    if (*len + 1 > *cap) {
        *cap = *cap ? *cap * 2 : 128;
        *arr = (fr_t *)realloc(*arr, sizeof(fr_t) * (size_t)*cap);
    }
    (*arr)[*len].node = node;
    (*arr)[*len].dist = d;
    (*len)++;
}

static void ch_push(ctx_t *c, int64_t node, float lb)
{
    if (c->nch + 1 > c->capch) {
        c->capch = c->capch ? c->capch * 2 : 128;
        c->ch_nodes = (int64_t *)realloc(c->ch_nodes, sizeof(int64_t) * (size_t)c->capch);
        c->ch_lb = (float *)realloc(c->ch_lb, sizeof(float) * (size_t)c->capch);
    }
    c->ch_nodes[c->nch] = node;
    c->ch_lb[c->nch] = lb;
    c->nch++;
}

static inline void lb_set(ctx_t *c, int64_t idx, float v)
{
    if (c->cached_lb[idx] == FLT_MAX)
        c->touched[c->n_touched++] = idx;
    c->cached_lb[idx] = v;
}

static inline float radius_of(const ox_tree *t, const ox_query_opts *o, int64_t node)
{
    /* si_cache radius, apply_level_cap (src/metric.rs:202-213), floor */
    float r = t->si[node];
    if (o->cap_default > 0.0f && o->cap_default < r)
        r = o->cap_default;
    if (r < o->radius_floor)
        r = o->radius_floor;
    return r;
}

typedef struct {
    float d;
    int64_t node;
    float lb;
} ord_t;

/* single_residual_knn_query, src/algo.rs:612-1242 (parity mode off). */
static void knn_one(const ox_tree *t, const ox_metric *mt, const int64_t *n2d, int64_t q,
                    int64_t q_ds, int32_t k, int32_t pred, const int64_t *smin,
                    const ox_query_opts *o, ctx_t *c, int64_t *out_idx, float *out_dist,
                    int64_t *evals)
{
    for (int32_t j = 0; j < k; j++) {
        out_idx[j] = -1;
        out_dist[j] = FLT_MAX;
    }
    int64_t ev = 0;
    const int64_t n = t->n;
    if (n == 0 || (pred && q_ds == 0)) { /* 629-638 */
        *evals = 0;
        return;
    }
    const int64_t max_idx = q_ds;
    for (int64_t i = 0; i < c->n_touched; i++) /* cached_lb.fill(MAX), 641 */
        c->cached_lb[c->touched[i]] = FLT_MAX;
    c->n_touched = 0;
    heap_t *h = &c->heap;
    h->len = 0;
    int64_t budget_idx = 0;
    int64_t budget_limit = o->n_budget > 0 ? o->budget[0] : INT64_MAX;
    int64_t survivors = 0, low_yield = 0;
    float kth = FLT_MAX;
    int all_roots_prunable = 1;
    c->nf = 0;
    c->nn = 0;
    /* roots: every node with parent == -1, 688-734 */
    for (int64_t root = 0; root < n; root++) {
        if (t->parents[root] != -1)
            continue;
        int64_t rds = n2d[root];
        int valid = pred ? rds < max_idx : 1;
        float rd = sqrtf(ox_dist_sq(mt, q, root));
        ev++;
        fr_push(&c->frontier, &c->nf, &c->capf, root, rd);
        if (valid) {
            nb_t x = {rd, root};
            heap_push(h, x);
            survivors++;
            if (rd < kth)
                kth = rd;
        }
        if (pred && smin) {
            if (smin[root] < max_idx)
                all_roots_prunable = 0;
        } else {
            all_roots_prunable = 0;
        }
    }
    int early = all_roots_prunable && pred; /* 737-749 */
    ord_t ord[256];
    int64_t ev_t[256];
    float dist_t[256];
    while (!early && c->nf > 0 &&
           (survivors < budget_limit || (pred && h->len < k))) { /* 828-829 */
        c->nch = 0;
        for (int64_t fi = 0; fi < c->nf; fi++) { /* gather, 862-995 */
            int64_t parent = c->frontier[fi].node;
            float pdist = c->frontier[fi].dist;
            float prad = radius_of(t, o, parent);
            float plb = pdist - prad;
            if (h->len == k && plb > kth)
                continue;
            int64_t ch = t->children[parent];
            while (ch != -1) {
                int64_t nx;
                if (pred && smin && smin[ch] >= max_idx) { /* subtree pruned, 900-916 */
                    nx = t->next[ch];
                    if (nx == ch || nx == -1)
                        break;
                    ch = nx;
                    if (ch == t->children[parent])
                        break;
                    continue;
                }
                int ok = pred ? n2d[ch] < max_idx : 1;
                if (!ok) { /* 927-956 */
                    int sub_valid = smin ? smin[ch] < max_idx : 1;
                    if (sub_valid) {
                        float dd = sqrtf(ox_dist_sq(mt, q, ch));
                        ev++;
                        fr_push(&c->next, &c->nn, &c->capn, ch, dd);
                    }
                    nx = t->next[ch];
                    if (nx == ch || nx == -1)
                        break;
                    ch = nx;
                    if (ch == t->children[parent])
                        break;
                    continue;
                }
                ch_push(c, ch, plb);
                if (c->cached_lb[ch] > plb)
                    lb_set(c, ch, plb);
                nx = t->next[ch];
                if (nx == ch)
                    break;
                ch = nx;
                if (ch == t->children[parent])
                    break;
            }
        }
        int64_t start = 0;
        while (start < c->nch && (survivors < budget_limit || (pred && h->len < k))) {
            int64_t active = c->nch - start;
            int64_t tile;
            if (o->dynamic_block) { /* 1001-1022 */
                int64_t wave = c->nf > 1 ? c->nf : 1;
                int64_t comb = active + wave;
                int64_t ad = comb <= 32 ? 16 : comb <= 128 ? 32 : comb <= 512 ? 64 : 128;
                tile = ad < o->stream_tile ? ad : o->stream_tile;
            } else {
                tile = o->stream_tile;
            }
            if (tile > 256)
                tile = 256;
            int64_t end = start + tile < c->nch ? start + tile : c->nch;
            int64_t len = end - start;
            if (h->len == k) { /* chunk lower-bound prune, 1032-1045 */
                float mn = c->ch_lb[start];
                for (int64_t j = start + 1; j < end; j++)
                    if (c->ch_lb[j] < mn)
                        mn = c->ch_lb[j];
                if (mn > kth) {
                    start = end;
                    continue;
                }
            }
            int64_t ne = 0;
            for (int64_t j = 0; j < len; j++) { /* cached_lb prune, 1053-1066 */
                int64_t node = c->ch_nodes[start + j];
                dist_t[j] = FLT_MAX;
                if (h->len == k && c->cached_lb[node] > kth)
                    continue;
                ev_t[ne++] = j;
            }
            for (int64_t e = 0; e < ne; e++) {
                int64_t j = ev_t[e];
                dist_t[j] = ox_dist_batch(mt, q, c->ch_nodes[start + j]);
            }
            ev += ne;
            int64_t no = 0; /* sort the evaluated tile by distance, stable, 1103-1109 */
            for (int64_t j = 0; j < len; j++) {
                if (dist_t[j] == FLT_MAX)
                    continue;
                ord_t x = {dist_t[j], c->ch_nodes[start + j], c->ch_lb[start + j]};
                int64_t p = no++;
                while (p > 0 && lt(x.d, ord[p - 1].d)) {
                    ord[p] = ord[p - 1];
                    p--;
                }
                ord[p] = x;
            }
            int64_t added = 0;
            for (int64_t e = 0; e < no; e++) { /* 1111-1189 */
                float dist = ord[e].d;
                int64_t child = ord[e].node;
                float crad = radius_of(t, o, child);
                if (h->len == k) {
                    float lbv = fminf(ord[e].lb, dist - crad); /* f32::min */
                    if (lbv > kth)
                        continue;
                }
                float nl = dist - crad;
                if (nl < c->cached_lb[child])
                    lb_set(c, child, nl);
                nb_t x = {dist, child};
                heap_push(h, x);
                if (h->len > k)
                    heap_pop(h);
                if (h->len > 0)
                    kth = h->a[0].d;
                survivors++;
                added++;
                int cont = survivors < budget_limit || (pred && h->len < k);
                if (cont) {
                    int sub_valid = (pred && smin) ? smin[child] < max_idx : 1;
                    if (sub_valid)
                        fr_push(&c->next, &c->nn, &c->capn, child, dist);
                } else {
                    break;
                }
            }
            if (o->n_budget > 0 && len > 0) { /* budget ladder, 1191-1219 */
                double ratio = (double)added / (double)len;
                if (ratio >= o->budget_up) {
                    if (budget_idx + 1 < o->n_budget) {
                        budget_idx++;
                        budget_limit = o->budget[budget_idx];
                    }
                    low_yield = 0;
                } else if (ratio < o->budget_down) {
                    low_yield++;
                    int need_more = pred && h->len < k;
                    if (low_yield >= 2 && !need_more)
                        break;
                } else {
                    low_yield = 0;
                }
            }
            start = end;
        }
        /* frontier = next_frontier, 1224-1225 */
        fr_t *tmp = c->frontier;
        int64_t tcap = c->capf;
        c->frontier = c->next;
        c->capf = c->capn;
        c->nf = c->nn;
        c->next = tmp;
        c->capn = tcap;
        c->nn = 0;
    }
    heap_into_sorted(h);
    int64_t outn = h->len < k ? h->len : k; /* to_py_arrays truncates, src/lib.rs:930-935 */
    for (int64_t j = 0; j < outn; j++) {
        out_idx[j] = n2d[h->a[j].node];
        out_dist[j] = h->a[j].d;
    }
    *evals = ev;
}

void ox_knn_residual(const ox_tree *t, const ox_metric *mt, const int64_t *n2d,
                     const int64_t *q_tree, const int64_t *q_ds, int64_t nq, int32_t k,
                     int32_t predecessor, const int64_t *subtree_min,
                     const ox_query_opts *opts, int64_t *out_idx, float *out_dist,
                     int64_t *evals_out)
{
    ox_query_opts def;
    if (!opts) {
        ox_query_opts_default(&def);
        opts = &def;
    }
    int64_t n = t->n;
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = opts->threads > 0 ? opts->threads : omp_get_max_threads();
#endif
    /* rayon par_chunks(64) with a thread-local context, src/algo.rs:427-460 */
#pragma omp parallel num_threads(nthreads)
    {
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.cached_lb = (float *)malloc(sizeof(float) * (size_t)(n + 1));
        c.touched = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
        for (int64_t i = 0; i <= n; i++)
            c.cached_lb[i] = FLT_MAX;
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < nq; i++) {
            int64_t e = 0;
            knn_one(t, mt, n2d, q_tree[i], q_ds[i], k, predecessor, subtree_min, opts, &c,
                    out_idx + i * k, out_dist + i * k, &e);
            if (evals_out)
                evals_out[i] = e;
        }
        free(c.cached_lb);
        free(c.touched);
        free(c.frontier);
        free(c.next);
        free(c.ch_nodes);
        free(c.ch_lb);
        free(c.heap.a);
    }
}

/* ---- Euclidean best-first kNN, single_knn_query, src/algo.rs:208-323 ---- */
typedef struct {
    float d;
    int64_t node;
} cand_t;

/* candidate heap: Candidate's Ord is reversed (min-heap by distance) */
typedef struct {
    cand_t *a;
    int64_t len, cap;
} cheap_t;

static void cpush(cheap_t *h, cand_t x)
{
    if (h->len + 1 > h->cap) {
        h->cap = h->cap ? h->cap * 2 : 256;
        h->a = (cand_t *)realloc(h->a, sizeof(cand_t) * (size_t)h->cap);
    }
    int64_t pos = h->len++;
    /* sift_up with reversed order: move up while x is strictly closer */
    while (pos > 0) {
        int64_t parent = (pos - 1) / 2;
        if (!(x.d < h->a[parent].d))
            break;
        h->a[pos] = h->a[parent];
        pos = parent;
    }
    h->a[pos] = x;
}

static cand_t cpop(cheap_t *h)
{
    cand_t top = h->a[0];
    cand_t last = h->a[--h->len];
    if (h->len > 0) {
        int64_t end = h->len, pos = 0, child = 1;
        int64_t lim = end >= 2 ? end - 2 : 0;
        while (end >= 2 && child <= lim) {
            /* greater in reversed order = smaller distance; right wins on
             * left <= right, i.e. right.d <= left.d */
            child += (!(h->a[child + 1].d > h->a[child].d)) ? 1 : 0;
            h->a[pos] = h->a[child];
            pos = child;
            child = 2 * pos + 1;
        }
        if (child == end - 1) {
            h->a[pos] = h->a[child];
            pos = child;
        }
        while (pos > 0) {
            int64_t parent = (pos - 1) / 2;
            if (!(last.d < h->a[parent].d))
                break;
            h->a[pos] = h->a[parent];
            pos = parent;
        }
        h->a[pos] = last;
    }
    return top;
}

static float euclid_pt(const float *x, const float *y, int32_t d)
{
    float acc = 0.0f;
    for (int32_t c = 0; c < d; c++) {
        float diff = x[c] - y[c];
        acc = acc + diff * diff;
    }
    return sqrtf(acc);
}

void ox_knn_euclidean(const ox_tree *t, const ox_metric *mt, const float *queries,
                      int64_t nq, int32_t k, int32_t predecessor, int64_t *out_idx,
                      float *out_dist, int32_t threads)
{
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = threads > 0 ? threads : omp_get_max_threads();
#endif
    const int64_t n = t->n;
    const int32_t d = mt->d;
#pragma omp parallel num_threads(nthreads)
    {
        cheap_t ch = {0};
        heap_t res = {0};
#pragma omp for schedule(dynamic, 16)
        for (int64_t qi = 0; qi < nq; qi++) {
            int64_t *oi = out_idx + qi * k;
            float *od = out_dist + qi * k;
            for (int32_t j = 0; j < k; j++) {
                oi[j] = -1;
                od[j] = FLT_MAX;
            }
            if (n == 0 || (predecessor && qi == 0))
                continue;
            const float *qp = queries + qi * d;
            ch.len = 0;
            res.len = 0;
            for (int64_t r = 0; r < n; r++) {
                if (t->parents[r] != -1)
                    continue;
                cand_t x = {euclid_pt(qp, mt->scaled + r * d, d), r};
                cpush(&ch, x);
            }
            float kth = FLT_MAX;
            while (ch.len > 0) {
                cand_t cd = cpop(&ch);
                int can_prune = predecessor ? res.len >= k : res.len == k;
                if (can_prune) {
                    int32_t lvl = t->levels[cd.node];
                    /* two.powi(level + 1) with i32 wrap on level = i32::MIN */
                    int32_t e = (int32_t)((uint32_t)lvl + 1u);
                    float radius = e < -160 ? 0.0f : (e > 160 ? INFINITY : ldexpf(1.0f, e));
                    float lbd = cd.d - radius;
                    if (lbd > kth)
                        continue;
                }
                int ok = predecessor ? cd.node < qi : 1;
                if (ok) {
                    nb_t x = {cd.d, cd.node};
                    heap_push(&res, x);
                    if (res.len > k)
                        heap_pop(&res);
                    if (res.len == k)
                        kth = res.a[0].d;
                }
                int64_t c = t->children[cd.node];
                while (c != -1) {
                    cand_t x = {euclid_pt(qp, mt->scaled + c * d, d), c};
                    cpush(&ch, x);
                    int64_t nx = t->next[c];
                    if (nx == c)
                        break;
                    c = nx;
                    if (c == t->children[cd.node])
                        break;
                }
            }
            heap_into_sorted(&res);
            int64_t outn = res.len < k ? res.len : k;
            for (int64_t j = 0; j < outn; j++) {
                oi[j] = res.a[j].node;
                od[j] = res.a[j].d;
            }
        }
        free(ch.a);
        free(res.a);
    }
}
