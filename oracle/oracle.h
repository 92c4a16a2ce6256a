/*
 * CPU oracle for the residual-correlation cover-tree path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference's rust-hilbert build + query
 * semantics.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / the timed CPU
 * baseline; the product library never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Floating point
 * follows the reference's f32 evaluation order exactly (compile with
 * -ffp-contract=off, no -ffast-math):
 *   - distance_sq_idx_f32      src/metric.rs:489-563
 *   - distances_f32_batch...   src/metric.rs:376-487
 *   - dot_product_simd_slice   src/metric.rs:693-733
 *   - dot_f32_simd_inline      src/metric.rs:601-643
 * The `wide` crate's f32x8 (SSE2 baseline target, no AVX/FMA features, as
 * maturin builds by default) is two f32x4 halves: mul_add is an unfused
 * multiply then add, and reduce_add sums each half left to right and adds
 * the two half sums.  That association is this oracle's choice and is
 * parity unpinned at the bit level (the crate cannot be built here).
 *
 * exp: OX_EXP_LIBM calls the platform expf, which is what Rust's f32::exp
 * lowers to on Linux (the reference's behaviour on this platform);
 * OX_EXP_CR rounds a double-precision exp to float (correctly rounded
 * except in cases within 2^-50 of a float rounding midpoint), which is what
 * the GPU build kernels compute.
 */
#ifndef COVERTREEX_ORACLE_H
#define COVERTREEX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OX_EXP_LIBM = 0, OX_EXP_CR = 1 };
enum { OX_KERNEL_RBF = 0, OX_KERNEL_MATERN52 = 1 };
enum { OX_METRIC_RESIDUAL = 0, OX_METRIC_EUCLIDEAN = 1 };

#define OX_LEVEL_UNSET INT32_MIN

typedef struct ox_metric {
    int64_t n;
    int32_t m, d;
    int32_t kind;        /* OX_METRIC_RESIDUAL / OX_METRIC_EUCLIDEAN */
    int32_t kernel_type; /* OX_KERNEL_* */
    int32_t exp_mode;    /* OX_EXP_* */
    float var;
    float *scaled;       /* n*d, coords * inv_ls (residual) or raw points (euclidean) */
    float *norms;        /* n, sum of squares of scaled rows */
    float *v;            /* n*m (residual only) */
    float *p;            /* n (residual only) */
} ox_metric;

typedef struct ox_tree {
    int64_t n, cap;
    int64_t *parents, *children, *next;
    int32_t *levels;
    float *si;
    int32_t min_level, max_level;
} ox_tree;

/* ---- metric ---------------------------------------------------------- */
ox_metric *ox_metric_residual(int64_t n, int32_t m, int32_t d, const float *coords,
                              const float *v, const float *p, float var,
                              const float *ls, int32_t n_ls, int32_t kernel_type,
                              int32_t exp_mode);
ox_metric *ox_metric_euclidean(int64_t n, int32_t d, const float *points);
void ox_metric_free(ox_metric *mt);
float ox_expf(int32_t mode, float x);
/* ox_expf of the count floats whose bit patterns follow first_bits */
void ox_expf_range(int32_t mode, uint32_t first_bits, int64_t count, float *out);
void ox_expf_array(int32_t mode, const float *x, int64_t n, float *out);
float ox_dist_sq(const ox_metric *mt, int64_t i, int64_t j);     /* build / roots */
float ox_dist_batch(const ox_metric *mt, int64_t q, int64_t j);  /* query tiles */
void ox_dist_sq_pairs(const ox_metric *mt, const int64_t *a, const int64_t *b,
                      int64_t n, float *out);
void ox_dist_batch_pairs(const ox_metric *mt, const int64_t *a, const int64_t *b,
                         int64_t n, float *out);

/* ---- ordering -------------------------------------------------------- */
void ox_morton_order(int64_t n, int32_t d, const float *coords, int64_t *order_out,
                     uint64_t *codes_out);
void ox_hilbert_codes(int64_t n, int32_t d, const double *points, uint64_t *codes_out);

/* ---- tree ------------------------------------------------------------ */
ox_tree *ox_tree_new(int64_t cap);
void ox_tree_free(ox_tree *t);
int64_t ox_batch_insert(ox_tree *t, const ox_metric *mt, int64_t start, int64_t end,
                        const float *grid_coords, int32_t grid_d, float grid_scale,
                        int32_t use_grid, int64_t *stats);
void ox_build(ox_tree *t, const ox_metric *mt, int64_t chunk, const float *grid_coords,
              int32_t grid_d, float grid_scale, int32_t use_grid, int64_t *stats);
/* Record the conflict graph of later ox_batch_insert calls into buf (cap
 * words, *used counts words; layout in build.c) as batch `batch`; NULL stops. */
void ox_set_conflict_log(int32_t *buf, int64_t cap, int64_t *used, int32_t batch);
void ox_si_cache(const ox_tree *t, const ox_metric *mt, float *out);
void ox_subtree_bounds(int64_t n, const int64_t *parents, const int64_t *n2d,
                       int64_t *min_out, int64_t *max_out);
void ox_greedy_mis(int64_t n, const int64_t *indptr, const int64_t *indices, uint8_t *out);
int64_t ox_luby_mis(int64_t n, const int64_t *indptr, const int64_t *indices, uint32_t seed,
                    uint8_t *out);
void ox_grid_edges_dense(int64_t n, int32_t d, const float *coords, float scale,
                         uint8_t *adj);

/* ---- queries --------------------------------------------------------- */
typedef struct ox_query_opts {
    int32_t n_budget;          /* 0 = unlimited (exact semantics) */
    int64_t budget[8];
    double budget_up, budget_down;
    int32_t stream_tile;       /* 64 */
    int32_t dynamic_block;     /* 1 */
    float radius_floor;        /* 1e-3 */
    float cap_default;         /* 2.0 */
    int32_t threads;           /* 0 = OpenMP default */
} ox_query_opts;

void ox_query_opts_default(ox_query_opts *o);
/* Residual kNN over a tree built on tree-ordered data.  q_tree: query as tree
 * index; q_ds: query dataset index (predecessor bound); n2d: node->dataset map
 * (the tree's Morton order).  Outputs are dataset indices padded -1/FLT_MAX. */
void ox_knn_residual(const ox_tree *t, const ox_metric *mt, const int64_t *n2d,
                     const int64_t *q_tree, const int64_t *q_ds, int64_t nq, int32_t k,
                     int32_t predecessor, const int64_t *subtree_min,
                     const ox_query_opts *opts, int64_t *out_idx, float *out_dist,
                     int64_t *evals_out);
void ox_knn_euclidean(const ox_tree *t, const ox_metric *mt, const float *queries,
                      int64_t nq, int32_t k, int32_t predecessor, int64_t *out_idx,
                      float *out_dist, int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
