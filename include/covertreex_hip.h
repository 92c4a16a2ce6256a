/*
 * covertreex_hip: MI355X (gfx950) engine for the parallel compressed cover tree
 * under the residual-correlation metric d = sqrt(1 - |rho|).
 *
 * C-ABI boundary.  Plain pointers and sizes only; every entry point returns a
 * cx_status and never throws.  On failure cx_last_error() returns a
 * thread-local message.  Host-pointer entry points copy their inputs to HBM;
 * the tree owns its device-resident, tree-ordered copies until cx_tree_free.
 * Each tree handle owns one HIP stream and a mutex: calls on one handle are
 * serialised, different handles may run concurrently.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * the covertreex repository).
 */
#ifndef COVERTREEX_HIP_H
#define COVERTREEX_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t cx_status;
enum {
    CX_OK = 0,
    CX_ERR_INVALID = 1,     /* bad argument (the reference raises PyTypeError / panics) */
    CX_ERR_HIP = 2,         /* HIP runtime error */
    CX_ERR_NO_DEVICE = 3,   /* no usable gfx950 device */
    CX_ERR_UNSUPPORTED = 4, /* shape outside what the kernels were built for */
};

/* kernel_type, covertreex/kernels.py:59-132 and src/metric.rs:44-57 */
enum { CX_KERNEL_RBF = 0, CX_KERNEL_MATERN52 = 1 };

/* batch_order, src/lib.rs:1404-1411 */
enum { CX_ORDER_HILBERT = 0 /* 21-bit Morton on D == 3, identity otherwise */, CX_ORDER_NATURAL = 1 };

/* Query precision modes for cx_tree_knn_residual*. */
enum {
    /* The reference's f32 search semantics (single_residual_knn_query,
     * src/algo.rs:612-1242, empty budget schedule) with its f32 distance
     * arithmetic: one wavefront per query. */
    CX_KNN_REFERENCE_F32 = 0,
    /* Batched mixed-precision engine: query tiles x candidate tiles on MFMA
     * with BF16 hi/lo V and a certified bound, surviving candidates rescored
     * in compensated FP32; neighbour sets are exact under the residual metric
     * evaluated on the stored inputs. */
    CX_KNN_MIXED = 1,
    /* The reference's default search: CX_KNN_REFERENCE_F32 with the budget
     * ladder [32, 64, 96], up 0.015, down 0.002 (src/algo.rs:648-683,
     * 1191-1219) -- approximate, as the reference's default is. */
    CX_KNN_REFERENCE_F32_BUDGETED = 2,
    /* Euclidean trees: exact k nearest (tiled scan, FP32 differences). */
    CX_KNN_EXACT = 1,
};

typedef struct cx_tree cx_tree;

/* Thread-local description of the last failure on this thread. */
const char *cx_last_error(void);

/* Number of visible HIP devices (0 when none). */
cx_status cx_device_count(int32_t *out);

/* Replaces build_pcct2_residual_tree (src/lib.rs:1373-1564), including the
 * si_cache pass (src/lib.rs:1548-1561) and the engine's
 * compute_subtree_bounds_py call (covertreex/engine.py:1038-1043).
 *   V: n x m float32 row-major (Nystrom factor), p: n residual variances,
 *   coords: n x d float32, ls: n_ls lengthscales (1 or d values),
 *   chunk: batch size (reference default 512), order_mode: CX_ORDER_*,
 *   grid_scale: grid_whiten_scale (reference default 1.0).
 * The tree is built on the device; node i holds dataset row order[i]. */
cx_status cx_tree_build_residual(const float *V, const float *p, const float *coords,
                                 int64_t n, int32_t m, int32_t d, float var,
                                 const float *ls, int32_t n_ls, int32_t kernel_type,
                                 int32_t chunk, int32_t order_mode, float grid_scale,
                                 cx_tree **out);

/* Replaces CoverTreeWrapper.knn_query_residual (src/lib.rs:369-507) +
 * to_py_arrays (src/lib.rs:907-938).  queries: dataset indices (host).
 * Outputs (host): out_idx nq x k dataset indices ascending by distance,
 * padded -1; out_dist nq x k padded FLT_MAX; out_err (optional, may be NULL)
 * nq x k certified bound on |d - d_exact| for CX_KNN_MIXED (the f32 error
 * of the reference arithmetic is not bounded: NaN in CX_KNN_REFERENCE_F32).
 * Out-of-range query indices map to node 0 as the reference does
 * (src/lib.rs:414-418).
 * Limits (CX_ERR_INVALID with a message when exceeded; the reference has none):
 * 1 <= k <= 256 in both modes, d <= 16 in CX_KNN_MIXED. */
cx_status cx_tree_knn_residual(cx_tree *t, const int64_t *queries, int64_t nq, int32_t k,
                               int32_t predecessor, int32_t mode, int64_t *out_idx,
                               float *out_dist, float *out_err);

/* Same with device-resident queries and outputs (HBM pointers), enqueued on
 * the tree's own stream and synchronised before returning.  That stream does
 * not wait for any other stream: the caller must have finished writing
 * d_queries (e.g. synchronised the stream that produced them) before the
 * call; the outputs are complete when it returns. */
cx_status cx_tree_knn_residual_device(cx_tree *t, const int64_t *d_queries, int64_t nq,
                                      int32_t k, int32_t predecessor, int32_t mode,
                                      int64_t *d_out_idx, float *d_out_dist,
                                      float *d_out_err);

/* Euclidean tree over n points of dimension d: chunked batch insert in
 * natural order with no grid edges.  Replaces CoverTreeWrapper(...).insert
 * (src/lib.rs:207-223 -> batch_insert, src/algo/batch/mod.rs:402-409), which
 * the reference's API path runs on one unchunked batch; chunk = n <= 1024
 * gives that semantics. */
cx_status cx_tree_build_euclidean(const float *points, int64_t n, int32_t d, int32_t chunk,
                                  cx_tree **out);

/* Replaces CoverTreeWrapper.knn_query (src/lib.rs:331-357 -> batch_knn_query,
 * src/algo.rs:179-323).  queries: nq x d float32 points (host); neighbours are
 * tree node indices (= input row indices, natural order).  mode:
 * CX_KNN_REFERENCE_F32 (the reference's best-first search) or CX_KNN_EXACT.
 * Predecessor mode admits node j for query position i when j < i. */
cx_status cx_tree_knn_euclidean(cx_tree *t, const float *queries, int64_t nq, int32_t k,
                                int32_t predecessor, int32_t mode, int64_t *out_idx,
                                float *out_dist);

/* Tree shape.  Replaces len(node_to_dataset) / get_min_level / get_max_level
 * (src/lib.rs:889-904). */
cx_status cx_tree_info(const cx_tree *t, int64_t *n, int32_t *m, int32_t *d,
                       int32_t *min_level, int32_t *max_level);

/* Host copies of the device tree arrays.  Replace get_parents, get_children,
 * get_next_node, get_levels (src/lib.rs:825-887) and the node_to_dataset list
 * returned by build_pcct2_residual_tree (src/lib.rs:1543). */
cx_status cx_tree_get_parents(const cx_tree *t, int64_t *out);
cx_status cx_tree_get_children(const cx_tree *t, int64_t *out);
cx_status cx_tree_get_next(const cx_tree *t, int64_t *out);
cx_status cx_tree_get_levels(const cx_tree *t, int32_t *out);
cx_status cx_tree_get_si_cache(const cx_tree *t, float *out);
cx_status cx_tree_get_order(const cx_tree *t, int64_t *out);
cx_status cx_tree_get_subtree_bounds(const cx_tree *t, int64_t *min_out, int64_t *max_out);

/* Replaces compute_subtree_bounds_py (src/lib.rs:991-1006 -> src/tree.rs:144-185):
 * min/max node_to_dataset value over each node's subtree, on the device. */
cx_status cx_subtree_bounds(const int64_t *parents, const int64_t *node_to_dataset, int64_t n,
                            int64_t *min_out, int64_t *max_out);

/* Replaces hilbert_order (src/lib.rs:956-962 -> src/pcct.rs:27-67): the 21-bit
 * Morton permutation for d == 3, identity otherwise, computed on the device. */
cx_status cx_morton_order(const float *coords, int64_t n, int32_t d, int64_t *order_out);

/* Replaces _hilbert_permutation (covertreex/algo/order/strategy.py:38-60),
 * the python-path batch order: bits = max(1, min(16, 64 / d)) per axis,
 * round-half-even quantisation in FP64, Skilling transpose, codes sorted
 * stably (ties by index).  points: n x d float64; d <= 64. */
cx_status cx_hilbert_order(const double *points, int64_t n, int32_t d, int64_t *order_out);

/* No counterpart in the reference: cx_hilbert_order for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  d_points (n x d float64) and d_order_out (n int64) are
 * device pointers; the work is enqueued on `stream` (a hipStream_t, NULL for
 * the default stream) and the order is complete when that stream's work is.
 * It allocates nothing: its work buffers are one fixed device array per GPU
 * (about 14 MiB, held until the process ends), so n is at most 262144
 * (CX_ERR_INVALID beyond) and calls that overlap on one GPU race. */
cx_status cx_hilbert_order_noalloc(const double *d_points, int64_t n, int32_t d,
                                   int64_t *d_order_out, void *stream);

/* Replaces run_mis_numba / _run_mis_numba_impl (covertreex/algo/mis.py:55-139,
 * covertreex/algo/_mis_numba.py:64-131): greedy in index order when the
 * max degree is <= 8, else Luby rounds with xorshift priorities (32-bit
 * arithmetic, ties to the lower index, first active node when a round has no
 * winner).  CSR graph (indptr n+1, indices), selected_out: n bytes (0/1). */
cx_status cx_luby_mis(const int64_t *indptr, const int64_t *indices, int64_t n, uint32_t seed,
                      uint8_t *selected_out, int64_t *iterations_out);

/* Replaces build_residual_backend's numeric core
 * (covertreex/metrics/residual/host_backend.py:201-240): FP64 kernel
 * matrices, Cholesky of K_mm + jitter*var*I, V = (L^-1 K_xm^T)^T and
 * p = max(var - |V|^2, 1e-9), all on the device, stored as FP32.
 *   points: n x d float64, inducing_idx: m sorted dataset indices,
 *   ls: n_ls lengthscales (1 or d), v_out: n x m float32, p_out: n float32.
 * Returns CX_ERR_INVALID when K_mm + jitter is not positive definite. */
cx_status cx_nystrom_residual(const double *points, int64_t n, int32_t d,
                              const int64_t *inducing_idx, int32_t m, double var,
                              const double *ls, int32_t n_ls, int32_t kernel_type, double jitter,
                              float *v_out, float *p_out);

/* Per-phase device timing of the last kNN call on this tree.  When enabled,
 * HIP events are recorded on the tree's stream at phase boundaries.
 * cx_tree_last_profile writes up to cap durations (ms):
 *   CX_KNN_MIXED:         [prep, query sort, screen pass 1, tau, filter,
 *                          rescore, overflowed queries past the rescore's end
 *                          (they run on a second stream beside it)]
 *   CX_KNN_REFERENCE_F32: [traversal]                                         */
/* Candidates per screen segment of the mixed engine (CX_KNN_MIXED): each
 * workgroup of the screen streams one segment's V rows against 64 queries.
 * 0 restores the default (the power of two of rows, at most 4096, whose
 * BF16 hi/lo fragments fill at most 2 MiB, halved while a call's grid would
 * hold fewer than 512 workgroups); other values must be multiples of 256 in
 * [256, 16384].  Results do not depend on it.  No counterpart in the
 * reference: a tile shape for the tuner (covertreex_amd/tuner.py).  The
 * getter reports the segment length the last mixed query used (0 before the
 * first). */
cx_status cx_tree_set_screen_segment(cx_tree *t, int32_t seg);
cx_status cx_tree_get_screen_segment(cx_tree *t, int32_t *seg_out);
cx_status cx_tree_set_profiling(cx_tree *t, int32_t enable);
cx_status cx_tree_last_profile(cx_tree *t, float *ms_out, int32_t cap, int32_t *n_out);

/* Incremental insert on a device-resident residual tree.  Replaces
 * CoverTreeWrapper.insert_residual (src/lib.rs:254-300): the n_new rows (V,
 * p, coords, host float32, dataset order) become dataset rows n .. n+n_new-1
 * and tree nodes n .. n+n_new-1 in their given order, batch-inserted in
 * chunks of `chunk` with the build's grid scale, kernel and lengthscales.
 * si_cache, subtree bounds and the child lists are recomputed. */
cx_status cx_tree_insert_residual(cx_tree *t, const float *V, const float *p,
                                  const float *coords, int64_t n_new, int32_t chunk);

/* Incremental insert on a Euclidean tree.  Replaces CoverTreeWrapper.insert
 * (src/lib.rs:207-223) applied to a grown tree: the n_new points become nodes
 * n .. n+n_new-1, batch-inserted in chunks of `chunk`. */
cx_status cx_tree_insert_euclidean(cx_tree *t, const float *points, int64_t n_new,
                                   int32_t chunk);

/* Point deletion on a device-resident tree.  Replaces batch_delete
 * (covertreex/algo/batch_delete.py:628-689, computation 105-525): the points
 * and every descendant of them are detached, the rest of the tree keeps its
 * links, and the detached points (not the removed ones) are inserted again
 * in groups of equal max(level, 0), lowest first, each in batches of the
 * build's chunk (this library's batch insert, not the reference's python
 * planner).  indices: dataset rows for a residual tree, node indices (as the
 * Euclidean kNN returns them) for a Euclidean tree; distinct, in range, fewer
 * than the tree's size.  Remaining dataset rows are renumbered in order (the
 * removed rows close up); tree order becomes [kept nodes in their order,
 * re-inserted nodes by group]. */
cx_status cx_tree_delete(cx_tree *t, const int64_t *indices, int64_t n_remove);

/* Per-batch build telemetry: the counterpart of the reference's
 * BatchInsertTelemetry / record_rust_batch payload (src/lib.rs:1274-1371),
 * for every batch of the build and of later inserts, CX_BATCH_FIELDS int64
 * per batch:                                                               */
#define CX_BATCH_FIELDS 16
#define CX_BF_START 0      /* first tree row of the batch */
#define CX_BF_SIZE 1       /* rows in the batch */
#define CX_BF_LEADERS 2    /* MIS-selected leaders, all levels (selected) */
#define CX_BF_SURVIVORS 3  /* rows attached at the bottom level */
#define CX_BF_ORPHANS 4    /* far non-leaders left as roots */
#define CX_BF_FAR_LEVELS 5 /* levels with a non-empty far set */
#define CX_BF_SCAN_EVALS 6 /* distance evaluations: near/far scan (traversal) */
#define CX_BF_MIS_EVALS 7  /* distance evaluations: conflict graph */
#define CX_BF_NL_EVALS 8   /* distance evaluations: non-leader attachment */
#define CX_BF_GRID_EDGES 9 /* conflicts found through the grid */
#define CX_BF_DIST_EDGES 10 /* conflicts found through d^2 <= r^2 */
#define CX_BF_NS_SCAN 11   /* device wall time (ns): near/far scan */
#define CX_BF_NS_MIS 12    /* device wall time (ns): near/far split, MIS, leaders */
#define CX_BF_NS_NL 13     /* device wall time (ns): non-leader attachment */
#define CX_BF_NS_TOTAL 14  /* device wall time (ns): whole batch */
#define CX_BF_NS_CONFLICT 15 /* device wall time (ns): conflict pass (far pairs) */
/* Writes min(cap, batches) rows of CX_BATCH_FIELDS; *n_batches = batches. */
cx_status cx_tree_build_telemetry(const cx_tree *t, int64_t *out, int64_t cap_batches,
                                  int64_t *n_batches);

/* Conflict records of batch inserts, for the reference's per-batch payload
 * (BatchInsertTelemetry's parents/levels/selected/dominated/conflict_indptr/
 * conflict_indices, src/algo/batch/mod.rs:802-956, sent by emit_rust_batch,
 * src/lib.rs:1274-1371, when a log writer is given).  Off by default:
 * cx_set_conflict_log(1) turns them on for every later build and insert in
 * the process.  cx_tree_conflict_log copies a tree's record (int32 words)
 * into out (up to cap words; *n_words = its length).  Entries, in insert
 * order:
 *   1, batch, level, nf, far node ids [nf], leader bits [W], covered
 *   non-leader bits [W], distance-conflict bits [nf][W] (d^2 <= r^2, j > i),
 *   grid-conflict bits [nf][W]   (one per level with a far set; W =
 *   ceil(nf / 32), bit j of a row in word j / 32)
 *   2, batch, count, bottom survivor node ids [count]   (one per batch)
 * COVERTREEX_HIP_CONFLICT_LOG_WORDS sets how many words one insert call may
 * record (32M by default; more fails with CX_ERR_INVALID). */
cx_status cx_set_conflict_log(int32_t enable);
cx_status cx_tree_conflict_log(const cx_tree *t, int32_t *out, int64_t cap, int64_t *n_words);

/* Query telemetry.  Replaces the COVERTREEX_RUST_QUERY_TELEMETRY switch and
 * CoverTreeWrapper.last_query_telemetry / clear_last_query_telemetry
 * (src/lib.rs:357-367, ResidualQueryTelemetry src/telemetry.rs:1-19).  When
 * enabled, each kNN call records CX_QT_SIZE int64 counters:                */
#define CX_QT_SIZE 160
#define CX_QT_LEVELS 64
#define CX_QT_MODE 0            /* the CX_KNN_* mode of the call */
#define CX_QT_QUERIES 1
#define CX_QT_DISTANCE_EVALS 2  /* residual distances evaluated (mixed: pairs screened) */
#define CX_QT_PRUNES_LB 3       /* prunes_lower_bound */
#define CX_QT_PRUNES_LB_CHUNKS 4 /* prunes_lower_bound_chunks */
#define CX_QT_PRED_FILTERED 5   /* predecessor_filtered */
#define CX_QT_SUBTREES_PRUNED 6 /* subtrees_pruned */
#define CX_QT_BUDGET_ESC 7      /* budget_escalations */
#define CX_QT_BUDGET_EARLY 8    /* budget_early_terminate */
#define CX_QT_LISTED 9          /* mixed: candidates listed for rescoring */
#define CX_QT_HARD_QUERIES 10   /* mixed: queries scored exhaustively (over all candidates) */
#define CX_QT_HARD_PAIRS 11     /* mixed: pairs scored for those queries */
#define CX_QT_MAX_LIST 12       /* mixed: longest list */
#define CX_QT_N_LEVELS 13       /* levels recorded below (<= CX_QT_LEVELS) */
#define CX_QT_FRONTIER 14       /* [CX_QT_LEVELS] frontier size per level, summed */
#define CX_QT_EXPANDED 78       /* [CX_QT_LEVELS] children gathered per level */
#define CX_QT_BLOCKS 142        /* [3] tiles of 16, 32, 64 candidates */
#define CX_QT_HARD_LISTED 145   /* mixed: pairs the last screen kept for those queries */
#define CX_QT_RESCREENED 146   /* mixed: overflowed queries given the exact-product screen */
cx_status cx_tree_set_query_telemetry(cx_tree *t, int32_t enable);
/* Cover-tree invariants (the properties batch insert establishes; the
 * oracle's tests check the same on the CPU, tests/test_oracle_tree.py):
 * writes CX_INV_FIELDS int64 counts.  Covering: d(node, parent)^2 <=
 * 4^(level+1); nesting: a node's parent was in the tree when the node was
 * inserted (same or earlier batch) and the parent chain reaches a root;
 * separation: nodes one
 * batch inserted at one level (strictly inside the level range) pairwise
 * d^2 > 4^level and, with grid conflicts, not grid neighbours; forest: child
 * lists and parent pointers agree, dropped nodes have no links.  Separation
 * covers the batches inserted since the build or the last delete.  No
 * counterpart entry point in the reference (its construction guarantees
 * these, src/algo/batch/mod.rs:447-800). */
#define CX_INV_FIELDS 8
#define CX_INV_COVERING 0
#define CX_INV_NESTING 1
#define CX_INV_SEPARATION 2
#define CX_INV_FOREST 3
#define CX_INV_NODES 4      /* linked nodes checked */
#define CX_INV_PAIRS 5      /* same-level pairs of one batch checked */
#define CX_INV_BATCHES 6    /* separation groups */
cx_status cx_tree_check_invariants(cx_tree *t, int64_t *out);

/* Test support, no counterpart in the reference: overwrite one node's parent
 * pointer and level on the device, leaving every child list as it is, so a
 * test can hand cx_tree_check_invariants a tree that breaks the covering,
 * nesting, separation or forest invariants.  Queries on such a tree are
 * meaningless. */
cx_status cx_tree_debug_set_node(cx_tree *t, int64_t node, int64_t parent, int32_t level);

/* The exp of the build's and the reference form's kernel values, which
 * replaces the platform expf the reference's metric calls
 * (src/metric.rs:183,196,458,464,534,541): binary32 exp correctly rounded,
 * computed in FP32 arithmetic.  Evaluates it on the device for the count
 * floats whose bit patterns follow first_bits (first_bits + count <= 2^32)
 * and writes the results to host memory out.  For checking the routine
 * against an oracle. */
cx_status cx_expf_cr_range(uint32_t first_bits, int64_t count, float *out);

/* *n_out = CX_QT_SIZE when the last call recorded telemetry, else 0. */
cx_status cx_tree_last_query_telemetry(const cx_tree *t, int64_t *out, int32_t cap,
                                       int32_t *n_out);
cx_status cx_tree_clear_last_query_telemetry(cx_tree *t);

void cx_tree_free(cx_tree *t);

#ifdef __cplusplus
}
#endif
#endif
