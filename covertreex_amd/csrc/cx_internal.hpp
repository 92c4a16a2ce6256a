// Internal host-side definitions shared by the translation units of
// libcovertreex_hip.so.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/covertreex_hip.h"

namespace cx {

// Error carried from deep inside the library to the C-ABI wrapper.
struct Error : std::runtime_error {
    cx_status code;
    Error(cx_status c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string &msg);

#define CX_HIP(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw ::cx::Error(CX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define CX_REQUIRE(cond, msg)                                   \
    do {                                                        \
        if (!(cond))                                            \
            throw ::cx::Error(CX_ERR_INVALID, std::string(msg)); \
    } while (0)

// The stream that orders this thread's DevBuf allocations while a call works
// on a tree (StreamScope): buffers come from the device's stream-ordered pool
// (hipMallocAsync / hipFreeAsync on that stream), so freeing one never waits
// for other trees' work on the device; with no scope, plain hipMalloc/hipFree.
inline thread_local hipStream_t tls_alloc_stream = nullptr;

struct StreamScope {
    hipStream_t prev;
    explicit StreamScope(hipStream_t s) : prev(tls_alloc_stream) { tls_alloc_stream = s; }
    ~StreamScope() { tls_alloc_stream = prev; }
    StreamScope(const StreamScope &) = delete;
    StreamScope &operator=(const StreamScope &) = delete;
};

// Keep freed pool memory cached in the process (no release threshold), once
// per device.
void ensure_stream_pool(int device);

// Stream-ordered buffers of kBigBlock bytes or more come from a process-wide
// cache of device blocks instead of the stream-ordered pool: a block is
// cached once the freeing stream has drained, and the next allocation of
// about its size on the same device (any stream) takes it.  Trees of a
// serving loop allocate and free the same large transient buffers (pair
// codes, candidate lists) request after request; a fresh allocation of that
// size from the stream-ordered pool has blocked the host for up to a second,
// a cached block costs a lookup.  Blocks stay cached for the process, up to
// 48 GiB of free blocks (the largest go first beyond that, and all of a
// device's when an allocation would otherwise fail).
constexpr size_t kBigBlock = (size_t)32 << 20;
void *big_alloc(size_t bytes, hipStream_t s);
// returns false when p is not a cached-block allocation
bool big_free(void *p, hipStream_t s);

// Process-wide pools of the per-tree runtime objects whose creation or
// destruction can wait on other work on the device (pinned host memory, and
// for symmetry streams and events): a tree takes them at first use and gives
// them back when it is destroyed.
hipStream_t acquire_stream(int device);
void release_stream(int device, hipStream_t s);
int32_t *acquire_host_word();
void release_host_word(int32_t *w);
hipEvent_t acquire_event();
void release_event(hipEvent_t e);

// Device-to-host copies into pageable memory through pinned staging chunks
// (a pool of two-chunk stagers, one per concurrent caller): the DMA of one
// chunk runs while the other is copied out on the host.  Copies below a few
// MiB go straight to hipMemcpyAsync.  Returns with the data in dst (s
// drained up to the copy).  (Uploads stay on HIP's own pageable path, which
// measured faster here than staging them the same way.)
void copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s);

// Owning device buffer.
template <typename T> struct DevBuf {
    T *ptr = nullptr;
    size_t count = 0;
    hipStream_t st = nullptr; // the stream ordering its allocation and free
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), count(o.count), st(o.st)
    {
        o.ptr = nullptr;
        o.count = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            ptr = o.ptr;
            count = o.count;
            st = o.st;
            o.ptr = nullptr;
            o.count = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    static T *raw_alloc(size_t n, hipStream_t s)
    {
        T *q = nullptr;
        const size_t bytes = n * sizeof(T) + 256;
        if (s && bytes >= kBigBlock)
            q = (T *)big_alloc(bytes, s);
        else if (s)
            CX_HIP(hipMallocAsync((void **)&q, bytes, s));
        else
            CX_HIP(hipMalloc((void **)&q, bytes));
        return q;
    }
    void alloc(size_t n)
    {
        release();
        count = n;
        st = tls_alloc_stream;
        if (n)
            ptr = raw_alloc(n, st);
    }
    void release()
    {
        if (ptr) {
            if (st) {
                if (!big_free(ptr, st))
                    (void)hipFreeAsync(ptr, st);
            } else {
                (void)hipFree(ptr);
            }
        }
        ptr = nullptr;
        count = 0;
    }
    size_t bytes() const { return count * sizeof(T); }
    // reallocate to n elements keeping the first min(count, n), ordered on
    // stream s (the old buffer's free follows the copy)
    void grow_keep(size_t n, hipStream_t s)
    {
        T *np = n ? raw_alloc(n, s) : nullptr;
        const size_t keep = count < n ? count : n;
        if (st && st != s)
            CX_HIP(hipStreamSynchronize(st)); // its work on the old buffer
        if (keep)
            CX_HIP(hipMemcpyAsync(np, ptr, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (!st)
            CX_HIP(hipStreamSynchronize(s));
        else
            st = s; // the old buffer is freed after the copy
        release();
        ptr = np;
        count = n;
        st = s;
    }
};

// Per-batch build telemetry (int64 fields per batch in TreeState::batch_stats)
constexpr int kBatchFields = 16;
enum BatchField {
    kBfStart = 0,     // first row (tree order) of the batch
    kBfSize = 1,      // rows in the batch
    kBfLeaders = 2,   // MIS-selected leaders over all levels
    kBfSurvivors = 3, // rows attached at the bottom level
    kBfOrphans = 4,   // far non-leaders dropped (left as roots)
    kBfFarLevels = 5, // levels with a non-empty far set
    kBfScanEvals = 6, // distance evaluations: near/far scan
    kBfMisEvals = 7,  // distance evaluations: MIS conflicts
    kBfNlEvals = 8,   // distance evaluations: non-leader attachment
    kBfGridEdges = 9, // conflicts found through the grid
    kBfDistEdges = 10, // conflicts found through d^2 <= r^2
    kBfTicksScan = 11, // wall-clock ticks: near/far scan
    kBfTicksMis = 12,  // wall-clock ticks: near/far split, MIS and leader linking
    kBfTicksNl = 13,   // wall-clock ticks: non-leader attachment
    kBfTicksTotal = 14, // wall-clock ticks: whole batch
    kBfTicksConflict = 15 // wall-clock ticks: conflict pass (far pairs' distances, grid)
};

// Device-resident tree and metric data (all arrays in tree order).
struct TreeState {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;

    int64_t n = 0;
    int32_t m = 0, d = 0;
    int32_t metric = 0; // 0 residual correlation, 1 Euclidean
    float var = 1.0f;
    int32_t kernel = 0;
    int32_t min_level = -20, max_level = 20;
    std::vector<float> inv_ls, ls;

    // metric (tree order)
    DevBuf<float> v, scaled, raw, norms, p;
    // tree
    DevBuf<int64_t> order, inv_order, parents, children, next;
    DevBuf<int32_t> levels, depth;
    DevBuf<float> si;
    DevBuf<int64_t> smin, smax;
    // roots (parent == -1) in index order; internal nodes (roots with children
    // are the only frontier entries that expand)
    DevBuf<int32_t> roots;
    int64_t n_roots = 0;
    int32_t max_depth = 0;
    // children in linked-list (prepend) order as CSR: child_list[child_off[i] ..
    // child_off[i+1]) are node i's children
    DevBuf<int32_t> child_off, child_list;
    int64_t n_internal = 0;

    // mixed-precision engine data, built lazily on first CX_KNN_MIXED query
    bool mixed_ready = false;
    DevBuf<uint16_t> vsw;          // npad x mp bf16 hi and lo, MFMA-fragment layout
    int32_t mp = 0;                // m padded to a multiple of 32
    int32_t mps = 0;               // m padded to the screen's K chunk (64)
    int32_t screen_seg = 0;        // candidates per screen segment (0: default)
    int32_t last_seg = 0;          // the segment length of the last mixed query
    DevBuf<float> vnorm;           // |V_i| rounded up, dataset order
    DevBuf<float> x_ds, isp_ds;    // raw coords and 1/sqrt(p) in dataset order
    DevBuf<uint8_t> codes;         // one byte per screened pair (a launch group)
    DevBuf<uint8_t> part;          // per-segment top-k lists of overflowed queries

    // exact Euclidean kNN: the points in Hilbert order cut into blocks of 64
    // with their bounding boxes (built on the first exact query)
    bool eu_ready = false;
    int64_t eu_nb = 0;
    DevBuf<float> eu_pts;       // eu_nb * 64 x d, Hilbert order
    DevBuf<int32_t> eu_idx;     // tree index of each sorted row, -1 past n
    DevBuf<float> eu_lo, eu_hi; // block then superblock boxes, axis-major
    DevBuf<int32_t> eu_min;     // smallest index per block then per superblock
    DevBuf<uint64_t> eu_codes;  // sorted Hilbert codes (the seed search)
    DevBuf<double> eu_bounds;   // per-axis quantisation bounds: lo[d], hi[d]

    // query telemetry: when enabled, kNN kernels add their counters into
    // telem (CX_QT_SIZE int64, zeroed per call; null when disabled) and the
    // entry point copies them to last_telem
    bool query_telemetry = false;
    DevBuf<int64_t> telem_buf;
    int64_t *telem = nullptr;
    std::vector<int64_t> last_telem;
    int32_t telem_mode = 0;
    int64_t telem_nq = 0;

    // build parameters kept for inserts, and the per-batch telemetry of every
    // insert so far (kBatchFields int64 per batch)
    int32_t chunk = 512;
    float grid_scale = 1.0f;
    std::vector<int64_t> batch_stats;
    // conflict records of the inserts made while they were turned on
    // (set_conflict_log; entry format at ConflictLog in build.hip)
    std::vector<int32_t> conflict_log;
    // (start, size) of every batch inserted since the build or the last delete
    // (tree positions): the invariant checker's separation groups
    std::vector<int64_t> live_batches;

    // reusable query workspace
    DevBuf<uint8_t> work;
    // pinned host word for small device-to-host reads (hipHostMalloc)
    int32_t *host_word = nullptr, *host_word2 = nullptr;
    hipEvent_t sync_event = nullptr, sync_event2 = nullptr; // small device-to-host reads
    // exact-product screen of overflowed queries: int8 slices of V and row
    // exponents (dataset order, built on first use), bounds workspace
    DevBuf<int8_t> vsl;
    DevBuf<int32_t> vexp;
    bool slices_ready = false;
    DevBuf<uint8_t> hs_buf;
    DevBuf<int32_t> list_big; // second listing of still-overflowing queries
    // second stream of the mixed kNN (overflowed queries beside the rescore)
    hipStream_t aux_stream = nullptr;
    hipEvent_t join_event = nullptr;
    int32_t last_hard = 0; // overflowed queries of the last mixed kNN
    float p_min = 1.0f;    // smallest p_diag, host-side at build
    double xs_max = 0.0;   // largest |coordinate / lengthscale|, host-side at build

    // optional per-phase timing of the last query: events recorded on the
    // tree's stream at phase boundaries (see cx_tree_last_profile)
    static constexpr int kMaxMarks = 8;
    bool profiling = false;
    hipEvent_t marks[kMaxMarks] = {};
    int n_marks = 0;
    int last_mode = -1;

    ~TreeState();
};

// build.hip
void build_residual_tree(TreeState &t, const float *V, const float *p, const float *coords,
                         int32_t chunk, int32_t order_mode, float grid_scale);
void build_euclidean_tree(TreeState &t, const float *points, int32_t chunk);
void insert_residual_rows(TreeState &t, const float *V, const float *p, const float *coords,
                          int64_t n_new, int32_t chunk);
void insert_euclidean_rows(TreeState &t, const float *points, int64_t n_new, int32_t chunk);
void delete_rows(TreeState &t, const int64_t *remove, int64_t n_remove, bool by_node);
// expf_cr of count consecutive bit patterns, computed on the device, into
// host memory
void expf_cr_range(uint32_t first_bits, int64_t count, float *out);
// Cover-tree invariants of the tree (CX_INV_* counts into out).
void check_invariants(TreeState &t, int64_t *out);
void morton_order_device(const float *coords_host, int64_t n, int32_t d, int64_t *order_host);
void subtree_bounds_device(const int64_t *parents_host, const int64_t *n2d_host, int64_t n,
                           int64_t *min_host, int64_t *max_host);

// knn_ref.hip
void begin_query_telemetry(TreeState &t, int32_t mode, int64_t nq);
void end_query_telemetry(TreeState &t);
void knn_reference_f32(TreeState &t, const int64_t *d_queries, int64_t nq, int32_t k,
                       int32_t predecessor, int32_t budgeted, int64_t *d_idx, float *d_dist,
                       float *d_err);

// knn_mixed.hip
// Candidates per screen segment of the mixed engine for nq queries on this
// tree, and whether a requested value is acceptable (0 or a multiple of 256
// in [256, 16384]).
// Turns the batch inserts' conflict records on or off (process-wide).
void set_conflict_log(bool enable);
int32_t screen_segment(const TreeState &t, int64_t nq);
bool screen_segment_valid(int32_t seg);
void knn_mixed(TreeState &t, const int64_t *d_queries, int64_t nq, int32_t k,
               int32_t predecessor, int64_t *d_idx, float *d_dist, float *d_err);

// knn_euclid.hip
void knn_euclidean(TreeState &t, const float *d_queries, int64_t nq, int32_t k, int32_t pred,
                   int32_t mode, int64_t *d_idx, float *d_dist);

// graph.hip
void hilbert_order_device(const double *points_host, int64_t n, int32_t d, int64_t *order_host);
void hilbert_order_noalloc(const double *d_points, int64_t n, int32_t d, int64_t *d_order,
                           hipStream_t st);
// Hilbert codes of n FP32 points (sorted) and the stable sorting permutation;
// d_lo / d_hi receive the per-axis bounds used for quantisation
void hilbert_sort_points_f32(const float *d_pts, int64_t n, int32_t d, uint64_t *d_sorted_codes,
                             int64_t *d_perm, double *d_lo, double *d_hi, hipStream_t st);
int64_t luby_mis_device(const int64_t *indptr_host, const int64_t *indices_host, int64_t n,
                        uint32_t seed, uint8_t *sel_host);

// nystrom.hip
void nystrom_device(const double *pts_host, int64_t n, int32_t d, const int64_t *ind_host,
                    int32_t m, double var, const double *ls_host, int32_t n_ls, int32_t kernel,
                    double jitter, float *v_host, float *p_host);

uint8_t *workspace(TreeState &t, size_t bytes);

// Record phase boundary `i` on the tree's stream when profiling is enabled.
void prof_mark(TreeState &t, int i);

} // namespace cx

struct cx_tree {
    cx::TreeState s;
};
