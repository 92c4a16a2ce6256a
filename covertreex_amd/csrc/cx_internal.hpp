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

// Owning device buffer.
template <typename T> struct DevBuf {
    T *ptr = nullptr;
    size_t count = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), count(o.count)
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
            o.ptr = nullptr;
            o.count = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n)
    {
        release();
        count = n;
        if (n)
            CX_HIP(hipMalloc((void **)&ptr, n * sizeof(T) + 256));
    }
    void release()
    {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    size_t bytes() const { return count * sizeof(T); }
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
    DevBuf<uint16_t> vhi, vlo;     // n x mp bf16, zero padded to mp
    int32_t mp = 0;                // m padded to a multiple of 32
    DevBuf<float> vnorm;           // |V_i| rounded up, dataset order
    DevBuf<float> x_ds, isp_ds;    // raw coords and 1/sqrt(p) in dataset order
    DevBuf<uint8_t> codes;         // one byte per screened pair (a launch group)
    DevBuf<uint8_t> part;          // per-segment top-k lists of overflowed queries

    // reusable query workspace
    DevBuf<uint8_t> work;
    // pinned host word for small device-to-host reads (hipHostMalloc)
    int32_t *host_word = nullptr;
    hipEvent_t sync_event = nullptr; // marks a small device-to-host read
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
void morton_order_device(const float *coords_host, int64_t n, int32_t d, int64_t *order_host);
void subtree_bounds_device(const int64_t *parents_host, const int64_t *n2d_host, int64_t n,
                           int64_t *min_host, int64_t *max_host);

// knn_ref.hip
void knn_reference_f32(TreeState &t, const int64_t *d_queries, int64_t nq, int32_t k,
                       int32_t predecessor, int32_t budgeted, int64_t *d_idx, float *d_dist,
                       float *d_err);

// knn_mixed.hip
void knn_mixed(TreeState &t, const int64_t *d_queries, int64_t nq, int32_t k,
               int32_t predecessor, int64_t *d_idx, float *d_dist, float *d_err);

// knn_euclid.hip
void knn_euclidean(TreeState &t, const float *d_queries, int64_t nq, int32_t k, int32_t pred,
                   int32_t mode, int64_t *d_idx, float *d_dist);

// graph.hip
void hilbert_order_device(const double *points_host, int64_t n, int32_t d, int64_t *order_host);
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
