// C-ABI entry points of libcovertreex_hip.so (declared in include/covertreex_hip.h).
#include <cmath>
#include <cstring>
#include <string>

#include "cx_internal.hpp"

namespace cx {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

TreeState::~TreeState()
{
    if (stream)
        (void)hipStreamSynchronize(stream);
    if (aux_stream)
        (void)hipStreamSynchronize(aux_stream);
    // DevBuf members release themselves; the streams go last
    v.release();
    scaled.release();
    raw.release();
    norms.release();
    p.release();
    order.release();
    inv_order.release();
    parents.release();
    children.release();
    next.release();
    levels.release();
    depth.release();
    si.release();
    smin.release();
    smax.release();
    roots.release();
    child_off.release();
    child_list.release();
    vsw.release();
    vnorm.release();
    x_ds.release();
    isp_ds.release();
    codes.release();
    part.release();
    vsl.release();
    vexp.release();
    hs_buf.release();
    list_big.release();
    work.release();
    telem_buf.release();
    eu_pts.release();
    eu_idx.release();
    eu_lo.release();
    eu_hi.release();
    eu_min.release();
    eu_codes.release();
    eu_bounds.release();
    if (host_word)
        release_host_word(host_word);
    if (host_word2)
        release_host_word(host_word2);
    if (sync_event)
        release_event(sync_event);
    if (sync_event2)
        release_event(sync_event2);
    if (join_event)
        release_event(join_event);
    for (int i = 0; i < kMaxMarks; i++)
        if (marks[i])
            (void)hipEventDestroy(marks[i]);
    if (aux_stream)
        release_stream(device, aux_stream);
    if (stream)
        release_stream(device, stream);
}

namespace {
std::mutex g_pool_mu;
std::vector<std::pair<int, hipStream_t>> g_streams;
std::vector<int32_t *> g_words;
std::vector<hipEvent_t> g_events;
} // namespace

hipStream_t acquire_stream(int device)
{
    {
        std::lock_guard<std::mutex> lock(g_pool_mu);
        for (size_t i = 0; i < g_streams.size(); i++)
            if (g_streams[i].first == device) {
                hipStream_t s = g_streams[i].second;
                g_streams.erase(g_streams.begin() + (ptrdiff_t)i);
                return s;
            }
    }
    hipStream_t s = nullptr;
    CX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void release_stream(int device, hipStream_t s)
{
    (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lock(g_pool_mu);
    g_streams.emplace_back(device, s);
}

int32_t *acquire_host_word()
{
    std::lock_guard<std::mutex> lock(g_pool_mu);
    if (g_words.empty()) {
        // one pinned page, cut into 64-byte words (kept for the process)
        char *page = nullptr;
        CX_HIP(hipHostMalloc((void **)&page, 4096, hipHostMallocDefault));
        for (int i = 0; i < 4096 / 64; i++)
            g_words.push_back((int32_t *)(page + 64 * i));
    }
    int32_t *w = g_words.back();
    g_words.pop_back();
    return w;
}

namespace {
constexpr size_t kStageChunk = (size_t)8 << 20;
constexpr size_t kStageMin = (size_t)4 << 20;
struct Stager {
    int device;
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false}; // a DMA recorded on ev[i] not yet waited for
};
std::mutex g_stage_mu;
std::vector<Stager *> g_stagers;

Stager *acquire_stager()
{
    int dev = 0;
    CX_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lock(g_stage_mu);
        for (size_t i = 0; i < g_stagers.size(); i++)
            if (g_stagers[i]->device == dev) {
                Stager *g = g_stagers[i];
                g_stagers.erase(g_stagers.begin() + (ptrdiff_t)i);
                return g;
            }
    }
    Stager *g = new Stager{dev};
    try {
        for (int i = 0; i < 2; i++) {
            CX_HIP(hipHostMalloc(&g->buf[i], kStageChunk, hipHostMallocPortable));
            CX_HIP(hipEventCreateWithFlags(&g->ev[i], hipEventDisableTiming));
        }
    } catch (...) {
        for (int i = 0; i < 2; i++) {
            if (g->buf[i])
                (void)hipHostFree(g->buf[i]);
            if (g->ev[i])
                (void)hipEventDestroy(g->ev[i]);
        }
        delete g;
        throw;
    }
    return g;
}

// Waits for the stager's DMAs (recorded on the caller's stream, which the
// caller still owns), so the next user never waits on a stale record.
struct StagerLease {
    Stager *g;
    explicit StagerLease(Stager *s) : g(s) {}
    ~StagerLease()
    {
        for (int i = 0; i < 2; i++)
            if (g->pending[i]) {
                (void)hipEventSynchronize(g->ev[i]);
                g->pending[i] = false;
            }
        std::lock_guard<std::mutex> lock(g_stage_mu);
        g_stagers.push_back(g);
    }
    StagerLease(const StagerLease &) = delete;
    StagerLease &operator=(const StagerLease &) = delete;
};
} // namespace

void copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    if (bytes < kStageMin) {
        CX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        CX_HIP(hipStreamSynchronize(s));
        return;
    }
    StagerLease lease(acquire_stager());
    Stager *g = lease.g;
    const size_t n = (bytes + kStageChunk - 1) / kStageChunk;
    auto issue = [&](size_t i) {
        const int slot = (int)(i & 1);
        const size_t off = i * kStageChunk;
        const size_t len = bytes - off < kStageChunk ? bytes - off : kStageChunk;
        CX_HIP(hipMemcpyAsync(g->buf[slot], (const char *)src + off, len, hipMemcpyDeviceToHost,
                              s));
        CX_HIP(hipEventRecord(g->ev[slot], s));
        g->pending[slot] = true;
    };
    issue(0);
    for (size_t i = 0; i < n; i++) {
        // chunk i + 1 comes down while chunk i is copied out (its buffer's
        // last chunk, i - 1, was copied out in the previous step)
        if (i + 1 < n)
            issue(i + 1);
        const int slot = (int)(i & 1);
        CX_HIP(hipEventSynchronize(g->ev[slot]));
        g->pending[slot] = false;
        const size_t off = i * kStageChunk;
        const size_t len = bytes - off < kStageChunk ? bytes - off : kStageChunk;
        std::memcpy((char *)dst + off, g->buf[slot], len);
    }
}

void release_host_word(int32_t *w)
{
    std::lock_guard<std::mutex> lock(g_pool_mu);
    g_words.push_back(w);
}

hipEvent_t acquire_event()
{
    {
        std::lock_guard<std::mutex> lock(g_pool_mu);
        if (!g_events.empty()) {
            hipEvent_t e = g_events.back();
            g_events.pop_back();
            return e;
        }
    }
    hipEvent_t e = nullptr;
    CX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

void release_event(hipEvent_t e)
{
    std::lock_guard<std::mutex> lock(g_pool_mu);
    g_events.push_back(e);
}

namespace {
struct BigBlock {
    void *p;
    size_t bytes;
    int device;
};
std::mutex g_big_mu;
std::vector<BigBlock> g_big_free;
std::vector<BigBlock> g_big_live;
size_t g_big_free_bytes = 0;
// cached free bytes kept per process; beyond it the largest free blocks are
// released
constexpr size_t kBigCacheCap = (size_t)48 << 30;
} // namespace

void *big_alloc(size_t bytes, hipStream_t)
{
    int dev = 0;
    CX_HIP(hipGetDevice(&dev));
    const size_t grain = (size_t)2 << 20;
    bytes = (bytes + grain - 1) & ~(grain - 1);
    // best fit among blocks up to half again (+64 MiB) the request
    const size_t slack = bytes / 2 + ((size_t)64 << 20);
    BigBlock b{};
    bool found = false;
    {
        std::lock_guard<std::mutex> lock(g_big_mu);
        size_t best = 0;
        for (size_t i = 0; i < g_big_free.size(); i++) {
            const BigBlock &c = g_big_free[i];
            if (c.device == dev && c.bytes >= bytes && c.bytes <= bytes + slack &&
                (!found || c.bytes < g_big_free[best].bytes)) {
                best = i;
                found = true;
            }
        }
        if (found) {
            b = g_big_free[best];
            g_big_free.erase(g_big_free.begin() + (ptrdiff_t)best);
            g_big_free_bytes -= b.bytes;
        }
    }
    if (!found) {
        b.bytes = bytes;
        b.device = dev;
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e == hipErrorOutOfMemory) {
            // give the cached blocks of this device back and try once more
            (void)hipGetLastError();
            std::vector<BigBlock> drop;
            {
                std::lock_guard<std::mutex> lock(g_big_mu);
                for (size_t i = 0; i < g_big_free.size();)
                    if (g_big_free[i].device == dev) {
                        drop.push_back(g_big_free[i]);
                        g_big_free_bytes -= g_big_free[i].bytes;
                        g_big_free.erase(g_big_free.begin() + (ptrdiff_t)i);
                    } else {
                        i++;
                    }
            }
            for (const BigBlock &c : drop)
                (void)hipFree(c.p);
            e = hipMalloc(&b.p, bytes);
        }
        CX_HIP(e);
    }
    std::lock_guard<std::mutex> lock(g_big_mu);
    g_big_live.push_back(b);
    return b.p;
}

bool big_free(void *p, hipStream_t s)
{
    BigBlock b{};
    {
        std::lock_guard<std::mutex> lock(g_big_mu);
        size_t i = 0;
        while (i < g_big_live.size() && g_big_live[i].p != p)
            i++;
        if (i == g_big_live.size())
            return false;
        b = g_big_live[i];
        g_big_live.erase(g_big_live.begin() + (ptrdiff_t)i);
    }
    // the freeing stream's work on the block ends before it is cached, so
    // any stream may take it at once (no cross-stream event, which a stream
    // captured into a graph elsewhere could not wait on); the frees come at
    // a tree's end or when a buffer grows, where the stream is drained or
    // about to be
    (void)hipStreamSynchronize(s);
    std::vector<BigBlock> drop;
    {
        std::lock_guard<std::mutex> lock(g_big_mu);
        g_big_free.push_back(b);
        g_big_free_bytes += b.bytes;
        while (g_big_free_bytes > kBigCacheCap) {
            size_t big = 0;
            for (size_t i = 1; i < g_big_free.size(); i++)
                if (g_big_free[i].bytes > g_big_free[big].bytes)
                    big = i;
            drop.push_back(g_big_free[big]);
            g_big_free_bytes -= g_big_free[big].bytes;
            g_big_free.erase(g_big_free.begin() + (ptrdiff_t)big);
        }
    }
    for (const BigBlock &c : drop)
        (void)hipFree(c.p);
    return true;
}

void ensure_stream_pool(int device)
{
    static std::mutex mu;
    static std::vector<int> done;
    std::lock_guard<std::mutex> lock(mu);
    for (int d : done)
        if (d == device)
            return;
    hipMemPool_t pool = nullptr;
    CX_HIP(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t thr = UINT64_MAX;
    CX_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
    done.push_back(device);
}

void prof_mark(TreeState &t, int i)
{
    if (!t.profiling || i < 0 || i >= TreeState::kMaxMarks)
        return;
    if (!t.marks[i])
        CX_HIP(hipEventCreate(&t.marks[i]));
    CX_HIP(hipEventRecord(t.marks[i], t.stream));
    if (i + 1 > t.n_marks)
        t.n_marks = i + 1;
}

uint8_t *workspace(TreeState &t, size_t bytes)
{
    if (t.work.count < bytes) {
        CX_HIP(hipStreamSynchronize(t.stream));
        t.work.alloc(bytes);
    }
    return t.work.ptr;
}

template <typename F> static cx_status guarded(F &&f)
{
    try {
        f();
        return CX_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return CX_ERR_INVALID;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return CX_ERR_INVALID;
    }
}

static void require_device()
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        throw Error(CX_ERR_NO_DEVICE, "no HIP device available");
}

static void copy_to_host(const TreeState &t, void *dst, const void *src, size_t bytes)
{
    CX_REQUIRE(dst != nullptr, "output pointer is NULL");
    if (bytes == 0)
        return;
    CX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, t.stream));
    CX_HIP(hipStreamSynchronize(t.stream));
}

// zeroed device counters for this call when query telemetry is on
void begin_query_telemetry(TreeState &t, int32_t mode, int64_t nq)
{
    t.last_telem.clear();
    t.telem = nullptr;
    if (!t.query_telemetry)
        return;
    if (!t.telem_buf.ptr)
        t.telem_buf.alloc(CX_QT_SIZE);
    CX_HIP(hipMemsetAsync(t.telem_buf.ptr, 0, sizeof(int64_t) * CX_QT_SIZE, t.stream));
    t.telem = t.telem_buf.ptr;
    t.telem_mode = mode;
    t.telem_nq = nq;
}

void end_query_telemetry(TreeState &t)
{
    if (!t.telem)
        return;
    t.last_telem.assign(CX_QT_SIZE, 0);
    CX_HIP(hipMemcpyAsync(t.last_telem.data(), t.telem, sizeof(int64_t) * CX_QT_SIZE,
                          hipMemcpyDeviceToHost, t.stream));
    CX_HIP(hipStreamSynchronize(t.stream));
    t.last_telem[CX_QT_MODE] = t.telem_mode;
    t.last_telem[CX_QT_QUERIES] = t.telem_nq;
    int64_t lv = 0;
    for (int i = 0; i < CX_QT_LEVELS; i++)
        if (t.last_telem[CX_QT_FRONTIER + i] || t.last_telem[CX_QT_EXPANDED + i])
            lv = i + 1;
    t.last_telem[CX_QT_N_LEVELS] = lv;
    t.telem = nullptr;
}


} // namespace cx

using namespace cx;

extern "C" {

const char *cx_last_error(void) { return g_last_error.c_str(); }

cx_status cx_device_count(int32_t *out)
{
    return guarded([&] {
        CX_REQUIRE(out != nullptr, "out is NULL");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess)
            count = 0;
        *out = count;
    });
}

cx_status cx_tree_build_residual(const float *V, const float *p, const float *coords, int64_t n,
                                 int32_t m, int32_t d, float var, const float *ls, int32_t n_ls,
                                 int32_t kernel_type, int32_t chunk, int32_t order_mode,
                                 float grid_scale, cx_tree **out)
{
    return guarded([&] {
        CX_REQUIRE(out != nullptr, "out is NULL");
        *out = nullptr;
        CX_REQUIRE(n >= 0 && n < (int64_t)1 << 31, "n must be in [0, 2^31)");
        CX_REQUIRE(m >= 1 && d >= 1 && d <= 64, "need m >= 1 and 1 <= d <= 64");
        CX_REQUIRE(n == 0 || (V && p && coords), "V, p and coords must be non-NULL");
        CX_REQUIRE(order_mode == CX_ORDER_HILBERT || order_mode == CX_ORDER_NATURAL,
                   "order_mode must be CX_ORDER_HILBERT or CX_ORDER_NATURAL");
        require_device();
        cx_tree *h = new cx_tree();
        try {
            TreeState &t = h->s;
            CX_HIP(hipGetDevice(&t.device));
            t.stream = acquire_stream(t.device);
            ensure_stream_pool(t.device);
            StreamScope ss_(t.stream);
            t.n = n;
            t.m = m;
            t.d = d;
            t.var = var;
            t.kernel = kernel_type == CX_KERNEL_MATERN52 ? 1 : 0; // src/metric.rs:50-56
            // smallest p (not NaN-aware: NaN compares false and stays out)
            t.p_min = 1.0f;
            for (int64_t i = 0; i < n; i++)
                t.p_min = p[i] < t.p_min ? p[i] : t.p_min;
            // inv_ls[i] = 1 / (|ls| < 1e-6 ? 1e-6 : ls), fallback ls[0] or 1
            // (ResidualMetric::new, src/metric.rs:93-101)
            const float fallback = (ls && n_ls > 0) ? ls[0] : 1.0f;
            t.inv_ls.resize(d);
            t.ls.resize(d);
            for (int32_t c = 0; c < d; c++) {
                float raw = (ls && c < n_ls) ? ls[c] : fallback;
                float safe = std::fabs(raw) < 1e-6f ? 1e-6f : raw;
                t.ls[c] = safe;
                t.inv_ls[c] = 1.0f / safe;
            }
            // largest |coordinate / lengthscale| (bounds the rounding of the
            // scaled coordinates the mixed-precision screen reads)
            double xs = 0.0;
            for (int64_t i = 0; i < n; i++)
                for (int32_t c = 0; c < d; c++) {
                    const double v = std::fabs((double)coords[i * d + c] * (double)t.inv_ls[c]);
                    xs = v > xs ? v : xs;
                }
            t.xs_max = xs;
            if (n > 0)
                build_residual_tree(t, V, p, coords, chunk, order_mode, grid_scale);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}


cx_status cx_tree_knn_residual_device(cx_tree *t, const int64_t *d_queries, int64_t nq, int32_t k,
                                      int32_t predecessor, int32_t mode, int64_t *d_out_idx,
                                      float *d_out_dist, float *d_out_err)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        CX_REQUIRE(k >= 1, "k must be >= 1");
        CX_REQUIRE(nq >= 0, "nq must be >= 0");
        CX_REQUIRE(nq == 0 || (d_queries && d_out_idx && d_out_dist), "NULL query/output pointer");
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        CX_REQUIRE(t->s.metric == 0, "residual kNN on a Euclidean tree");
        CX_HIP(hipSetDevice(t->s.device));
        if (nq == 0)
            return;
        t->s.n_marks = 0;
        t->s.last_mode = mode;
        begin_query_telemetry(t->s, mode, nq);
        if (mode == CX_KNN_REFERENCE_F32 || mode == CX_KNN_REFERENCE_F32_BUDGETED)
            knn_reference_f32(t->s, d_queries, nq, k, predecessor,
                              mode == CX_KNN_REFERENCE_F32_BUDGETED, d_out_idx, d_out_dist,
                              d_out_err);
        else if (mode == CX_KNN_MIXED)
            knn_mixed(t->s, d_queries, nq, k, predecessor, d_out_idx, d_out_dist, d_out_err);
        else
            throw Error(CX_ERR_INVALID, "unknown kNN mode");
        end_query_telemetry(t->s);
        CX_HIP(hipStreamSynchronize(t->s.stream));
    });
}

cx_status cx_tree_knn_residual(cx_tree *t, const int64_t *queries, int64_t nq, int32_t k,
                               int32_t predecessor, int32_t mode, int64_t *out_idx,
                               float *out_dist, float *out_err)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        CX_REQUIRE(k >= 1, "k must be >= 1");
        CX_REQUIRE(nq >= 0, "nq must be >= 0");
        CX_REQUIRE(nq == 0 || (queries && out_idx && out_dist), "NULL query/output pointer");
        if (nq == 0)
            return;
        StreamScope ss_(t->s.stream);
        DevBuf<int64_t> q(nq), oi((size_t)nq * k);
        DevBuf<float> od((size_t)nq * k), oe(out_err ? (size_t)nq * k : 0);
        hipStream_t st = t->s.stream;
        CX_HIP(hipMemcpyAsync(q.ptr, queries, sizeof(int64_t) * nq, hipMemcpyHostToDevice, st));
        cx_status s = cx_tree_knn_residual_device(t, q.ptr, nq, k, predecessor, mode, oi.ptr,
                                                  od.ptr, out_err ? oe.ptr : nullptr);
        if (s != CX_OK)
            throw Error(s, g_last_error);
        copy_d2h(out_idx, oi.ptr, sizeof(int64_t) * nq * k, st);
        copy_d2h(out_dist, od.ptr, sizeof(float) * nq * k, st);
        if (out_err)
            copy_d2h(out_err, oe.ptr, sizeof(float) * nq * k, st);
    });
}

cx_status cx_tree_build_euclidean(const float *points, int64_t n, int32_t d, int32_t chunk,
                                  cx_tree **out)
{
    return guarded([&] {
        CX_REQUIRE(out != nullptr, "out is NULL");
        *out = nullptr;
        CX_REQUIRE(n >= 0 && n < (int64_t)1 << 31, "n must be in [0, 2^31)");
        CX_REQUIRE(d >= 1 && d <= 64, "need 1 <= d <= 64");
        CX_REQUIRE(n == 0 || points, "points must be non-NULL");
        require_device();
        cx_tree *h = new cx_tree();
        try {
            TreeState &t = h->s;
            CX_HIP(hipGetDevice(&t.device));
            t.stream = acquire_stream(t.device);
            ensure_stream_pool(t.device);
            StreamScope ss_(t.stream);
            t.n = n;
            t.d = d;
            t.metric = 1;
            if (n > 0)
                build_euclidean_tree(t, points, chunk);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

cx_status cx_tree_knn_euclidean(cx_tree *t, const float *queries, int64_t nq, int32_t k,
                                int32_t predecessor, int32_t mode, int64_t *out_idx,
                                float *out_dist)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        CX_REQUIRE(t->s.metric == 1, "Euclidean kNN on a residual tree");
        CX_REQUIRE(k >= 1 && nq >= 0, "need k >= 1 and nq >= 0");
        CX_REQUIRE(nq == 0 || (queries && out_idx && out_dist), "NULL query/output pointer");
        CX_REQUIRE(mode == CX_KNN_REFERENCE_F32 || mode == CX_KNN_EXACT, "unknown kNN mode");
        if (nq == 0)
            return;
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        CX_HIP(hipSetDevice(t->s.device));
        TreeState &s = t->s;
        s.n_marks = 0;
        s.last_mode = mode;
        hipStream_t st = s.stream;
        DevBuf<float> q((size_t)nq * s.d), od((size_t)nq * k);
        DevBuf<int64_t> oi((size_t)nq * k);
        CX_HIP(hipMemcpyAsync(q.ptr, queries, sizeof(float) * nq * s.d, hipMemcpyHostToDevice,
                              st));
        knn_euclidean(s, q.ptr, nq, k, predecessor, mode, oi.ptr, od.ptr);
        copy_d2h(out_idx, oi.ptr, sizeof(int64_t) * nq * k, st);
        copy_d2h(out_dist, od.ptr, sizeof(float) * nq * k, st);
    });
}

cx_status cx_tree_info(const cx_tree *t, int64_t *n, int32_t *m, int32_t *d, int32_t *min_level,
                       int32_t *max_level)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        if (n)
            *n = t->s.n;
        if (m)
            *m = t->s.m;
        if (d)
            *d = t->s.d;
        if (min_level)
            *min_level = t->s.min_level;
        if (max_level)
            *max_level = t->s.max_level;
    });
}

#define CX_GETTER(name, field, type)                                                  \
    cx_status name(const cx_tree *t, type *out)                                       \
    {                                                                                 \
        return guarded([&] {                                                          \
            CX_REQUIRE(t != nullptr, "tree is NULL");                                 \
            copy_to_host(t->s, out, t->s.field.ptr, sizeof(type) * (size_t)t->s.n);   \
        });                                                                           \
    }

CX_GETTER(cx_tree_get_parents, parents, int64_t)
CX_GETTER(cx_tree_get_children, children, int64_t)
CX_GETTER(cx_tree_get_next, next, int64_t)
CX_GETTER(cx_tree_get_levels, levels, int32_t)
CX_GETTER(cx_tree_get_si_cache, si, float)
CX_GETTER(cx_tree_get_order, order, int64_t)

cx_status cx_tree_get_subtree_bounds(const cx_tree *t, int64_t *min_out, int64_t *max_out)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        copy_to_host(t->s, min_out, t->s.smin.ptr, sizeof(int64_t) * (size_t)t->s.n);
        copy_to_host(t->s, max_out, t->s.smax.ptr, sizeof(int64_t) * (size_t)t->s.n);
    });
}

cx_status cx_subtree_bounds(const int64_t *parents, const int64_t *node_to_dataset, int64_t n,
                            int64_t *min_out, int64_t *max_out)
{
    return guarded([&] {
        CX_REQUIRE(n >= 0 && n < (int64_t)1 << 31, "n must be in [0, 2^31)");
        if (n == 0)
            return;
        CX_REQUIRE(parents && node_to_dataset && min_out && max_out, "NULL pointer");
        for (int64_t i = 0; i < n; i++)
            CX_REQUIRE(parents[i] >= -1 && parents[i] < n, "parent index out of range");
        require_device();
        subtree_bounds_device(parents, node_to_dataset, n, min_out, max_out);
    });
}

cx_status cx_morton_order(const float *coords, int64_t n, int32_t d, int64_t *order_out)
{
    return guarded([&] {
        CX_REQUIRE(n >= 0 && n < (int64_t)1 << 31, "n must be in [0, 2^31)");
        if (n == 0)
            return;
        CX_REQUIRE(coords && order_out && d >= 1, "NULL pointer or d < 1");
        require_device();
        morton_order_device(coords, n, d, order_out);
    });
}

cx_status cx_nystrom_residual(const double *points, int64_t n, int32_t d,
                              const int64_t *inducing_idx, int32_t m, double var, const double *ls,
                              int32_t n_ls, int32_t kernel_type, double jitter, float *v_out,
                              float *p_out)
{
    return guarded([&] {
        CX_REQUIRE(n >= 1 && n < (int64_t)1 << 31, "n must be in [1, 2^31)");
        CX_REQUIRE(d >= 1 && m >= 1 && m <= n, "need d >= 1 and 1 <= m <= n");
        CX_REQUIRE(points && inducing_idx && v_out && p_out, "NULL pointer");
        for (int32_t j = 0; j < m; j++)
            CX_REQUIRE(inducing_idx[j] >= 0 && inducing_idx[j] < n,
                       "inducing index out of range");
        require_device();
        nystrom_device(points, n, d, inducing_idx, m, var, ls, n_ls,
                       kernel_type == CX_KERNEL_MATERN52 ? 1 : 0, jitter, v_out, p_out);
    });
}

cx_status cx_expf_cr_range(uint32_t first_bits, int64_t count, float *out)
{
    return guarded([&] {
        CX_REQUIRE(count >= 0 && count <= ((int64_t)1 << 32) - (int64_t)first_bits,
                   "count must be in [0, 2^32 - first_bits]");
        if (count == 0)
            return;
        CX_REQUIRE(out != nullptr, "NULL pointer");
        require_device();
        expf_cr_range(first_bits, count, out);
    });
}

cx_status cx_tree_set_screen_segment(cx_tree *t, int32_t seg)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        CX_REQUIRE(screen_segment_valid(seg),
                   "screen segment must be 0 or a multiple of 256 in [256, 16384]");
        std::lock_guard<std::mutex> lock(t->s.mu);
        t->s.screen_seg = seg;
    });
}

cx_status cx_tree_get_screen_segment(cx_tree *t, int32_t *seg_out)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr && seg_out != nullptr, "NULL argument");
        std::lock_guard<std::mutex> lock(t->s.mu);
        *seg_out = t->s.last_seg;
    });
}

cx_status cx_tree_set_profiling(cx_tree *t, int32_t enable)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        t->s.profiling = enable != 0;
        t->s.n_marks = 0;
    });
}

cx_status cx_tree_last_profile(cx_tree *t, float *ms_out, int32_t cap, int32_t *n_out)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr && ms_out != nullptr && n_out != nullptr, "NULL pointer");
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        TreeState &s = t->s;
        const int phases = s.n_marks > 1 ? s.n_marks - 1 : 0;
        *n_out = phases;
        for (int i = 0; i < phases && i < cap; i++) {
            float ms = 0.0f;
            CX_HIP(hipEventElapsedTime(&ms, s.marks[i], s.marks[i + 1]));
            ms_out[i] = ms;
        }
    });
}

cx_status cx_tree_insert_residual(cx_tree *t, const float *V, const float *p,
                                  const float *coords, int64_t n_new, int32_t chunk)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        CX_REQUIRE(n_new >= 0, "n_new must be >= 0");
        CX_REQUIRE(n_new == 0 || (V && p && coords), "V, p and coords must be non-NULL");
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        TreeState &s = t->s;
        CX_REQUIRE(s.metric == 0, "insert_residual on a Euclidean tree");
        CX_REQUIRE(s.n > 0, "insert into an empty tree: build it with the rows instead");
        CX_HIP(hipSetDevice(s.device));
        if (n_new == 0)
            return;
        for (int64_t i = 0; i < n_new; i++)
            s.p_min = p[i] < s.p_min ? p[i] : s.p_min;
        for (int64_t i = 0; i < n_new; i++)
            for (int32_t c = 0; c < s.d; c++) {
                const double v = std::fabs((double)coords[i * s.d + c] * (double)s.inv_ls[c]);
                s.xs_max = v > s.xs_max ? v : s.xs_max;
            }
        insert_residual_rows(s, V, p, coords, n_new, chunk);
    });
}

cx_status cx_tree_insert_euclidean(cx_tree *t, const float *points, int64_t n_new, int32_t chunk)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        CX_REQUIRE(n_new >= 0, "n_new must be >= 0");
        CX_REQUIRE(n_new == 0 || points, "points must be non-NULL");
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        TreeState &s = t->s;
        CX_REQUIRE(s.metric == 1, "insert_euclidean on a residual tree");
        CX_REQUIRE(s.n > 0, "insert into an empty tree: build it with the points instead");
        CX_HIP(hipSetDevice(s.device));
        if (n_new == 0)
            return;
        insert_euclidean_rows(s, points, n_new, chunk);
    });
}

cx_status cx_tree_delete(cx_tree *t, const int64_t *indices, int64_t n_remove)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        CX_REQUIRE(n_remove >= 0 && (n_remove == 0 || indices), "NULL indices");
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        TreeState &s = t->s;
        CX_REQUIRE(n_remove < s.n, "cannot delete every point of a tree");
        CX_HIP(hipSetDevice(s.device));
        if (n_remove == 0)
            return;
        delete_rows(s, indices, n_remove, s.metric == 1);
    });
}

cx_status cx_tree_check_invariants(cx_tree *t, int64_t *out)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr && out != nullptr, "NULL pointer");
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        CX_HIP(hipSetDevice(t->s.device));
        check_invariants(t->s, out);
    });
}

cx_status cx_tree_debug_set_node(cx_tree *t, int64_t node, int64_t parent, int32_t level)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        TreeState &s = t->s;
        CX_REQUIRE(node >= 0 && node < s.n, "node out of range");
        CX_REQUIRE(parent >= -1 && parent < s.n, "parent out of range");
        std::lock_guard<std::mutex> lock(s.mu);
        StreamScope ss_(s.stream);
        CX_HIP(hipSetDevice(s.device));
        CX_HIP(hipMemcpyAsync(s.parents.ptr + node, &parent, sizeof(int64_t),
                              hipMemcpyHostToDevice, s.stream));
        CX_HIP(hipMemcpyAsync(s.levels.ptr + node, &level, sizeof(int32_t),
                              hipMemcpyHostToDevice, s.stream));
        CX_HIP(hipStreamSynchronize(s.stream));
    });
}

cx_status cx_tree_build_telemetry(const cx_tree *t, int64_t *out, int64_t cap_batches,
                                  int64_t *n_batches)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr && n_batches != nullptr, "NULL pointer");
        const TreeState &s = t->s;
        const int64_t nb = (int64_t)s.batch_stats.size() / CX_BATCH_FIELDS;
        *n_batches = nb;
        if (!out)
            return;
        int khz = 0;
        CX_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s.device));
        const double ns_per_tick = khz > 0 ? 1e6 / (double)khz : 0.0;
        for (int64_t b = 0; b < nb && b < cap_batches; b++)
            for (int f = 0; f < CX_BATCH_FIELDS; f++) {
                int64_t v = s.batch_stats[(size_t)(b * CX_BATCH_FIELDS + f)];
                if (f >= CX_BF_NS_SCAN && f <= CX_BF_NS_CONFLICT)
                    v = (int64_t)((double)v * ns_per_tick);
                out[b * CX_BATCH_FIELDS + f] = v;
            }
    });
}

cx_status cx_set_conflict_log(int32_t enable)
{
    return guarded([&] { set_conflict_log(enable != 0); });
}

cx_status cx_tree_conflict_log(const cx_tree *t, int32_t *out, int64_t cap, int64_t *n_words)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr && n_words != nullptr, "NULL pointer");
        const std::vector<int32_t> &log = t->s.conflict_log;
        *n_words = (int64_t)log.size();
        if (out && cap > 0)
            std::memcpy(out, log.data(),
                        sizeof(int32_t) * (size_t)std::min<int64_t>(cap, (int64_t)log.size()));
    });
}

cx_status cx_tree_set_query_telemetry(cx_tree *t, int32_t enable)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        t->s.query_telemetry = enable != 0;
    });
}

cx_status cx_tree_last_query_telemetry(const cx_tree *t, int64_t *out, int32_t cap,
                                       int32_t *n_out)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr && n_out != nullptr, "NULL pointer");
        const TreeState &s = t->s;
        *n_out = (int32_t)s.last_telem.size();
        for (int32_t i = 0; out && i < cap && i < (int32_t)s.last_telem.size(); i++)
            out[i] = s.last_telem[(size_t)i];
    });
}

cx_status cx_tree_clear_last_query_telemetry(cx_tree *t)
{
    return guarded([&] {
        CX_REQUIRE(t != nullptr, "tree is NULL");
        std::lock_guard<std::mutex> lock(t->s.mu);
        StreamScope ss_(t->s.stream);
        t->s.last_telem.clear();
    });
}

cx_status cx_hilbert_order(const double *points, int64_t n, int32_t d, int64_t *order_out)
{
    return guarded([&] {
        CX_REQUIRE(n >= 0 && n < (int64_t)1 << 31, "n must be in [0, 2^31)");
        CX_REQUIRE(d >= 1 && d <= 64, "need 1 <= d <= 64 (codes must fit 64 bits)");
        if (n == 0)
            return;
        CX_REQUIRE(points && order_out, "NULL pointer");
        require_device();
        hilbert_order_device(points, n, d, order_out);
    });
}

cx_status cx_hilbert_order_noalloc(const double *d_points, int64_t n, int32_t d,
                                   int64_t *d_order_out, void *stream)
{
    return guarded([&] {
        CX_REQUIRE(n >= 0, "n must be >= 0");
        CX_REQUIRE(d >= 1 && d <= 64, "need 1 <= d <= 64 (codes must fit 64 bits)");
        if (n == 0)
            return;
        CX_REQUIRE(d_points && d_order_out, "NULL pointer");
        require_device();
        hilbert_order_noalloc(d_points, n, d, d_order_out, (hipStream_t)stream);
    });
}

cx_status cx_luby_mis(const int64_t *indptr, const int64_t *indices, int64_t n, uint32_t seed,
                      uint8_t *selected_out, int64_t *iterations_out)
{
    return guarded([&] {
        CX_REQUIRE(n >= 0 && n < (int64_t)1 << 31, "n must be in [0, 2^31)");
        CX_REQUIRE(iterations_out != nullptr, "NULL pointer");
        *iterations_out = 0;
        if (n == 0)
            return;
        CX_REQUIRE(indptr && selected_out, "NULL pointer");
        CX_REQUIRE(indptr[0] == 0, "indptr[0] must be 0");
        for (int64_t v = 0; v < n; v++)
            CX_REQUIRE(indptr[v + 1] >= indptr[v], "indptr must be non-decreasing");
        CX_REQUIRE(indptr[n] == 0 || indices, "NULL indices");
        for (int64_t e = 0; e < indptr[n]; e++)
            CX_REQUIRE(indices[e] >= 0 && indices[e] < n, "neighbour index out of range");
        require_device();
        *iterations_out = luby_mis_device(indptr, indices, n, seed, selected_out);
    });
}

void cx_tree_free(cx_tree *t) { delete t; }

} // extern "C"
