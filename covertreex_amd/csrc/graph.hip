// Python-path build primitives on the device:
//   Hilbert order  -> _hilbert_permutation, covertreex/algo/order/strategy.py:38-99
//   Luby MIS       -> _run_mis_numba_impl, covertreex/algo/_mis_numba.py:64-131
// Both are integer work: bit-exact with the reference's semantics.
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "cx_internal.hpp"
#include "hilbert.hpp"

namespace cx {

// ---- Hilbert ---------------------------------------------------------------

__global__ void minmax_f64_kernel(const double *__restrict__ x, int64_t n, int32_t d, double *lo,
                                  double *hi)
{
    // one block per dimension
    const int32_t c = blockIdx.x;
    __shared__ double smin[256], smax[256];
    double mn = __builtin_huge_val(), mx = -__builtin_huge_val();
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        double v = x[i * d + c];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            smin[threadIdx.x] = smin[threadIdx.x + o] < smin[threadIdx.x] ? smin[threadIdx.x + o]
                                                                          : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + o] > smax[threadIdx.x] ? smax[threadIdx.x + o]
                                                                          : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        lo[c] = smin[0];
        hi[c] = smax[0];
    }
}

__global__ void hilbert_codes_kernel(const double *__restrict__ x, int64_t n, int32_t d,
                                     int32_t bits, const double *__restrict__ lo,
                                     const double *__restrict__ hi, uint64_t *codes, int64_t *idx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint64_t v[64];
    for (int32_t c = 0; c < d; c++)
        v[c] = hilbert_quantize(x[i * d + c], lo[c], hi[c], bits);
    codes[i] = hilbert_from_quantized(v, d, bits);
    idx[i] = i;
}

__global__ void f32_to_f64_kernel(const float *__restrict__ a, int64_t count,
                                  double *__restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        b[i] = (double)a[i];
}

void hilbert_sort_points_f32(const float *d_pts, int64_t n, int32_t d, uint64_t *d_sorted_codes,
                             int64_t *d_perm, double *d_lo, double *d_hi, hipStream_t st)
{
    const int32_t bits = hilbert_bits(d);
    DevBuf<double> x((size_t)n * d);
    DevBuf<uint64_t> codes(n);
    DevBuf<int64_t> idx(n);
    const int64_t cnt = n * d;
    f32_to_f64_kernel<<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(d_pts, cnt, x.ptr);
    minmax_f64_kernel<<<(unsigned)d, 256, 0, st>>>(x.ptr, n, d, d_lo, d_hi);
    hilbert_codes_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(x.ptr, n, d, bits, d_lo,
                                                                      d_hi, codes.ptr, idx.ptr);
    CX_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, codes.ptr, d_sorted_codes,
                                              idx.ptr, d_perm, (int)n, 0, 64, st));
    DevBuf<uint8_t> tmp(tmp_bytes);
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, codes.ptr, d_sorted_codes,
                                              idx.ptr, d_perm, (int)n, 0, 64, st));
}

void hilbert_order_device(const double *points_host, int64_t n, int32_t d, int64_t *order_host)
{
    hipStream_t st = nullptr;
    const int32_t bits = hilbert_bits(d);
    DevBuf<double> x((size_t)n * d), lo(d), hi(d);
    DevBuf<uint64_t> codes(n), sorted(n);
    DevBuf<int64_t> idx(n), order(n);
    CX_HIP(hipMemcpy(x.ptr, points_host, sizeof(double) * n * d, hipMemcpyHostToDevice));
    minmax_f64_kernel<<<(unsigned)d, 256, 0, st>>>(x.ptr, n, d, lo.ptr, hi.ptr);
    hilbert_codes_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(x.ptr, n, d, bits, lo.ptr,
                                                                      hi.ptr, codes.ptr, idx.ptr);
    CX_HIP(hipGetLastError());
    // lexsort((arange, codes)): stable radix sort on the code
    size_t tmp_bytes = 0;
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, codes.ptr, sorted.ptr, idx.ptr,
                                              order.ptr, (int)n, 0, 64, st));
    DevBuf<uint8_t> tmp(tmp_bytes);
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, codes.ptr, sorted.ptr, idx.ptr,
                                              order.ptr, (int)n, 0, 64, st));
    CX_HIP(hipMemcpy(order_host, order.ptr, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
}

// ---- Hilbert order without allocation ---------------------------------------
//
// The same codes and stable sort on device-resident points, with every work
// buffer carved from one file-scope device array instead of device
// allocations, for callers that cannot allocate (graph capture, tight
// serving loops).  Layout: per-axis bounds (2 x 64 doubles), codes, sorted
// codes and indices (kNoAllocMaxPoints each), then the radix sort's
// temporary storage.  Every GPU holds its own copy of the array until the
// process ends; calls that overlap on one GPU share it and race.

constexpr int64_t kNoAllocMaxPoints = (int64_t)1 << 18;
constexpr size_t kNoAllocSortBytes = (size_t)8 << 20;
constexpr size_t kNoAllocFixedBytes =
    2 * 64 * sizeof(double) + 3 * (size_t)kNoAllocMaxPoints * sizeof(uint64_t);
constexpr size_t kNoAllocScratchBytes = kNoAllocFixedBytes + kNoAllocSortBytes;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// the current device's address of `scratch` (looked up once per device)
static unsigned char *scratch_for_current_device()
{
    int dev = 0;
    CX_HIP(hipGetDevice(&dev));
    CX_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *ptr = nullptr;
        CX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        CX_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned no address for scratch");
        g_scratch_ptr[dev] = ptr;
    }
    return (unsigned char *)g_scratch_ptr[dev];
}

void hilbert_order_noalloc(const double *d_points, int64_t n, int32_t d, int64_t *d_order,
                           hipStream_t st)
{
    CX_REQUIRE(n <= kNoAllocMaxPoints, "allocation-free Hilbert order supports at most 262144 "
                                       "points (its fixed scratch array)");
    unsigned char *base = scratch_for_current_device();
    double *lo = (double *)base, *hi = lo + 64;
    uint64_t *codes = (uint64_t *)(hi + 64);
    uint64_t *sorted = codes + kNoAllocMaxPoints;
    int64_t *idx = (int64_t *)(sorted + kNoAllocMaxPoints);
    void *tmp = base + kNoAllocFixedBytes;
    size_t tmp_bytes = 0;
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, codes, sorted, idx, d_order,
                                              (int)n, 0, 64, st));
    CX_REQUIRE(tmp_bytes <= kNoAllocSortBytes, "radix sort storage exceeds the scratch array");
    const int32_t bits = hilbert_bits(d);
    minmax_f64_kernel<<<(unsigned)d, 256, 0, st>>>(d_points, n, d, lo, hi);
    hilbert_codes_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(d_points, n, d, bits, lo,
                                                                      hi, codes, idx);
    CX_HIP(hipGetLastError());
    CX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, codes, sorted, idx, d_order, (int)n,
                                              0, 64, st));
}

// ---- Luby MIS ----------------------------------------------------------------

// xorshift priority (_mis_numba.py:19-28) with 32-bit wrapping as in the JAX
// path (covertreex/algo/mis.py:146-168).  The reference's priority is
// (x + 1) / 2^32 in float64, exact and strictly increasing in x, so the
// 32-bit word x orders the nodes the same way.
__device__ __forceinline__ uint32_t luby_priority(int64_t node, uint32_t seed, int64_t iteration)
{
    uint32_t x = (uint32_t)node ^ seed;
    x ^= (uint32_t)((uint64_t)iteration * 0x9E3779B9ull);
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

__global__ void luby_round_kernel(const int64_t *__restrict__ indptr,
                                  const int64_t *__restrict__ indices, int64_t n, uint32_t seed,
                                  int64_t iteration, const uint8_t *__restrict__ active,
                                  uint8_t *win, int32_t *any_win)
{
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n)
        return;
    win[v] = 0;
    if (!active[v])
        return;
    const uint32_t pv = luby_priority(v, seed, iteration);
    for (int64_t e = indptr[v]; e < indptr[v + 1]; e++) {
        const int64_t u = indices[e];
        if (!active[u])
            continue;
        const uint32_t pu = luby_priority(u, seed, iteration);
        if (pu > pv || (pu == pv && u < v))
            return;
    }
    win[v] = 1;
    *any_win = 1;
}

__global__ void luby_fallback_kernel(const uint8_t *__restrict__ active, int64_t n,
                                     const int32_t *__restrict__ any_win, uint8_t *win)
{
    // no winner: the first active node wins (_mis_numba.py:110-114)
    if (*any_win || threadIdx.x != 0 || blockIdx.x != 0)
        return;
    for (int64_t v = 0; v < n; v++)
        if (active[v]) {
            win[v] = 1;
            return;
        }
}

__global__ void luby_apply_kernel(const int64_t *__restrict__ indptr,
                                  const int64_t *__restrict__ indices, int64_t n,
                                  const uint8_t *__restrict__ win, uint8_t *active, uint8_t *dom,
                                  uint8_t *sel)
{
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n || !win[v])
        return;
    sel[v] = 1;
    for (int64_t e = indptr[v]; e < indptr[v + 1]; e++)
        dom[indices[e]] = 1;
}

__global__ void luby_retire_kernel(int64_t n, const uint8_t *__restrict__ win,
                                   const uint8_t *__restrict__ dom, uint8_t *active,
                                   int32_t *any_active)
{
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n)
        return;
    if (win[v] || dom[v])
        active[v] = 0;
    if (active[v])
        *any_active = 1;
}

// greedy over index order for graphs of max degree <= 8 (_mis_numba.py:36-62)
__global__ void greedy_mis_kernel(const int64_t *__restrict__ indptr,
                                  const int64_t *__restrict__ indices, int64_t n, uint8_t *active,
                                  uint8_t *sel)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    for (int64_t v = 0; v < n; v++) {
        if (!active[v])
            continue;
        sel[v] = 1;
        for (int64_t e = indptr[v]; e < indptr[v + 1]; e++)
            active[indices[e]] = 0;
        active[v] = 0;
    }
}

int64_t luby_mis_device(const int64_t *indptr_host, const int64_t *indices_host, int64_t n,
                        uint32_t seed, uint8_t *sel_host)
{
    hipStream_t st = nullptr;
    const int64_t ne = indptr_host[n];
    int64_t maxdeg = 0;
    for (int64_t v = 0; v < n; v++)
        maxdeg = indptr_host[v + 1] - indptr_host[v] > maxdeg ? indptr_host[v + 1] - indptr_host[v]
                                                              : maxdeg;
    DevBuf<int64_t> indptr(n + 1), indices(ne > 0 ? ne : 1);
    DevBuf<uint8_t> active(n), win(n), dom(n), sel(n);
    DevBuf<int32_t> flags(2);
    CX_HIP(hipMemcpy(indptr.ptr, indptr_host, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    if (ne > 0)
        CX_HIP(hipMemcpy(indices.ptr, indices_host, sizeof(int64_t) * ne, hipMemcpyHostToDevice));
    CX_HIP(hipMemset(active.ptr, 1, n));
    CX_HIP(hipMemset(sel.ptr, 0, n));
    int64_t iterations = 0;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (maxdeg <= 8) {
        greedy_mis_kernel<<<1, 64, 0, st>>>(indptr.ptr, indices.ptr, n, active.ptr, sel.ptr);
        CX_HIP(hipGetLastError());
        iterations = 1;
    } else {
        for (;;) {
            iterations++;
            CX_HIP(hipMemset(flags.ptr, 0, 2 * sizeof(int32_t)));
            CX_HIP(hipMemset(dom.ptr, 0, n));
            luby_round_kernel<<<blocks, 256, 0, st>>>(indptr.ptr, indices.ptr, n, seed,
                                                      iterations, active.ptr, win.ptr, flags.ptr);
            luby_fallback_kernel<<<1, 64, 0, st>>>(active.ptr, n, flags.ptr, win.ptr);
            luby_apply_kernel<<<blocks, 256, 0, st>>>(indptr.ptr, indices.ptr, n, win.ptr,
                                                      active.ptr, dom.ptr, sel.ptr);
            luby_retire_kernel<<<blocks, 256, 0, st>>>(n, win.ptr, dom.ptr, active.ptr,
                                                       flags.ptr + 1);
            CX_HIP(hipGetLastError());
            int32_t any_active = 0;
            CX_HIP(hipMemcpy(&any_active, flags.ptr + 1, sizeof(int32_t), hipMemcpyDeviceToHost));
            if (!any_active || iterations > n + 1)
                break;
        }
    }
    CX_HIP(hipMemcpy(sel_host, sel.ptr, n, hipMemcpyDeviceToHost));
    return iterations;
}

} // namespace cx
