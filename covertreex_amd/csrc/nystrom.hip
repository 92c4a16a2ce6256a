// Nystrom factor on the device in FP64 (build_residual_backend,
// covertreex/metrics/residual/host_backend.py:165-246):
//   K_mm = k(U, U) + jitter * var * I           (208-210)
//   L    = cholesky(K_mm)                       (211)
//   V    = (L^-1 K_xm^T)^T, one row per thread  (213-215)
//   p    = max(var - |V_i|^2, 1e-9)             (217-218)
// then FP32 storage of V and p (236-240).  Lengthscales may be per dimension
// (ARD); the reference supports a scalar only.
#include "cx_internal.hpp"

namespace cx {

__device__ __forceinline__ double kernel_f64(const double *__restrict__ a,
                                             const double *__restrict__ b,
                                             const double *__restrict__ inv_ls, int32_t d,
                                             double var, int32_t kernel)
{
    double sq = 0.0;
    for (int32_t c = 0; c < d; c++) {
        double df = (a[c] - b[c]) * inv_ls[c];
        sq += df * df;
    }
    if (kernel == 1) {
        double r = __builtin_sqrt(sq);
        double x = 2.23606797749978969641 * r; // sqrt(5)
        return var * (1.0 + x + x * x / 3.0) * ::exp(-x);
    }
    return var * ::exp(-0.5 * sq);
}

__global__ void kmm_kernel(const double *__restrict__ pts, const int64_t *__restrict__ ind,
                           int32_t m, int32_t d, const double *__restrict__ inv_ls, double var,
                           int32_t kernel, double jitter, double *__restrict__ kmm)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m * m)
        return;
    int32_t i = (int32_t)(t / m), j = (int32_t)(t % m);
    double v = kernel_f64(pts + ind[i] * d, pts + ind[j] * d, inv_ls, d, var, kernel);
    if (i == j)
        v += jitter;
    kmm[t] = v;
}

// Left-looking Cholesky of an m x m SPD matrix in place (lower triangle), one
// workgroup.  status[0] = 1 when a pivot is not positive.
__global__ void __launch_bounds__(1024) cholesky_kernel(double *__restrict__ a, int32_t m,
                                                        int32_t *status)
{
    __shared__ double s_piv;
    for (int32_t j = 0; j < m; j++) {
        if (threadIdx.x == 0) {
            double s = a[(int64_t)j * m + j];
            for (int32_t k = 0; k < j; k++) {
                double l = a[(int64_t)j * m + k];
                s -= l * l;
            }
            if (!(s > 0.0)) {
                status[0] = 1;
                s = 1.0;
            }
            s = __builtin_sqrt(s);
            a[(int64_t)j * m + j] = s;
            s_piv = s;
        }
        __syncthreads();
        const double piv = s_piv;
        for (int32_t i = j + 1 + threadIdx.x; i < m; i += blockDim.x) {
            double s = a[(int64_t)i * m + j];
            for (int32_t k = 0; k < j; k++)
                s -= a[(int64_t)i * m + k] * a[(int64_t)j * m + k];
            a[(int64_t)i * m + j] = s / piv;
        }
        __syncthreads();
    }
}

// Forward substitution v = L^-1 k(x_i, U) for one row per thread; v lives in
// a column-major scratch (vt[j * n + i]) so a wavefront's accesses coalesce,
// while every thread of a wavefront reads the same L entry.
__global__ void __launch_bounds__(256)
    trsv_rows_kernel(const double *__restrict__ pts, int64_t n, int32_t d,
                     const int64_t *__restrict__ ind, int32_t m, const double *__restrict__ inv_ls,
                     double var, int32_t kernel, const double *__restrict__ L,
                     double *__restrict__ vt, float *__restrict__ v_out, float *__restrict__ p_out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double *x = pts + i * d;
    double norm2 = 0.0;
    for (int32_t j = 0; j < m; j++) {
        double s = kernel_f64(x, pts + ind[j] * d, inv_ls, d, var, kernel);
        const double *lj = L + (int64_t)j * m;
        for (int32_t l = 0; l < j; l++)
            s -= lj[l] * vt[(int64_t)l * n + i];
        s /= lj[j];
        vt[(int64_t)j * n + i] = s;
        v_out[i * m + j] = (float)s;
        norm2 += s * s;
    }
    double p = var - norm2;
    p_out[i] = (float)(p > 1e-9 ? p : 1e-9);
}

void nystrom_device(const double *pts_host, int64_t n, int32_t d, const int64_t *ind_host,
                    int32_t m, double var, const double *ls_host, int32_t n_ls, int32_t kernel,
                    double jitter, float *v_host, float *p_host)
{
    // a stream of its own (concurrent callers overlap on the device) that also
    // orders the work buffers' allocations
    struct OwnStream {
        int dev = 0;
        hipStream_t s = nullptr;
        ~OwnStream()
        {
            if (s)
                release_stream(dev, s);
        }
    } own;
    CX_HIP(hipGetDevice(&own.dev));
    own.s = acquire_stream(own.dev);
    ensure_stream_pool(own.dev);
    hipStream_t st = own.s;
    StreamScope scope(st);
    std::vector<double> inv_ls(d);
    for (int32_t c = 0; c < d; c++) {
        double l = (ls_host && n_ls > 0) ? ls_host[c < n_ls ? c : 0] : 1.0;
        inv_ls[c] = 1.0 / (l > 1e-12 ? l : 1e-12); // max(lengthscale, 1e-12), 23-24
    }
    DevBuf<double> pts((size_t)n * d), il(d), kmm((size_t)m * m), vt((size_t)n * m);
    DevBuf<int64_t> ind(m);
    DevBuf<float> v((size_t)n * m), p(n);
    DevBuf<int32_t> status(1);
    CX_HIP(hipMemcpyAsync(pts.ptr, pts_host, sizeof(double) * n * d, hipMemcpyHostToDevice, st));
    CX_HIP(hipMemcpyAsync(il.ptr, inv_ls.data(), sizeof(double) * d, hipMemcpyHostToDevice, st));
    CX_HIP(hipMemcpyAsync(ind.ptr, ind_host, sizeof(int64_t) * m, hipMemcpyHostToDevice, st));
    CX_HIP(hipMemsetAsync(status.ptr, 0, sizeof(int32_t), st));
    const int64_t mm = (int64_t)m * m;
    kmm_kernel<<<(unsigned)((mm + 255) / 256), 256, 0, st>>>(pts.ptr, ind.ptr, m, d, il.ptr, var,
                                                            kernel, jitter * var, kmm.ptr);
    cholesky_kernel<<<1, 1024, 0, st>>>(kmm.ptr, m, status.ptr);
    trsv_rows_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
        pts.ptr, n, d, ind.ptr, m, il.ptr, var, kernel, kmm.ptr, vt.ptr, v.ptr, p.ptr);
    CX_HIP(hipGetLastError());
    int32_t bad = 0;
    CX_HIP(hipMemcpyAsync(&bad, status.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CX_HIP(hipStreamSynchronize(st));
    if (bad)
        throw Error(CX_ERR_INVALID, "inducing kernel matrix is not positive definite");
    CX_HIP(hipMemcpyAsync(v_host, v.ptr, sizeof(float) * n * m, hipMemcpyDeviceToHost, st));
    CX_HIP(hipMemcpyAsync(p_host, p.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    CX_HIP(hipStreamSynchronize(st));
}

} // namespace cx
