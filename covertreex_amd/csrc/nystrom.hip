// Nystrom factor on the device in FP64 (build_residual_backend,
// covertreex/metrics/residual/host_backend.py:165-246):
//   K_mm = k(U, U) + jitter * var * I           (208-210)
//   L    = cholesky(K_mm)                       (211), blocked, one workgroup
//   V    = (L^-1 K_xm^T)^T                      (213-215), blocked forward
//          substitution in column blocks of 32 (rows resident in LDS up to
//          m = 256, one row per thread above)
//   p    = max(var - |V_i|^2, 1e-9)             (217-218)
// then FP32 storage of V and p (236-240).  Lengthscales may be per dimension
// (ARD); the reference supports a scalar only.
#include <mutex>

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

// Blocked left-looking Cholesky of an m x m SPD matrix in place (lower
// triangle), one workgroup of 1024 threads, panels of kNb columns:
//   1. the panel's top rows L[jb : jb+nb, 0 : jb] (final) go to LDS;
//   2. panel update A[jb:m, jb:jb+nb] -= L[jb:m, 0:jb] L[jb:jb+nb, 0:jb]^T,
//      one thread per (row, column) entry, rows read from global memory (the
//      matrix stays in L2);
//   3. the nb x nb diagonal block is factored in LDS (one column at a time);
//   4. the rows below are solved against it by forward substitution, one
//      thread per row.
// The sums are those of column-by-column Cholesky in another order.
// status[0] = 1 when a pivot is not positive.
constexpr int kNb = 32;
constexpr int kCholThreads = 1024;
constexpr int kCholMaxM = 512;

__global__ void __launch_bounds__(kCholThreads) cholesky_kernel(double *__restrict__ a, int32_t m,
                                                                int32_t *status)
{
    // top rows of the panel, padded by one double against bank conflicts
    __shared__ double s_top[kNb * (kCholMaxM + 1)];
    __shared__ double s_diag[kNb * (kNb + 1)];
    const int tid = threadIdx.x;
    const int ld = kCholMaxM + 1;
    for (int32_t jb = 0; jb < m; jb += kNb) {
        const int nb = m - jb < kNb ? m - jb : kNb;
        for (int t = tid; t < nb * jb; t += kCholThreads) {
            const int r = t / jb, k = t % jb;
            s_top[r * ld + k] = a[(int64_t)(jb + r) * m + k];
        }
        __syncthreads();
        // 2. panel update over rows jb .. m-1, columns jb .. jb+nb-1
        const int64_t rows = m - jb;
        for (int64_t t = tid; t < rows * nb; t += kCholThreads) {
            const int32_t i = jb + (int32_t)(t / nb), c = (int32_t)(t % nb);
            if (i < jb + c)
                continue; // above the diagonal of the block
            const double *li = a + (int64_t)i * m;
            double s = li[jb + c];
            for (int32_t k = 0; k < jb; k++)
                s -= li[k] * s_top[c * ld + k];
            a[(int64_t)i * m + jb + c] = s;
        }
        __syncthreads();
        // 3. factor the diagonal block in LDS
        for (int t = tid; t < nb * nb; t += kCholThreads) {
            const int r = t / nb, c = t % nb;
            s_diag[r * (kNb + 1) + c] = a[(int64_t)(jb + r) * m + jb + c];
        }
        __syncthreads();
        if (tid < 64) {
            for (int j = 0; j < nb; j++) {
                if (tid == 0) {
                    double s = s_diag[j * (kNb + 1) + j];
                    for (int k = 0; k < j; k++) {
                        const double l = s_diag[j * (kNb + 1) + k];
                        s -= l * l;
                    }
                    if (!(s > 0.0)) {
                        status[0] = 1;
                        s = 1.0;
                    }
                    s_diag[j * (kNb + 1) + j] = __builtin_sqrt(s);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                const double piv = s_diag[j * (kNb + 1) + j];
                const int i = j + 1 + tid;
                if (i < nb) {
                    double s = s_diag[i * (kNb + 1) + j];
                    for (int k = 0; k < j; k++)
                        s -= s_diag[i * (kNb + 1) + k] * s_diag[j * (kNb + 1) + k];
                    s_diag[i * (kNb + 1) + j] = s / piv;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
        }
        __syncthreads();
        for (int t = tid; t < nb * nb; t += kCholThreads) {
            const int r = t / nb, c = t % nb;
            if (c <= r)
                a[(int64_t)(jb + r) * m + jb + c] = s_diag[r * (kNb + 1) + c];
        }
        // 4. rows below the block: x L_bb^T = a  (forward substitution)
        for (int32_t i = jb + nb + tid; i < m; i += kCholThreads) {
            double *ri = a + (int64_t)i * m + jb;
            double x[kNb];
#pragma unroll
            for (int c = 0; c < kNb; c++)
                x[c] = c < nb ? ri[c] : 0.0;
#pragma unroll
            for (int c = 0; c < kNb; c++) {
                if (c < nb) {
                    double s = x[c];
                    for (int k = 0; k < c; k++)
                        s -= x[k] * s_diag[c * (kNb + 1) + k];
                    x[c] = s / s_diag[c * (kNb + 1) + c];
                }
            }
#pragma unroll
            for (int c = 0; c < kNb; c++)
                if (c < nb)
                    ri[c] = x[c];
        }
        __syncthreads();
    }
}

// lt[k * m + j] = L[j][k] for k <= j (L^T, upper triangle), so that a column
// block's entries of one row k of L^T are contiguous
__global__ void transpose_lower_kernel(const double *__restrict__ l, int32_t m,
                                       double *__restrict__ lt)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m * m)
        return;
    const int32_t j = (int32_t)(t / m), k = (int32_t)(t % m);
    lt[(int64_t)k * m + j] = k <= j ? l[t] : 0.0;
}

// One column block of V = K_xm L^-T, one row per thread: the block's kNb
// columns c of row i start at k(x_i, u_c), take the earlier columns' terms
// sum_{l < jb} V[i][l] L[c][l] (V read column-major from vt, L^T rows
// contiguous and the same for the whole wavefront), then the block is solved
// by forward substitution against L's diagonal block.  The row's squared norm
// accumulates in FP64 across blocks; the last block writes p.
__global__ void __launch_bounds__(64)
    trsm_block_kernel(const double *__restrict__ pts, int64_t n, int32_t d,
                      const int64_t *__restrict__ ind, int32_t m,
                      const double *__restrict__ inv_ls, double var, int32_t kernel,
                      const double *__restrict__ lt, int32_t jb, double *__restrict__ vt,
                      double *__restrict__ norm2, float *__restrict__ v_out,
                      float *__restrict__ p_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int nb = m - jb < kNb ? m - jb : kNb;
    const double *x = pts + i * d;
    double acc[kNb];
#pragma unroll
    for (int c = 0; c < kNb; c++)
        acc[c] = c < nb ? kernel_f64(x, pts + ind[jb + c] * d, inv_ls, d, var, kernel) : 0.0;
    for (int32_t l = 0; l < jb; l++) {
        const double v = vt[(int64_t)l * n + i];
        const double *row = lt + (int64_t)l * m + jb;
#pragma unroll
        for (int c = 0; c < kNb; c++)
            acc[c] -= v * (c < nb ? row[c] : 0.0);
    }
    double s2 = jb > 0 ? norm2[i] : 0.0;
#pragma unroll
    for (int c = 0; c < kNb; c++) {
        if (c < nb) {
            double s = acc[c];
            for (int k = 0; k < c; k++)
                s -= acc[k] * lt[(int64_t)(jb + k) * m + jb + c];
            s /= lt[(int64_t)(jb + c) * m + jb + c];
            acc[c] = s;
            vt[(int64_t)(jb + c) * n + i] = s;
            v_out[i * m + jb + c] = (float)s;
            s2 += s * s;
        }
    }
    if (jb + nb < m) {
        norm2[i] = s2;
    } else {
        const double p = var - s2;
        p_out[i] = (float)(p > 1e-9 ? p : 1e-9);
    }
}

// V = K_xm L^-T for a block of R rows per workgroup, the rows' FP64 values
// resident in LDS for the whole solve (written once, as FP32, at the end):
//   1. sX[r][c] = k(x_{r0+r}, u_c) for every column (rows padded by one
//      double, so a wavefront reading one column of many rows spreads over
//      the banks);
//   2. per column block of kNb: every (row, column) entry subtracts the
//      earlier columns' terms sum_{l < jb} X[r][l] L[c][l], l ascending,
//      with L's rows staged through LDS in 32 x 32 tiles; then L's diagonal
//      block goes to LDS and one thread per row solves against it;
//   3. V rows out in FP32 and p = max(var - |V_i|^2, 1e-9) (the squared norm
//      summed in column order).
// Every value is the one column-by-column forward substitution gives: the
// same operations in the same order for each entry.
template <int R>
__global__ void __launch_bounds__(256)
    trsm_rows_kernel(const double *__restrict__ pts, int64_t n, int32_t d,
                     const int64_t *__restrict__ ind, int32_t m,
                     const double *__restrict__ inv_ls, double var, int32_t kernel,
                     const double *__restrict__ lt, float *__restrict__ v_out,
                     float *__restrict__ p_out)
{
    extern __shared__ double sX[]; // R x (m + 1)
    __shared__ double sL[kNb * kNb]; // staged L tile: sL[l * kNb + c] = L[jb + c][l0 + l]
    const int tid = threadIdx.x;
    const int ld = m + 1;
    const int64_t r0 = (int64_t)blockIdx.x * R;
    const int rows = n - r0 < R ? (int)(n - r0) : R;
    for (int t = tid; t < R * m; t += 256) {
        const int r = t / m, c = t % m;
        sX[r * ld + c] = r < rows ? kernel_f64(pts + (r0 + r) * d, pts + ind[c] * d, inv_ls,
                                               d, var, kernel)
                                  : 0.0;
    }
    __syncthreads();
    constexpr int kRg = 256 / kNb;      // row groups
    constexpr int kPer = R / kRg;       // rows per thread in the update
    const int c = tid % kNb, rg = tid / kNb;
    for (int32_t jb = 0; jb < m; jb += kNb) {
        const int nb = m - jb < kNb ? m - jb : kNb;
        double acc[kPer];
#pragma unroll
        for (int i = 0; i < kPer; i++)
            acc[i] = c < nb ? sX[(rg + kRg * i) * ld + jb + c] : 0.0;
        for (int32_t l0 = 0; l0 < jb; l0 += kNb) {
            // L[jb + c][l0 + l] for l < 32 (l0 + 32 <= jb: blocks are whole)
            for (int t = tid; t < kNb * kNb; t += 256) {
                const int l = t / kNb, cc = t % kNb;
                sL[t] = cc < nb ? lt[(int64_t)(l0 + l) * m + jb + cc] : 0.0;
            }
            __syncthreads();
#pragma unroll 8
            for (int l = 0; l < kNb; l++) {
                const double lv = sL[l * kNb + c];
#pragma unroll
                for (int i = 0; i < kPer; i++)
                    acc[i] -= sX[(rg + kRg * i) * ld + l0 + l] * lv;
            }
            __syncthreads();
        }
        if (c < nb)
#pragma unroll
            for (int i = 0; i < kPer; i++)
                sX[(rg + kRg * i) * ld + jb + c] = acc[i];
        // L's diagonal block: sL[k * kNb + cc] = L[jb + cc][jb + k]
        for (int t = tid; t < kNb * kNb; t += 256) {
            const int k = t / kNb, cc = t % kNb;
            sL[t] = (k < nb && cc < nb) ? lt[(int64_t)(jb + k) * m + jb + cc] : 0.0;
        }
        __syncthreads();
        if (tid < rows) {
            double x[kNb];
            double *xr = sX + tid * ld + jb;
#pragma unroll
            for (int cc = 0; cc < kNb; cc++)
                x[cc] = cc < nb ? xr[cc] : 0.0;
#pragma unroll
            for (int cc = 0; cc < kNb; cc++) {
                if (cc < nb) {
                    double s = x[cc];
                    for (int k = 0; k < cc; k++)
                        s -= x[k] * sL[k * kNb + cc];
                    x[cc] = s / sL[cc * kNb + cc];
                }
            }
#pragma unroll
            for (int cc = 0; cc < kNb; cc++)
                if (cc < nb)
                    xr[cc] = x[cc];
        }
        __syncthreads();
    }
    for (int t = tid; t < rows * m; t += 256) {
        const int r = t / m, cc = t % m;
        v_out[(r0 + r) * m + cc] = (float)sX[r * ld + cc];
    }
    if (tid < rows) {
        double s2 = 0.0;
        for (int cc = 0; cc < m; cc++) {
            const double x = sX[tid * ld + cc];
            s2 += x * x;
        }
        const double pv = var - s2;
        p_out[r0 + tid] = (float)(pv > 1e-9 ? pv : 1e-9);
    }
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
    CX_REQUIRE(m <= kCholMaxM, "the Nystrom factor supports at most 512 inducing points");
    DevBuf<double> pts((size_t)n * d), il(d), kmm((size_t)m * m), lt((size_t)m * m);
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
    cholesky_kernel<<<1, kCholThreads, 0, st>>>(kmm.ptr, m, status.ptr);
    transpose_lower_kernel<<<(unsigned)((mm + 255) / 256), 256, 0, st>>>(kmm.ptr, m, lt.ptr);
    // V and p.  Up to m = 256 the rows' values stay in LDS (32 rows per
    // workgroup, two workgroups per CU): C2 (n = 32,768, m = 256) 0.42 ms
    // against 1.34 ms for the column-block launches.  Above, a workgroup's
    // rows no longer leave room for a second workgroup per CU and the
    // column-block launches are faster (m = 512, n = 262,144: 9.8 ms against
    // 14.8-15.9 ms with 16 or 32 resident rows).  Both give the same bits.
    DevBuf<double> vt, norm2;
    if (m <= 256) {
        static std::once_flag once; // more than 64 KiB of dynamic LDS
        std::call_once(once, [] {
            CX_HIP(hipFuncSetAttribute((const void *)trsm_rows_kernel<32>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(sizeof(double) * 32 * (256 + 1))));
        });
        trsm_rows_kernel<32><<<(unsigned)((n + 31) / 32), 256, sizeof(double) * 32 * (m + 1),
                               st>>>(pts.ptr, n, d, ind.ptr, m, il.ptr, var, kernel, lt.ptr,
                                     v.ptr, p.ptr);
    } else {
        vt.alloc((size_t)n * m);
        norm2.alloc(n);
        for (int32_t jb = 0; jb < m; jb += kNb)
            trsm_block_kernel<<<(unsigned)((n + 63) / 64), 64, 0, st>>>(
                pts.ptr, n, d, ind.ptr, m, il.ptr, var, kernel, lt.ptr, jb, vt.ptr, norm2.ptr,
                v.ptr, p.ptr);
    }
    CX_HIP(hipGetLastError());
    int32_t bad = 0;
    CX_HIP(hipMemcpyAsync(&bad, status.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CX_HIP(hipStreamSynchronize(st));
    if (bad)
        throw Error(CX_ERR_INVALID, "inducing kernel matrix is not positive definite");
    copy_d2h(v_host, v.ptr, sizeof(float) * n * m, st);
    copy_d2h(p_host, p.ptr, sizeof(float) * n, st);
}

} // namespace cx
