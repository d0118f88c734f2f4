// Truncated two-site SVD on the device in FP64 (rocSOLVER gesvd).
//
// Restates dense_tensor_svd_trunc_lapack_f64 (reference
// src/m_tensor/dense_tensor.mojo:1483-1592), which copies theta to the host
// and calls LAPACK dgesdd through native/svd_shim.c.  Here the matrix stays
// in HBM: a row-major [m, n] theta is the column-major [n, m] matrix theta^T,
// whose SVD theta^T = U' S V'^T gives theta = V' S U'^T.  Passing the caller's
// Vt buffer as rocSOLVER's U (ldu = n) and the U buffer as its V^T (ldv = k)
// therefore yields row-major U [m, k] and Vt [k, n] with no transposes.
// Only the k singular values come back to the host, for the truncation rule
// (:1528-1548).
#include "ct_internal.h"
#include "ct_device.h"

namespace ct {

namespace {

// dst[r][c] = src[r][c] for c < chi (src row length k)
__global__ void compact_cols_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t rows, int64_t k,
                                    int64_t chi) {
    const int64_t n = rows * chi;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / chi, c = i - r * chi;
        dst[i] = src[r * k + c];
    }
}

__global__ __launch_bounds__(256) void sumsq_partial(const double* __restrict__ x, int64_t n, double* __restrict__ part) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += x[i] * x[i];
    acc = block_sum256(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void sum_final(const double* __restrict__ part, int np, double* __restrict__ out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
    acc = block_sum256(acc);
    if (threadIdx.x == 0) *out = acc;
}

}  // namespace

int svd_trunc(Handle* h, const double* A, int64_t m, int64_t n, int64_t chi_max, double eps, double* U, double* S,
              double* Vt, SvdResult* r) {
    const int64_t k = m < n ? m : n;
    if (k == 0) {
        r->chi = 0;
        return 0;
    }
    void *wa, *wu, *wv, *we;
    CT_CHECK(ws_get(h, WS_SVD, (size_t)(m * n + 2 * k + 64) * 8, &wa));
    CT_CHECK(ws_get(h, WS_SVD2, (size_t)(m * k) * 8, &wu));
    CT_CHECK(ws_get(h, WS_SVD3, (size_t)(k * n) * 8, &wv));
    CT_CHECK(ws_get(h, WS_TMP4, 1024 * 8, &we));
    double* a = (double*)wa;
    double* s = a + m * n;
    double* e = s + k;
    rocblas_int* info = (rocblas_int*)(e + k + 8);
    double* part = (double*)we;

    // |theta|^2 in FP64 on the device (split_two_site_theta_svd, dmrg.mojo:176)
    hipLaunchKernelGGL(sumsq_partial, dim3(512), dim3(256), 0, h->stream, A, m * n, part);
    hipLaunchKernelGGL(sum_final, dim3(1), dim3(256), 0, h->stream, (const double*)part, 512, part + 600);
    CT_HIP(hipMemcpyAsync(a, A, (size_t)(m * n) * 8, hipMemcpyDeviceToDevice, h->stream));
    rocblas_status st = rocsolver_dgesvd(h->blas, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)n,
                                         (rocblas_int)m, a, (rocblas_int)n, s, (double*)wv, (rocblas_int)n,
                                         (double*)wu, (rocblas_int)k, e, rocblas_outofplace, info);
    if (st != rocblas_status_success) {
        set_last_error(std::string("rocsolver_dgesvd: ") + rocblas_status_to_string(st));
        return CT_STATUS_ROCBLAS_ERROR_BASE + (int)st;
    }
    // read back sigma, info and |theta|^2
    double* hp = h->h_pinned;
    CT_HIP(hipMemcpyAsync(hp, s, (size_t)k * 8, hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipMemcpyAsync(hp + k, part + 600, 8, hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipMemcpyAsync(hp + k + 1, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipStreamSynchronize(h->stream));
    int hinfo = 0;
    memcpy(&hinfo, hp + k + 1, sizeof(int));
    if (hinfo != 0) {
        set_last_error("gesvd failed, info=" + std::to_string(hinfo));
        return CT_STATUS_SVD_FAILED;
    }
    const double norm_sq = hp[k];
    std::vector<double> sig(hp, hp + k);

    // truncation rule (dense_tensor.mojo:1528-1548)
    int64_t chi = k < chi_max ? k : chi_max;
    if (eps > 0.0) {
        double tot = 0.0;
        for (int64_t i = 0; i < k; ++i) tot += sig[i] * sig[i];
        if (tot <= 0.0) {
            chi = 0;
        } else {
            double kept = 0.0;
            const int64_t lim = k < chi_max ? k : chi_max;
            chi = lim;
            for (int64_t t = 1; t <= lim; ++t) {
                kept += sig[t - 1] * sig[t - 1];
                if ((tot - kept) / tot <= eps) {
                    chi = t;
                    break;
                }
            }
        }
    }
    r->chi = chi;
    r->norm_sq = norm_sq;
    double kept = 0.0;
    for (int64_t i = 0; i < chi; ++i) kept += sig[i] * sig[i];
    r->discarded = norm_sq > 0.0 ? (norm_sq - kept) / norm_sq : 0.0;
    if (chi == 0) return 0;

    // U [m, chi] from [m, k]; Vt rows 0..chi-1 are a contiguous prefix; S prefix
    hipLaunchKernelGGL(compact_cols_kernel, dim3(1024), dim3(256), 0, h->stream, (const double*)wu, U, m, k, chi);
    CT_HIP(hipGetLastError());
    CT_HIP(hipMemcpyAsync(Vt, wv, (size_t)(chi * n) * 8, hipMemcpyDeviceToDevice, h->stream));
    CT_HIP(hipMemcpyAsync(S, s, (size_t)chi * 8, hipMemcpyDeviceToDevice, h->stream));
    return 0;
}

}  // namespace ct
