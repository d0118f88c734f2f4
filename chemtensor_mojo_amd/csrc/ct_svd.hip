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
#include <algorithm>
#include <cmath>

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

namespace {

// dst[i][c] = src[k-1-i][c] for i < chi (rows of length k): the eigenvectors
// of the chi largest eigenvalues, in descending order.
__global__ void take_top_rows_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t k, int64_t chi) {
    const int64_t n = chi * k;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / k, c = t - i * k;
        dst[t] = src[(k - 1 - i) * k + c];
    }
}

}  // namespace

// Two-site split through the Gram matrix (the same factors as
// split_two_site_theta_svd, dmrg.mojo:143-252, with S already absorbed):
//   left-to-right: G = theta theta^T [m x m]; U = eigenvectors of the chi
//                  largest eigenvalues (A_i = U, orthonormal columns);
//                  A_{i+1} = U^T theta = S Vt.
//   right-to-left: G = theta^T theta [n x n]; Vt = top eigenvectors
//                  (A_{i+1} = Vt, orthonormal rows); A_i = theta Vt^T = U S.
// sigma_i^2 are the eigenvalues, so the truncation rule (dense_tensor.mojo:1528-1548)
// and the discarded weight use them directly.  No division by sigma occurs, so
// small singular values cannot destroy orthonormality.  Cost: one GEMM for G,
// one symmetric eigensolve of size min-side, one GEMM for the other factor.
int svd_split_gram(Handle* h, const double* theta, int64_t m, int64_t n, int64_t chi_max, double eps, bool l2r,
                   double* Ai, double* Aj, SvdResult* r) {
    const int64_t k = l2r ? m : n;
    void *wg, *wsel, *wp, *wi;
    CT_CHECK(ws_get(h, WS_SVD, (size_t)(k * k + 2 * k + 64) * 8, &wg));
    CT_CHECK(ws_get(h, WS_SVD2, (size_t)(k * k) * 8, &wsel));
    CT_CHECK(ws_get(h, WS_TMP4, 1024 * 8, &wp));
    double* G = (double*)wg;
    double* lam = G + k * k;
    double* e = lam + k;
    int* info = (int*)(e + k + 8);
    double* part = (double*)wp;
    (void)wi;
    hipLaunchKernelGGL(sumsq_partial, dim3(512), dim3(256), 0, h->stream, theta, m * n, part);
    hipLaunchKernelGGL(sum_final, dim3(1), dim3(256), 0, h->stream, (const double*)part, 512, part + 600);
    {
        GemmDesc g;
        g.M = (int)k; g.N = (int)k; g.ldc = k;
        if (l2r) {  // G[i][i'] = sum_j theta[i][j] theta[i'][j]
            g.K = (int)n;
            g.a_kc = 1; g.sAm = n; g.sAk = 1;
            g.b_kc = 1; g.sBn = n; g.sBk = 1;
        } else {    // G[j][j'] = sum_i theta[i][j] theta[i][j']
            g.K = (int)m;
            g.a_kc = 0; g.sAm = 1; g.sAk = n;
            g.b_kc = 0; g.sBk = n; g.sBn = 1;
        }
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, theta, CT_F64, theta, CT_F64, G, g));
    }
    void* wz;
    CT_CHECK(ws_get(h, WS_SVD3, (size_t)(k * k) * 8, &wz));
    double* Z = (double*)wz;  // eigenvector j in row j
    bool staged = false;
    CT_CHECK(syevd_stage1(h, G, k, lam, Z, info, &staged));
    double* hp = h->h_pinned;
    CT_HIP(hipMemcpyAsync(hp, lam, (size_t)k * 8, hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipMemcpyAsync(hp + k, part + 600, 8, hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipMemcpyAsync(hp + k + 1, info, 2 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipStreamSynchronize(h->stream));
    int hinfo[2] = {0, 0};
    memcpy(hinfo, hp + k + 1, sizeof(hinfo));
    if (hinfo[1] != 0) {
        set_last_error("tridiagonalization: grid-wide wait timed out");
        return CT_STATUS_SVD_FAILED;
    }
    if (hinfo[0] != 0) {
        set_last_error("symmetric eigensolver failed, info=" + std::to_string(hinfo[0]));
        return CT_STATUS_SVD_FAILED;
    }
    const double norm_sq = hp[k];
    const int64_t kk = m < n ? m : n;  // number of singular values
    std::vector<double> sig(kk);
    for (int64_t i = 0; i < kk; ++i) {
        const double l = hp[k - 1 - i];
        sig[i] = l > 0.0 ? std::sqrt(l) : 0.0;
    }
    int64_t chi = kk < chi_max ? kk : chi_max;
    if (eps > 0.0) {
        double tot = 0.0;
        for (int64_t i = 0; i < kk; ++i) tot += sig[i] * sig[i];
        if (tot <= 0.0) {
            chi = 0;
        } else {
            double kept = 0.0;
            const int64_t lim = chi;
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
    if (staged) CT_CHECK(syevd_stage2(h, k, Z, chi));  // only the kept eigenvectors
    double* sel = (double*)wsel;  // [chi][k], descending
    hipLaunchKernelGGL(take_top_rows_kernel, dim3(1024), dim3(256), 0, h->stream, (const double*)Z, sel, k, chi);
    CT_HIP(hipGetLastError());
    if (l2r) {
        const int64_t shape[2] = {chi, m};
        const int perm[2] = {1, 0};
        CT_CHECK(permute(h, CT_F64, sel, 2, shape, perm, Ai));  // U [m][chi]
        GemmDesc g;  // S Vt [chi][n] = sel[chi][m] theta[m][n]
        g.M = (int)chi; g.N = (int)n; g.K = (int)m; g.ldc = n;
        g.a_kc = 1; g.sAm = m; g.sAk = 1;
        g.b_kc = 0; g.sBk = n; g.sBn = 1;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, sel, CT_F64, theta, CT_F64, Aj, g));
    } else {
        CT_HIP(hipMemcpyAsync(Aj, sel, (size_t)(chi * n) * 8, hipMemcpyDeviceToDevice, h->stream));  // Vt
        GemmDesc g;  // U S [m][chi] = theta[m][n] sel[chi][n]^T
        g.M = (int)m; g.N = (int)chi; g.K = (int)n; g.ldc = chi;
        g.a_kc = 1; g.sAm = n; g.sAk = 1;
        g.b_kc = 1; g.sBn = n; g.sBk = 1;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, theta, CT_F64, sel, CT_F64, Ai, g));
    }
    return 0;
}

namespace {

// x[r][c] *= s[r]   (S absorbed into Vt, left-to-right: dmrg.mojo:206-230)
__global__ void scale_rows_kernel(double* __restrict__ x, const double* __restrict__ s, int64_t rows, int64_t cols) {
    const int64_t n = rows * cols;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= s[i / cols];
}

// x[r][c] *= s[c]   (S absorbed into U, right-to-left: dmrg.mojo:231-250)
__global__ void scale_cols_kernel(double* __restrict__ x, const double* __restrict__ s, int64_t rows, int64_t cols) {
    const int64_t n = rows * cols;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= s[i % cols];
}

}  // namespace

int theta_split(Handle* h, int mode, const double* theta, int64_t m, int64_t n, int64_t chi_max, double eps, bool l2r,
                double* Ai, double* Aj, SvdResult* r) {
    if (mode == CT_SVD_GRAM) return svd_split_gram(h, theta, m, n, chi_max, eps, l2r, Ai, Aj, r);
    const int64_t k = std::min(std::min(m, n), chi_max);
    void* sb;
    CT_CHECK(ws_get(h, WS_TMP3, (size_t)(k + 1) * 8, &sb));
    double* S = (double*)sb;
    CT_CHECK(svd_trunc(h, theta, m, n, chi_max, eps, Ai, S, Aj, r));
    const int64_t chi = r->chi;
    if (chi <= 0) return 0;
    const int grid = 1024;
    if (l2r)
        hipLaunchKernelGGL(scale_rows_kernel, dim3(grid), dim3(256), 0, h->stream, Aj, (const double*)S, chi, n);
    else
        hipLaunchKernelGGL(scale_cols_kernel, dim3(grid), dim3(256), 0, h->stream, Ai, (const double*)S, m, chi);
    CT_HIP(hipGetLastError());
    return 0;
}

}  // namespace ct
