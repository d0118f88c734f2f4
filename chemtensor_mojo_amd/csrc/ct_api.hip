// extern "C" entry points of libdmrg_hip.so (declared in include/chemtensor_hip.h)
// and the handle / workspace management behind them.
#include <algorithm>
#include <new>

#include "ct_internal.h"

namespace ct {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

int buf_get(Handle* h, Buffer& b, size_t bytes, void** out) {
    if (bytes == 0) bytes = 16;
    if (h->arena) {  // an allocation-free call never grows a buffer
        if (b.bytes >= bytes) {
            *out = b.ptr;
            return 0;
        }
        set_last_error("allocation-free call: a buffer would have to grow");
        return CT_STATUS_ALLOC;
    }
    if (b.bytes < bytes) {
        // never inside a stream capture: the synchronisation and the new
        // address would break the graph being recorded
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (h->stream && hipStreamIsCapturing(h->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
            set_last_error("workspace growth during graph capture");
            return CT_STATUS_NOT_SUPPORTED;
        }
        // earlier work on the stream may still read the old buffer
        CT_HIP(hipStreamSynchronize(h->stream));
        ++h->ws_gen;
        if (b.ptr) CT_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
        size_t want = bytes + bytes / 8;  // slack for slowly growing bond dimensions
        want = (want + 4095) & ~(size_t)4095;
        hipError_t e = hipMalloc(&b.ptr, want);
        if (e != hipSuccess) {
            b.ptr = nullptr;
            set_last_error("workspace allocation of " + std::to_string(want) + " bytes failed: " +
                           hipGetErrorString(e));
            return CT_STATUS_ALLOC;
        }
        b.bytes = want;
    }
    *out = b.ptr;
    return 0;
}

void buf_free(Buffer& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

bool arena_reserve(Arena& a, int slot, size_t bytes) {
    const size_t start = (a.used + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 16;
    if (start > a.cap || bytes > a.cap - start) return false;
    a.off[slot] = start;
    a.len[slot] = bytes;
    a.used = start + bytes;
    return true;
}

int ws_get(Handle* h, int slot, size_t bytes, void** out) {
    if (h->arena) {
        Arena& a = *h->arena;
        if (bytes == 0) bytes = 16;
        if (a.len[slot] < bytes) {
            set_last_error("allocation-free call: workspace slot " + std::to_string(slot) + " needs " +
                           std::to_string(bytes) + " bytes, " + std::to_string(a.len[slot]) + " were planned");
            return CT_STATUS_ALLOC;
        }
        *out = a.base + a.off[slot];
        return 0;
    }
    return buf_get(h, h->ws[slot], bytes, out);
}

}  // namespace ct

static constexpr size_t kPinnedBytes = 1 << 20;

extern "C" {

const char* ct_last_error(void) { return ct::g_last_error.c_str(); }

const char* ct_version(void) { return "chemtensor-mi355x 0.1 (gfx950)"; }

int ct_handle_create(int device, void* stream, ct_handle_t* out) {
    CT_ARG(out != nullptr, 3, "out is NULL");
    int ndev = 0;
    CT_HIP(hipGetDeviceCount(&ndev));
    CT_ARG(device >= 0 && device < ndev, 1, "no such device");
    CT_HIP(hipSetDevice(device));
    ct_handle_s* h = new ct_handle_s();
    h->device = device;
    // NULL is the default (null) stream, as everywhere in HIP: the library
    // then orders with every other user of the default stream.
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
    if (rocblas_create_handle(&h->blas) != rocblas_status_success) {
        delete h;
        ct::set_last_error("rocblas_create_handle failed");
        return CT_STATUS_ROCBLAS_ERROR_BASE;
    }
    rocblas_set_stream(h->blas, h->stream);
    int rc = ct::hip_rc(hipMalloc(&h->d_scalars, 4096));
    if (rc == 0) rc = ct::hip_rc(hipHostMalloc(&h->h_pinned, kPinnedBytes, hipHostMallocDefault));
    // the allocation-free entry point's scratch array on this device (once per device)
    if (rc == 0) rc = ct::noalloc_scratch_init(device);
    if (rc != 0) {
        (void)ct_handle_destroy(h);
        return rc;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->num_cu = prop.multiProcessorCount;
    *out = h;
    return 0;
}

int ct_handle_destroy(ct_handle_t h) {
    if (!h) return 0;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    for (auto& b : h->ws)
        if (b.ptr) (void)hipFree(b.ptr);
    if (h->d_scalars) (void)hipFree(h->d_scalars);
    if (h->h_pinned) (void)hipHostFree(h->h_pinned);
    if (h->blas) rocblas_destroy_handle(h->blas);
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

int ct_handle_set_stream(ct_handle_t h, void* stream) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_HIP(hipStreamSynchronize(h->stream));
    if (h->own_stream) CT_HIP(hipStreamDestroy(h->stream));
    h->stream = (hipStream_t)stream;  // NULL = the default stream
    h->own_stream = false;
    rocblas_set_stream(h->blas, h->stream);
    return 0;
}

int ct_synchronize(ct_handle_t h) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

int ct_tensordot(ct_handle_t h, int prec, int dtype, const void* A, int rankA, const int64_t* shapeA, const void* B,
                 int rankB, const int64_t* shapeB, int ndim_mult, int axrange_A, int axrange_B, void* C, int rankC,
                 const int64_t* shapeC) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(prec >= 0 && prec <= CT_PREC_BF16X6, 2, "unknown precision");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 3, "unknown dtype");
    return ct::tensordot(h, prec, dtype, A, rankA, shapeA, B, rankB, shapeB, ndim_mult, axrange_A, axrange_B, C,
                         rankC, shapeC);
}

int ct_split_bf16(ct_handle_t h, int dtype, const void* X, int64_t rows, int64_t K, int64_t s_row, int64_t s_k,
                  uint16_t* out, int64_t ld) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 2, "unknown dtype");
    CT_ARG(rows >= 0 && rows < (1ll << 31), 4, "rows out of range");
    CT_ARG(K >= 0 && K < (1ll << 31), 5, "K out of range");
    CT_ARG(ld >= 2 * ct::split_k32(K), 9, "ld must be >= 2*roundup(K, 32)");
    return ct::split_bf16(h, dtype, X, (int)rows, (int)K, s_row, s_k, 1, 0, out, ld, 0, nullptr);
}

int ct_gemm_bf16x3_nt(ct_handle_t h, const uint16_t* A, int64_t lda, const uint16_t* B, int64_t ldb, int dtype_c,
                      void* C, int64_t M, int64_t N, int64_t K, int64_t ldc) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(lda >= 2 * ct::split_k32(K), 3, "lda must be >= 2*roundup(K, 32)");
    CT_ARG(ldb >= 2 * ct::split_k32(K), 5, "ldb must be >= 2*roundup(K, 32)");
    CT_ARG(dtype_c == CT_F64 || dtype_c == CT_F32, 6, "unknown dtype");
    CT_ARG(M >= 0 && M < (1ll << 31), 8, "M out of range");
    CT_ARG(N >= 0 && N < (1ll << 31), 9, "N out of range");
    CT_ARG(K >= 0 && K < (1ll << 31), 10, "K out of range");
    CT_ARG(ldc >= N, 11, "ldc must be >= N");
    ct::SplitOp a, b;
    a.p = A; a.ld = lda;
    b.p = B; b.ld = ldb;
    return ct::gemm_bf16x3_nt(h, a, b, dtype_c, C, (int)M, (int)N, (int)K, ldc, 0, 1, 1.0, 0.0, nullptr);
}

int ct_split_bf16x6(ct_handle_t h, int dtype, const void* X, int64_t rows, int64_t K, int64_t s_row, int64_t s_k,
                    uint16_t* out, int64_t ld, uint16_t* outl, int64_t ldl) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 2, "unknown dtype");
    CT_ARG(rows >= 0 && rows < (1ll << 31), 4, "rows out of range");
    CT_ARG(K >= 0 && K < (1ll << 31), 5, "K out of range");
    CT_ARG(ld >= 2 * ct::split_k32(K), 9, "ld must be >= 2*roundup(K, 32)");
    CT_ARG(outl != nullptr, 10, "outl is NULL");
    CT_ARG(ldl >= ct::split_k32(K), 11, "ldl must be >= roundup(K, 32)");
    return ct::split_bf16(h, dtype, X, (int)rows, (int)K, s_row, s_k, 1, 0, out, ld, 0, nullptr, outl, ldl, 0);
}

int ct_gemm_bf16x6_nt(ct_handle_t h, const uint16_t* A, int64_t lda, const uint16_t* Al, int64_t ldal,
                      const uint16_t* B, int64_t ldb, const uint16_t* Bl, int64_t ldbl, int dtype_c, void* C,
                      int64_t M, int64_t N, int64_t K, int64_t ldc) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(lda >= 2 * ct::split_k32(K), 3, "lda must be >= 2*roundup(K, 32)");
    CT_ARG(Al != nullptr, 4, "Al is NULL");
    CT_ARG(ldal >= ct::split_k32(K), 5, "ldal must be >= roundup(K, 32)");
    CT_ARG(ldb >= 2 * ct::split_k32(K), 7, "ldb must be >= 2*roundup(K, 32)");
    CT_ARG(Bl != nullptr, 8, "Bl is NULL");
    CT_ARG(ldbl >= ct::split_k32(K), 9, "ldbl must be >= roundup(K, 32)");
    CT_ARG(dtype_c == CT_F64 || dtype_c == CT_F32, 10, "unknown dtype");
    CT_ARG(M >= 0 && M < (1ll << 31), 12, "M out of range");
    CT_ARG(N >= 0 && N < (1ll << 31), 13, "N out of range");
    CT_ARG(K >= 0 && K < (1ll << 31), 14, "K out of range");
    CT_ARG(ldc >= N, 15, "ldc must be >= N");
    ct::SplitOp a, b;
    a.p = A; a.ld = lda; a.l = Al; a.ldl = ldal;
    b.p = B; b.ld = ldb; b.l = Bl; b.ldl = ldbl;
    return ct::gemm_bf16x6_nt(h, a, b, dtype_c, C, (int)M, (int)N, (int)K, ldc, 0, 1, 1.0, 0.0, nullptr);
}

int ct_permute(ct_handle_t h, int dtype, const void* in, int rank, const int64_t* shape, const int* perm, void* out) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 2, "unknown dtype");
    CT_ARG(rank >= 1 && rank <= 8, 4, "rank must be 1..8");
    bool seen[8] = {false};
    for (int i = 0; i < rank; ++i) {
        CT_ARG(perm[i] >= 0 && perm[i] < rank && !seen[perm[i]], 6, "perm is not a permutation");
        seen[perm[i]] = true;
    }
    return ct::permute(h, dtype, in, rank, shape, perm, out);
}

int ct_norm_sq(ct_handle_t h, int dtype, const void* x, int64_t n, double* result) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(result != nullptr, 5, "result is NULL");
    return ct::dot_host(h, dtype, x, x, n, result);
}

int ct_dot(ct_handle_t h, int dtype, const void* x, const void* y, int64_t n, double* result) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(result != nullptr, 6, "result is NULL");
    return ct::dot_host(h, dtype, x, y, n, result);
}

int ct_axpy(ct_handle_t h, int dtype, double alpha, const void* x, void* y, int64_t n) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    return ct::axpy(h, dtype, alpha, x, y, n);
}

int ct_scale(ct_handle_t h, int dtype, double alpha, void* x, int64_t n) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    return ct::scale(h, dtype, alpha, x, n);
}

int ct_convert(ct_handle_t h, int dtype_in, const void* in, int dtype_out, void* out, int64_t n) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    return ct::convert(h, dtype_in, in, dtype_out, out, n);
}

int ct_heff_apply(ct_handle_t h, int prec, int dtype, const void* L, const void* Wi, const void* Wj, const void* R,
                  const void* theta, void* out, int64_t Dl, int64_t Dlb, int64_t d, int64_t Dr, int64_t Drb,
                  int64_t wL, int64_t wM, int64_t wR) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(prec >= 0 && prec <= CT_PREC_BF16X6, 2, "unknown precision");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 3, "unknown dtype");
    CT_ARG(Dl > 0 && Dlb > 0 && d > 0 && Dr > 0 && Drb > 0 && wL > 0 && wM > 0 && wR > 0, 11, "dimensions must be > 0");
    ct::HeffDims hd{Dl, Dlb, d, Dr, Drb, wL, wM, wR};
    void* w2;
    CT_CHECK(ct::ws_get(h, ct::WS_W2, (size_t)(wL * d * d * d * d * wR) * 8, &w2));
    CT_CHECK(ct::heff_build_w2(h, dtype, Wi, Wj, wL, wM, wR, d, (double*)w2));
    ct::HeffEnv env;
    CT_CHECK(ct::heff_prepare(h, prec, dtype, L, R, (const double*)w2, hd, &env));
    return ct::heff_apply(h, env, dtype, theta, out, nullptr);
}

struct ct_heff_s {
    ct::Handle* h = nullptr;
    int dtype = CT_F64;
    ct::HeffEnv env;
    ct::HeffStore store;
};

int ct_heff_create(ct_handle_t h, int prec, int dtype, const void* L, const void* Wi, const void* Wj, const void* R,
                   int64_t Dl, int64_t Dlb, int64_t d, int64_t Dr, int64_t Drb, int64_t wL, int64_t wM, int64_t wR,
                   ct_heff_t* op) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(prec >= 0 && prec <= CT_PREC_BF16X6, 2, "unknown precision");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 3, "unknown dtype");
    CT_ARG(Dl > 0 && Dlb > 0 && d > 0 && Dr > 0 && Drb > 0 && wL > 0 && wM > 0 && wR > 0, 8, "dimensions must be > 0");
    CT_ARG(op != nullptr, 16, "op is NULL");
    *op = nullptr;
    ct_heff_s* o = new (std::nothrow) ct_heff_s();
    if (!o) {
        ct::set_last_error("ct_heff_create: out of host memory");
        return CT_STATUS_ALLOC;
    }
    o->h = h;
    o->dtype = dtype;
    ct::HeffDims hd{Dl, Dlb, d, Dr, Drb, wL, wM, wR};
    void* w2 = nullptr;
    int rc = ct::buf_get(h, o->store.w2, (size_t)(wL * d * d * d * d * wR) * 8, &w2);
    if (rc == 0) rc = ct::heff_build_w2(h, dtype, Wi, Wj, wL, wM, wR, d, (double*)w2);
    if (rc == 0) rc = ct::heff_prepare(h, prec, dtype, L, R, (const double*)w2, hd, &o->env, &o->store);
    if (rc != 0) {
        (void)hipStreamSynchronize(h->stream);
        delete o;
        return rc;
    }
    *op = o;
    return 0;
}

int ct_heff_apply_prepared(ct_heff_t op, const void* theta, void* out) {
    CT_ARG(op != nullptr, 1, "op is NULL");
    return ct::heff_apply(op->h, op->env, op->dtype, theta, out, nullptr);
}

int ct_heff_destroy(ct_heff_t op) {
    if (!op) return 0;
    CT_HIP(hipStreamSynchronize(op->h->stream));  // kernels may still read the operator's buffers
    delete op;
    return 0;
}

int ct_env_update_left(ct_handle_t h, int prec, int dtype, const void* L, const void* A, const void* W, void* out,
                       int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wR) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(prec >= 0 && prec <= CT_PREC_BF16X6, 2, "unknown precision");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 3, "unknown dtype");
    CT_ARG(Dl > 0 && d > 0 && Dr > 0 && wL > 0 && wR > 0, 8, "dimensions must be > 0");
    return ct::env_left(h, prec, dtype, L, A, W, out, Dl, d, Dr, wL, wR);
}

int ct_env_update_right(ct_handle_t h, int prec, int dtype, const void* R, const void* A, const void* W, void* out,
                        int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wR) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(prec >= 0 && prec <= CT_PREC_BF16X6, 2, "unknown precision");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 3, "unknown dtype");
    CT_ARG(Dl > 0 && d > 0 && Dr > 0 && wL > 0 && wR > 0, 8, "dimensions must be > 0");
    return ct::env_right(h, prec, dtype, R, A, W, out, Dl, d, Dr, wL, wR);
}

int ct_two_site_theta(ct_handle_t h, int prec, int dtype, const void* A_i, const void* A_j, void* theta, int64_t Dl,
                      int64_t d, int64_t Dm, int64_t Dr) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(Dl > 0 && d > 0 && Dm > 0 && Dr > 0, 7, "dimensions must be > 0");
    ct::GemmDesc g;
    g.M = (int)(Dl * d); g.N = (int)(d * Dr); g.K = (int)Dm;
    g.a_kc = 1; g.sAm = Dm; g.sAk = 1;
    g.b_kc = 0; g.sBk = d * Dr; g.sBn = 1;
    g.ldc = d * Dr;
    return ct::gemm(h, prec, dtype, A_i, dtype, A_j, dtype, theta, g);
}

int ct_lanczos_two_site(ct_handle_t h, int prec, int dtype, const void* L, const void* Wi, const void* Wj,
                        const void* R, void* theta, int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wM,
                        int64_t wR, int max_iter, double tol, int max_rotations, int refine, double* E0,
                        int* num_iter, double* E_lanczos, double* residual_norm) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(prec >= 0 && prec <= CT_PREC_BF16X6, 2, "unknown precision");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 3, "unknown dtype");
    CT_ARG(Dl > 0 && d > 0 && Dr > 0 && wL > 0 && wM > 0 && wR > 0, 10, "dimensions must be > 0");
    CT_ARG(max_iter >= 1 && max_iter <= 64, 16, "max_iter must be in [1, 64]");
    CT_ARG(max_rotations >= -1, 18, "max_rotations must be >= -1");
    const int64_t nL = wL * Dl * Dl, nR = wR * Dr * Dr, nWi = wL * d * d * wM, nWj = wM * d * d * wR;
    const int64_t dim = Dl * d * d * Dr;
    const double *L64 = (const double*)L, *R64 = (const double*)R, *Wi64 = (const double*)Wi,
                 *Wj64 = (const double*)Wj;
    double* th64 = (double*)theta;
    if (dtype == CT_F32) {
        // float32 in and out, FP64 inside (Krylov basis, tridiagonal solve, refinement)
        void *bl, *br, *bw, *bt;
        CT_CHECK(ct::ws_get(h, ct::WS_BND0, (size_t)nL * 8, &bl));
        CT_CHECK(ct::ws_get(h, ct::WS_BND1, (size_t)nR * 8, &br));
        CT_CHECK(ct::ws_get(h, ct::WS_BND2, (size_t)(nWi + nWj) * 8, &bw));
        CT_CHECK(ct::ws_get(h, ct::WS_BND3, (size_t)dim * 8, &bt));
        CT_CHECK(ct::convert(h, CT_F32, L, CT_F64, bl, nL));
        CT_CHECK(ct::convert(h, CT_F32, R, CT_F64, br, nR));
        CT_CHECK(ct::convert(h, CT_F32, Wi, CT_F64, bw, nWi));
        CT_CHECK(ct::convert(h, CT_F32, Wj, CT_F64, (double*)bw + nWi, nWj));
        CT_CHECK(ct::convert(h, CT_F32, theta, CT_F64, bt, dim));
        L64 = (const double*)bl;
        R64 = (const double*)br;
        Wi64 = (const double*)bw;
        Wj64 = (const double*)bw + nWi;
        th64 = (double*)bt;
    }
    ct::HeffDims hd{Dl, Dl, d, Dr, Dr, wL, wM, wR};
    void* w2;
    CT_CHECK(ct::ws_get(h, ct::WS_W2, (size_t)(wL * d * d * d * d * wR) * 8, &w2));
    CT_CHECK(ct::heff_build_w2(h, CT_F64, Wi64, Wj64, wL, wM, wR, d, (double*)w2));
    ct::HeffEnv env, env64;
    CT_CHECK(ct::heff_prepare(h, prec, CT_F64, L64, R64, (const double*)w2, hd, &env));
    CT_CHECK(ct::heff_prepare(h, CT_PREC_FP64, CT_F64, L64, R64, (const double*)w2, hd, &env64));
    ct::LanczosResult r;
    CT_CHECK(ct::lanczos(h, env, &env64, th64, max_iter, tol, max_rotations, refine, &r));
    if (dtype == CT_F32) {
        CT_CHECK(ct::convert(h, CT_F64, th64, CT_F32, theta, dim));
        CT_HIP(hipStreamSynchronize(h->stream));
    }
    if (E0) *E0 = r.E0;
    if (num_iter) *num_iter = r.num_iter;
    if (E_lanczos) *E_lanczos = r.E_lanczos;
    if (residual_norm) *residual_norm = r.res_norm;
    return 0;
}

int ct_svd_trunc(ct_handle_t h, int dtype, const void* A, int64_t m, int64_t n, int64_t chi_max, double eps, void* U,
                 void* S, void* Vt, int64_t* chi, double* discarded_weight) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 2, "unknown dtype");
    CT_ARG(m >= 0 && n >= 0, 4, "dimensions must be >= 0");
    CT_ARG(chi_max >= 1, 6, "chi_max must be >= 1");
    ct::SvdResult r;
    if (dtype == CT_F64) {
        CT_CHECK(ct::svd_trunc(h, (const double*)A, m, n, chi_max, eps, (double*)U, (double*)S, (double*)Vt, &r));
    } else {
        // float32 in and out, the factorisation and truncation rule in FP64
        const int64_t k = std::min(std::min(m, n), chi_max);
        void *ba, *bu, *bs, *bv;
        CT_CHECK(ct::ws_get(h, ct::WS_BND0, (size_t)(m * n) * 8, &ba));
        CT_CHECK(ct::ws_get(h, ct::WS_BND1, (size_t)(m * k) * 8, &bu));
        CT_CHECK(ct::ws_get(h, ct::WS_BND2, (size_t)k * 8, &bs));
        CT_CHECK(ct::ws_get(h, ct::WS_BND3, (size_t)(k * n) * 8, &bv));
        CT_CHECK(ct::convert(h, CT_F32, A, CT_F64, ba, m * n));
        CT_CHECK(ct::svd_trunc(h, (const double*)ba, m, n, chi_max, eps, (double*)bu, (double*)bs, (double*)bv, &r));
        CT_CHECK(ct::convert(h, CT_F64, bu, CT_F32, U, m * r.chi));
        CT_CHECK(ct::convert(h, CT_F64, bs, CT_F32, S, r.chi));
        CT_CHECK(ct::convert(h, CT_F64, bv, CT_F32, Vt, r.chi * n));
        CT_HIP(hipStreamSynchronize(h->stream));
    }
    if (chi) *chi = r.chi;
    if (discarded_weight) *discarded_weight = r.discarded;
    return 0;
}

int ct_theta_split(ct_handle_t h, int svd_mode, int dtype, const void* theta, int64_t Dl, int64_t d, int64_t Dr,
                   int64_t chi_max, double eps, int left_to_right, void* A_i, void* A_j, int64_t* chi,
                   double* discarded_weight) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(svd_mode == CT_SVD_GESVD || svd_mode == CT_SVD_GRAM, 2, "unknown svd_mode");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 3, "unknown dtype");
    CT_ARG(Dl > 0 && d > 0 && Dr > 0, 5, "dimensions must be > 0");
    CT_ARG(chi_max >= 1, 8, "chi_max must be >= 1");
    const int64_t m = Dl * d, n = d * Dr;
    ct::SvdResult r;
    if (dtype == CT_F64) {
        CT_CHECK(ct::theta_split(h, svd_mode, (const double*)theta, m, n, chi_max, eps, left_to_right != 0,
                                 (double*)A_i, (double*)A_j, &r));
    } else {
        const int64_t k = std::min(std::min(m, n), chi_max);
        void *bt, *bi, *bj;
        CT_CHECK(ct::ws_get(h, ct::WS_BND0, (size_t)(m * n) * 8, &bt));
        CT_CHECK(ct::ws_get(h, ct::WS_BND1, (size_t)(m * k) * 8, &bi));
        CT_CHECK(ct::ws_get(h, ct::WS_BND2, (size_t)(k * n) * 8, &bj));
        CT_CHECK(ct::convert(h, CT_F32, theta, CT_F64, bt, m * n));
        CT_CHECK(ct::theta_split(h, svd_mode, (const double*)bt, m, n, chi_max, eps, left_to_right != 0, (double*)bi,
                                 (double*)bj, &r));
        CT_CHECK(ct::convert(h, CT_F64, bi, CT_F32, A_i, m * r.chi));
        CT_CHECK(ct::convert(h, CT_F64, bj, CT_F32, A_j, r.chi * n));
        CT_HIP(hipStreamSynchronize(h->stream));
    }
    if (chi) *chi = r.chi;
    if (discarded_weight) *discarded_weight = r.discarded;
    return 0;
}

int ct_sytrd(ct_handle_t h, int64_t n, double* A, double* d, double* e, double* tau) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(ct::sytrd_coop_supported(n), 2, "n must be in [2, 2048]");
    void* w;
    CT_CHECK(ct::ws_get(h, ct::WS_TMP4, 1024 * 8, &w));
    int* err = (int*)w;
    CT_HIP(hipMemsetAsync(err, 0, sizeof(int), h->stream));
    bool ok = false;
    CT_CHECK(ct::sytrd_coop(h, A, n, d, e, tau, err, &ok));
    if (!ok) {
        ct::set_last_error("tridiagonalization grid does not fit on the device");
        return CT_STATUS_NOT_SUPPORTED;
    }
    int herr = 0;
    CT_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipStreamSynchronize(h->stream));
    if (herr != 0) {
        ct::set_last_error("tridiagonalization: grid-wide wait timed out");
        return CT_STATUS_SVD_FAILED;
    }
    return 0;
}

int ct_syevd(ct_handle_t h, int64_t n, double* A, double* w, double* Z, int method) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(n >= 1, 2, "n must be >= 1");
    CT_ARG(method == CT_EIG_AUTO || method == CT_EIG_ROCSOLVER, 6, "unknown method");
    void* ws;
    CT_CHECK(ct::ws_get(h, ct::WS_TMP4, 1024 * 8, &ws));
    int* info = (int*)ws;
    CT_CHECK(ct::syevd(h, A, n, w, Z, info, method == CT_EIG_ROCSOLVER));
    int hinfo[2] = {0, 0};
    CT_HIP(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipStreamSynchronize(h->stream));
    if (hinfo[1] != 0) {
        ct::set_last_error("tridiagonalization: grid-wide wait timed out");
        return CT_STATUS_SVD_FAILED;
    }
    if (hinfo[0] != 0) {
        ct::set_last_error("symmetric eigensolver failed, info=" + std::to_string(hinfo[0]));
        return CT_STATUS_SVD_FAILED;
    }
    return 0;
}

int ct_heff_adjoint(ct_handle_t h, const double* L, const double* Wi, const double* Wj, const double* R,
                    const double* theta, double* gWi, double* gWj, int64_t Dl, int64_t d, int64_t Dr, int64_t wL,
                    int64_t wM, int64_t wR) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(Dl > 0 && d > 0 && Dr > 0 && wL > 0 && wM > 0 && wR > 0, 9, "dimensions must be > 0");
    ct::HeffDims hd{Dl, Dl, d, Dr, Dr, wL, wM, wR};
    return ct::heff_adjoint(h, L, Wi, Wj, R, theta, gWi, gWj, hd);
}

int ct_env_adjoint(ct_handle_t h, const double* L, const double* A, const double* R, double* gW, int64_t Dl,
                   int64_t d, int64_t Dr, int64_t wL, int64_t wR) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(Dl > 0 && d > 0 && Dr > 0 && wL > 0 && wR > 0, 6, "dimensions must be > 0");
    return ct::env_adjoint(h, L, A, R, gW, Dl, d, Dr, wL, wR);
}

}  // extern "C"
