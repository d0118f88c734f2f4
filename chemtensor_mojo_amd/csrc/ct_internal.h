// Internal declarations of the MI355X (gfx950) two-site DMRG library.
// Everything here is private to libdmrg_hip.so; the public C ABI is
// include/chemtensor_hip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/chemtensor_hip.h"

namespace ct {

// Status codes (public values are in chemtensor_hip.h).
struct Error {
    int code;
    std::string msg;
};

void set_last_error(const std::string& msg);

#define CT_HIP(expr)                                                                    \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            ::ct::set_last_error(std::string(#expr) + ": " + hipGetErrorString(_e));     \
            return CT_STATUS_HIP_ERROR_BASE + (int)_e;                                  \
        }                                                                               \
    } while (0)

// Expression form of CT_HIP: returns 0 or the library status for e.
inline int hip_rc(hipError_t e) {
    if (e == hipSuccess) return 0;
    set_last_error(std::string("HIP error: ") + hipGetErrorString(e));
    return CT_STATUS_HIP_ERROR_BASE + (int)e;
}

#define CT_CHECK(expr)                  \
    do {                                \
        int _s = (expr);                \
        if (_s != 0) return _s;         \
    } while (0)

#define CT_ARG(cond, k, msg)                                          \
    do {                                                              \
        if (!(cond)) {                                                \
            ::ct::set_last_error(std::string("argument ") + #k + ": " + (msg)); \
            return -(k);                                              \
        }                                                             \
    } while (0)

// Device scratch buffer owned by a handle.  Grows on demand (after a stream
// sync), never shrinks until the handle is destroyed.
struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Workspaces carved from a fixed device region instead of hipMalloc: while a
// handle's `arena` is set, ws_get returns the region reserved for the slot
// (planned before any launch) and never allocates; a request beyond the
// reservation is an error.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    size_t off[48] = {};
    size_t len[48] = {};
};
// Reserves `bytes` (256-byte aligned) for `slot` in the arena; false when the
// arena is full.
bool arena_reserve(Arena& a, int slot, size_t bytes);

struct Handle {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    rocblas_handle blas = nullptr;
    // named workspaces (H_eff intermediates, reductions, SVD, Lanczos basis ...)
    Buffer ws[48];
    Arena* arena = nullptr;       // set only for the duration of an allocation-free call
    double* d_scalars = nullptr;  // small FP64 scratch for reductions (pinned-size)
    double* h_pinned = nullptr;   // pinned host mirror for scalar read-back
    int num_cu = 256;
    // bumped whenever a workspace is reallocated: device pointers recorded
    // before (e.g. in a captured graph) are stale once it changes
    uint64_t ws_gen = 0;
};

enum WsSlot {
    WS_T1 = 0,      // first GEMM output
    WS_T2 = 1,      // MPO-mixing output
    WS_CONV_A = 2,  // operand conversions (mixed precision)
    WS_CONV_B = 3,
    WS_PARTIAL = 4, // reduction partial sums
    WS_KRYLOV = 5,  // Lanczos basis
    WS_W = 6,       // Lanczos work vector
    WS_SVD = 7,     // SVD copies
    WS_SVD2 = 8,
    WS_SVD3 = 9,
    WS_TMP = 10,
    WS_TMP2 = 11,
    WS_MIX = 12,    // mixing coefficient matrix
    WS_LZ = 13,     // Lanczos device state
    WS_TMP3 = 14,
    WS_TMP4 = 15,
    WS_LSPLIT = 16,  // prepared (converted / split) left environment
    WS_RSPLIT = 17,  // prepared right environment
    WS_VSPLIT = 18,  // split Krylov vector (3xBF16 matvec) / its FP32 copy (FP32 matvec)
    WS_USPLIT = 19,  // split mixing output (3xBF16 matvec)
    WS_ENV_T1 = 20,  // environment update intermediates
    WS_ENV_T2 = 21,
    WS_ENV_CONV = 22,
    WS_W2 = 23,      // fused MPO coefficients
    WS_SPLITK = 24,  // split-K partial sums of the 3xBF16 GEMM
    WS_W2PAD64 = 25, // zero-padded fused MPO coefficients, FP64 (FP64 H_eff)
    WS_W2PAD32 = 26, // zero-padded fused MPO coefficients, FP32 (FP32 / 3xBF16 H_eff)
    WS_EIG = 27,     // tridiagonalization exchange buffers and arrival flags
    WS_EIG2 = 28,    // tridiagonal off-diagonal and reflector scalars
    WS_EIG3 = 29,    // back-transformation: reflector rows Vt
    WS_EIG4 = 30,    // back-transformation: T^T Vt per block
    WS_EIG5 = 31,    // back-transformation: V^T V and T per block
    WS_MIXA = 32,    // large-w mixing: packed input
    WS_MIXB = 33,    // large-w mixing: GEMM output
    WS_MIXC = 34,    // large-w mixing: split coefficients (3xBF16)
    WS_BND0 = 35,    // FP64 copies of float32 arguments at the C-ABI boundary
    WS_BND1 = 36,
    WS_BND2 = 37,
    WS_BND3 = 38,
    WS_LSPLIT_L = 39,  // l planes of the 6xBF16 operands: left environment,
    WS_RSPLIT_L = 40,  // right environment,
    WS_VSPLIT_L = 41,  // Krylov vector,
    WS_USPLIT_L = 42,  // mixing output,
    WS_CONV_AL = 43,   // generic operand conversions
    WS_CONV_BL = 44,
    WS_MIXA_L = 45,    // large-w mixing, 6xBF16: l planes of the split input
    WS_MIXC_L = 46,    // and of the split coefficients
    WS_COUNT = 47
};

int ws_get(Handle* h, int slot, size_t bytes, void** out);
// Same growth policy for a buffer owned by another object (synchronises the
// handle's stream before a reallocation).
int buf_get(Handle* h, Buffer& b, size_t bytes, void** out);
void buf_free(Buffer& b);

// ---------------------------------------------------------------- GEMM
// C[b][m, n] = alpha * sum_k A[b](m, k) * B[b](k, n) + beta * C[b][m, n]
// A element (m, k) at A + b*bsA + m*sAm + k*sAk; exactly one of sAm, sAk is 1.
// B element (k, n) at B + b*bsB + k*sBk + n*sBn; exactly one of sBk, sBn is 1.
// C element (m, n) at C + b*bsC + m*ldc + n.
struct GemmDesc {
    int M = 0, N = 0, K = 0, batch = 1;
    int a_kc = 0;  // 1: A stored [m][k] (sAk == 1); 0: A stored [k][m] (sAm == 1)
    int b_kc = 0;  // 1: B stored [n][k] (sBk == 1); 0: B stored [k][n] (sBn == 1)
    int64_t sAm = 0, sAk = 0, sBk = 0, sBn = 0, ldc = 0;
    int64_t bsA = 0, bsB = 0, bsC = 0;
    double alpha = 1.0, beta = 0.0;
    const int* active = nullptr;  // optional device flag: kernel is a no-op when *active == 0
};

// dtype codes for operands: CT_F64 / CT_F32
int gemm(Handle* h, int prec, int tA, const void* A, int tB, const void* B, int tC, void* C,
         const GemmDesc& g);
// Bytes of split-K partial sums (workspace WS_SPLITK) that gemm() requests for
// g at precision CT_PREC_FP64 or CT_PREC_FP32 (0 when it does not split K).
size_t gemm_splitk_bytes(int prec, const GemmDesc& g);

// 3xBF16 operand in "split rows": row r of batch b starts at p + b*bs + r*ld;
// for each 32-wide block kb of the contraction index it holds 32 hi = bf16(x)
// values followed by the 32 lo = bf16(x - hi) values (one 128-byte line), with
// k >= K zero-filled up to split_k32(K).  ld >= 2*split_k32(K).
// A 6xBF16 operand adds the "l plane": element (row, k) at l + b*bsl + row*ldl + k
// holds bf16(x - hi - lo), so that hi + lo + l carries every bit of the FP32
// value; ldl >= split_k32(K), zero-filled like the split rows.
struct SplitOp {
    const uint16_t* p = nullptr;
    int64_t ld = 0, bs = 0;
    const uint16_t* l = nullptr;
    int64_t ldl = 0, bsl = 0;
};
inline int64_t split_k32(int64_t K) { return (K + 31) / 32 * 32; }
// Splits X (element (row, k) of batch b at X + b*bsX + row*s_row + k*s_k) into
// split rows at out (ld_out, bs_out), zero-filling k in [K, split_k32(K)).
// With outl, the l plane (ldl, bsl) is written as well.
int split_bf16(Handle* h, int tX, const void* X, int rows, int K, int64_t s_row, int64_t s_k, int batch,
               int64_t bsX, uint16_t* out, int64_t ld_out, int64_t bs_out, const int* active,
               uint16_t* outl = nullptr, int64_t ldl = 0, int64_t bsl = 0);
// Same with rows contiguous (s_row == 1) and a two-level k: element k of a row
// at X + (k / kin)*s_kc + (k % kin)*s_k.
int split_bf16_kk(Handle* h, int tX, const void* X, int rows, int K, int kin, int64_t s_k, int64_t s_kc, int batch,
                  int64_t bsX, uint16_t* out, int64_t ld_out, int64_t bs_out, const int* active,
                  uint16_t* outl = nullptr, int64_t ldl = 0, int64_t bsl = 0);
// Zeroes k in [K, split_k32(K)) of every row of a split-rows operand (and of
// its l plane when outl is given).
int split_zero_pad(Handle* h, uint16_t* out, int64_t rows, int64_t K, int64_t ld, uint16_t* outl = nullptr,
                   int64_t ldl = 0);
// C[b][m][n] = alpha * sum_k (Ah Bh + Ah Bl + Al Bh)[m,k][n,k] + beta * C
int gemm_bf16x3_nt(Handle* h, const SplitOp& A, const SplitOp& B, int tC, void* C, int M, int N, int K,
                   int64_t ldc, int64_t bsC, int batch, double alpha, double beta, const int* active);
// 6xBF16 (both operands carry their l planes):
// C = alpha * sum_k (Al Bh + Ah Bl + Am Bm + Am Bh + Ah Bm + Ah Bh) + beta * C,
// with h, m, l = the split rows' hi, lo and the l plane.
int gemm_bf16x6_nt(Handle* h, const SplitOp& A, const SplitOp& B, int tC, void* C, int M, int N, int K,
                   int64_t ldc, int64_t bsC, int batch, double alpha, double beta, const int* active);
// Generic 3xBF16 / 6xBF16 GEMM on GemmDesc operands (splits both into workspace).
int gemm_bf16x3(Handle* h, int tA, const void* A, int tB, const void* B, int tC, void* C, const GemmDesc& g);
int gemm_bf16x6(Handle* h, int tA, const void* A, int tB, const void* B, int tC, void* C, const GemmDesc& g);

// ------------------------------------------------------------- mixing
// Per pixel (p, q): Y[co, p, eo, q] (+)= sum_{ci, ei} coef[(ci*EIN+ei)*NOUT + co*EOUT+eo] * X[ci, p, ei, q]
// with element strides given explicitly (q stride 1 on both sides).
struct MixDesc {
    int CIN = 0, EIN = 0, COUT = 0, EOUT = 0;
    int64_t P = 0, Q = 0;
    int64_t sXc = 0, sXp = 0, sXe = 0;
    int64_t sYc = 0, sYp = 0, sYe = 0;
    const int* active = nullptr;
};
int mix(Handle* h, int tX, const void* X, int tY, void* Y, const double* d_coef, const MixDesc& m,
        const void* padded = nullptr);
// Writes the zero-padded KM x KM copy of the [K][NO] coefficients that the
// mixing kernel reads (accumulation type tAcc: CT_F64 for FP64 input, CT_F32
// otherwise) into workspace `slot`; pass it to mix / mix_split as `padded` to
// skip the per-call padding.
int mix_pad_coef(Handle* h, const double* d_coef, int K, int NO, int tAcc, void* dst);
size_t mix_pad_bytes(int K, int NO, int tAcc);
// Same, FP32 input, output written as 3xBF16 split rows.
// Output in split rows: m.sYp / m.sYe give the row offset (a multiple of the
// row stride) of pixel p / output channel element eo, m.sYc the k offset of
// output channel co; q adds to k.
// With Yl, the l planes of a 6xBF16 operand are written too: the l value of
// split-rows element (row offset o, k) goes to Yl[o / 2 + k], i.e. the l plane
// has half the split rows' row stride.
int mix_split(Handle* h, const float* X, uint16_t* Y, const double* d_coef, const MixDesc& m,
              const void* padded = nullptr, uint16_t* Yl = nullptr);

// ------------------------------------------------------------- misc kernels
int permute(Handle* h, int dtype, const void* in, int rank, const int64_t* shape, const int* perm, void* out);
// out = (tout) in, n elements; with `active`, a no-op on the device when *active == 0
int convert(Handle* h, int tin, const void* in, int tout, void* out, int64_t n, const int* active = nullptr);
int fill(Handle* h, int dtype, void* x, int64_t n, double value);
int set_identity_env(Handle* h, int dtype, void* x, int64_t w, int64_t D);
// reductions into a device FP64 scalar (no host sync)
int dot_dev(Handle* h, int dtype, const void* x, const void* y, int64_t n, double* d_out, const int* active);
// host-returning reductions (synchronise the stream)
int dot_host(Handle* h, int dtype, const void* x, const void* y, int64_t n, double* out);
int axpy(Handle* h, int dtype, double alpha, const void* x, void* y, int64_t n);
int scale(Handle* h, int dtype, double alpha, void* x, int64_t n);

// ---------------------------------------------------------- contractions
struct HeffDims {
    int64_t Dl, Dlb, d, Dr, Drb, wL, wM, wR;
};

// Environments prepared for one precision: FP64/FP32 keep plain pointers
// (tEnv = element type); 3xBF16 / 6xBF16 hold the split, k-contiguous planes
// (6xBF16 with their l planes)
//   Ls: [wL][Dlb][Dl]  (A operand of L^T theta)
//   Rs: [Drb][wR*Dr]   (B operand of U R)
struct HeffEnv {
    int prec = CT_PREC_FP64;
    int tEnv = CT_F64;
    const void* L = nullptr;
    const void* R = nullptr;
    SplitOp Ls, Rs;
    const double* W2 = nullptr;  // fused MPO coefficients (device, FP64)
    const void* W2pad = nullptr; // W2 zero-padded in the mixing accumulation type
    HeffDims dims{};
};

// Builds the fused two-site MPO coefficients
//   coef[((wl*d + s1)*d + s2) * (d*d*wR) + wr*d*d + s1'*d + s2'] = sum_wm Wi[wl,s1,s1',wm] Wj[wm,s2,s2',wr]
// into d_out (device FP64, wL*d*d*d*d*wR values).  W tensors are FP64 or FP32 (tW).
int heff_build_w2(Handle* h, int tW, const void* Wi, const void* Wj, int64_t wL, int64_t wM, int64_t wR, int64_t d,
                  double* d_out);
// Prepares env for prec.  For 3xBF16 the split planes are written to the
// WS_LSPLIT / WS_RSPLIT workspaces (valid until the next prepare).  For FP32
// with FP64 input the converted copies go to the same workspaces.
// Buffers owned by a prepared H_eff operator (ct_heff_create) instead of the
// handle's workspaces, so the operator survives other calls on the handle.
struct HeffStore {
    Buffer w2, pad, l, r, ll, rl;
    ~HeffStore() { buf_free(w2); buf_free(pad); buf_free(l); buf_free(r); buf_free(ll); buf_free(rl); }
};
int heff_prepare(Handle* h, int prec, int tEnv, const void* L, const void* R, const double* d_W2,
                 const HeffDims& hd, HeffEnv* env, HeffStore* store = nullptr);
// out = H_eff theta.  tVec: element type of theta and out.
int heff_apply(Handle* h, const HeffEnv& env, int tVec, const void* theta, void* out, const int* active);
// The two GEMMs of heff_apply's FP64 / FP32 path: step 1 (T1 = L^T theta,
// batched over wL) and step 4 (out = U R).
GemmDesc heff_step1_desc(const HeffDims& D);
GemmDesc heff_step4_desc(const HeffDims& D);

// Resolves the device address of the allocation-free entry point's static
// scratch array on `device` (which must be current); called once per device
// by ct_handle_create.
int noalloc_scratch_init(int device);

int env_left(Handle* h, int prec, int dtype, const void* L, const void* A, const void* W, void* out,
             int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wR);
int env_right(Handle* h, int prec, int dtype, const void* R, const void* A, const void* W, void* out,
              int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wR);
int tensordot(Handle* h, int prec, int dtype, const void* A, int rankA, const int64_t* shapeA, const void* B,
              int rankB, const int64_t* shapeB, int ndim_mult, int axrange_A, int axrange_B, void* C, int rankC,
              const int64_t* shapeC);
int env_adjoint(Handle* h, const double* L, const double* A, const double* R, double* gW, int64_t Dl, int64_t d,
                int64_t Dr, int64_t wL, int64_t wR);
int heff_adjoint(Handle* h, const double* L, const double* Wi, const double* Wj, const double* R,
                 const double* theta, double* gWi, double* gWj, const HeffDims& hd);

// Lanczos (device resident) ------------------------------------------------
struct LanczosResult {
    double E0 = 0.0;        // reported energy (refined when refine != 0)
    double E_lanczos = 0.0; // lowest Ritz value of T (reference definition)
    double E_rq = 0.0;      // FP64 Rayleigh quotient <psi|H|psi> (refinement)
    double res_norm = 0.0;  // ||H psi - E_rq psi|| (refinement)
    int num_iter = 0;
    int matvecs = 0;        // H_eff applications (Krylov + refinement)
    int kept_start = 0;     // the refined result lay above the start vector's energy: v0 returned
};
// env: environments prepared for the Krylov matvecs; env64: FP64 environments
// for the refinement matvec (only used when refine != 0).
int lanczos(Handle* h, const HeffEnv& env, const HeffEnv* env64, double* d_theta, int max_iter, double tol,
            int max_rot, int refine, LanczosResult* res);
// The same solve split in two: lanczos_enqueue only enqueues kernels on
// h->stream (no allocation once the shapes have been seen, no host
// synchronisation, so it can be captured into a graph); lanczos_finish reads
// the device state back (synchronises) and fills *res.
int lanczos_enqueue(Handle* h, const HeffEnv& env, const HeffEnv* env64, double* d_theta, int max_iter, double tol,
            int max_rot, int refine);
int lanczos_finish(Handle* h, int refine, bool have_env64, LanczosResult* res);

// SVD / truncation --------------------------------------------------------
struct SvdResult {
    int64_t chi = 0;
    double discarded = 0.0;  // (|theta|^2 - sum kept s^2) / |theta|^2
    double norm_sq = 0.0;
};
// A: [m, n] row-major FP64 (device, preserved).  U: [m, chi] row-major, S[chi], Vt[chi, n].
// split_two_site_theta_svd: A_i, A_j with S absorbed (into A_j when l2r).
int theta_split(Handle* h, int mode, const double* theta, int64_t m, int64_t n, int64_t chi_max, double eps, bool l2r,
                double* Ai, double* Aj, SvdResult* r);
// Gram-matrix split writing the site tensors with S absorbed (see ct_svd.hip).
int svd_split_gram(Handle* h, const double* theta, int64_t m, int64_t n, int64_t chi_max, double eps, bool l2r,
                   double* Ai, double* Aj, SvdResult* r);
int svd_trunc(Handle* h, const double* A, int64_t m, int64_t n, int64_t chi_max, double eps, double* U,
              double* S, double* Vt, SvdResult* r);

// Symmetric eigensolver (ct_eig.hip) --------------------------------------
// Householder tridiagonalization of the symmetric n x n FP64 matrix A (row
// stride n, both triangles) in one cooperative kernel, output in LAPACK
// dsytrd's lower format (column-major view): d[n], e[n-1], tau[n-1], row k of
// A holds v_k at columns k+2.. .  *ok = false (nothing launched) when n is
// outside [2, 2048] or the grid cannot be co-resident.  *d_err (device) is
// set to 1 if a grid-wide wait timed out.
bool sytrd_coop_supported(int64_t n);
int sytrd_coop(Handle* h, double* A, int64_t n, double* d, double* e, double* tau, int* d_err, bool* ok);
// Eigendecomposition of the symmetric n x n matrix A (destroyed): ascending
// eigenvalues w[n]; eigenvector j in row j of Z[n][n].  info[0] (device):
// rocSOLVER info; info[1]: tridiagonalization wait timeout.  Both are read by
// the caller after its next stream synchronisation.
// force_rocsolver: rocSOLVER dsyevd regardless of n (comparison runs).
int syevd(Handle* h, double* A, int64_t n, double* w, double* Z, int* info, bool force_rocsolver = false);
// The same in two stages, so that a caller can read the eigenvalues first and
// back-transform only the eigenvectors it keeps: stage 1 leaves the
// eigenvectors of the tridiagonal matrix in Z (final ones when *staged is
// false); stage 2 turns rows n-rows..n-1 of Z into eigenvectors of A.
int syevd_stage1(Handle* h, double* A, int64_t n, double* w, double* Z, int* info, bool* staged,
                 bool force_rocsolver = false);
int syevd_stage2(Handle* h, int64_t n, double* Z, int64_t rows);

}  // namespace ct

// The public opaque handle type is the internal handle.
struct ct_handle_s : public ct::Handle {};
