/*
 * chemtensor_hip.h - C ABI of libdmrg_hip.so, the MI355X (gfx950) two-site
 * DMRG hot path: dense tensordot, the two-site effective-Hamiltonian matvec,
 * the left/right environment updates, the device-resident Lanczos solver,
 * the truncated two-site SVD and the sweep driver.
 *
 * The library is the sibling of the reference's only native component,
 * native/svd_shim.c, and follows its conventions (native/svd_shim.c:7-23):
 *   - plain extern "C" functions returning an int status;
 *   - the caller allocates every input and output buffer (here: device memory
 *     of the handle's device, row-major contiguous, int64_t dimensions);
 *   - the library never frees caller memory.
 *
 * Status: 0 = ok; -k = the k-th argument is invalid (LAPACK style, as the
 * reference's LAPACKE `info`); CT_STATUS_* > 0 for runtime failures.
 * ct_last_error() returns a message for the calling thread's last failure.
 *
 * All calls are asynchronous on the handle's stream unless they return a host
 * scalar (documented per call), in which case they synchronise that stream.
 * A handle is not thread-safe.
 */
#ifndef CHEMTENSOR_HIP_H
#define CHEMTENSOR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums ----------------------------------------------------------- */
/* compute precision of the contractions */
enum {
    CT_PREC_FP64 = 0,   /* FP64 MFMA, FP64 accumulate (the oracle-equivalent path) */
    CT_PREC_FP32 = 1,   /* FP32 MFMA, FP32 accumulate                              */
    CT_PREC_BF16X3 = 2, /* a*b ~ ah*bh + ah*bl + al*bh on BF16 MFMA, FP32 accumulate */
    CT_PREC_BF16X6 = 3  /* FP32-accurate on BF16 MFMA: x = h + m + l (three bf16 values
                         * carrying the FP32 significand), a*b ~ the six products with
                         * terms down to 2^-16 |ab|, FP32 accumulate; no counterpart in
                         * the reference (its arithmetic is FP32, which this matches) */
};
/* element type of buffers */
enum { CT_F64 = 0, CT_F32 = 1 };
/* two-site split: full SVD (rocSOLVER gesvd), or eigendecomposition of the
 * Gram matrix theta theta^T / theta^T theta (ct_syevd's path + 2 GEMMs) */
enum { CT_SVD_GESVD = 0, CT_SVD_GRAM = 1 };
/* ct_syevd methods: the library's path (cooperative tridiagonalization for
 * n <= 2048) or rocSOLVER dsyevd throughout (for comparison). */
enum { CT_EIG_AUTO = 0, CT_EIG_ROCSOLVER = 1 };

enum {
    CT_STATUS_OK = 0,
    CT_STATUS_NOT_SUPPORTED = 1,
    CT_STATUS_ZERO_NORM = 2,        /* Lanczos start vector has zero norm (dmrg.mojo:610-611) */
    CT_STATUS_EMPTY_KRYLOV = 3,     /* (dmrg.mojo:706-707) */
    CT_STATUS_SVD_FAILED = 4,       /* rocSOLVER info != 0 (src/mylinalg/svd.mojo:75-76) */
    CT_STATUS_ALLOC = 5,
    CT_STATUS_SHAPE = 6,            /* "Dim mismatch" / wrong C shape (dense_tensor.mojo:1163-1167, 1294-1299) */
    CT_STATUS_SCRATCH_LIMIT = 7,    /* an allocation-free call would need more than CT_NOALLOC_SCRATCH_BYTES */
    CT_STATUS_ROCBLAS_ERROR_BASE = 2000,
    CT_STATUS_HIP_ERROR_BASE = 1000
};

typedef struct ct_handle_s* ct_handle_t;
typedef struct ct_dmrg_s* ct_dmrg_t;

/* ---- handle ------------------------------------------------------------
 * Replaces the reference's DeviceContext for this path (src/algorithms/dmrg.mojo:378).
 * All work is enqueued on `stream`; NULL is the default (null) stream. */
int ct_handle_create(int device, void* stream, ct_handle_t* out);
int ct_handle_destroy(ct_handle_t h);
int ct_handle_set_stream(ct_handle_t h, void* stream);
int ct_synchronize(ct_handle_t h);
const char* ct_last_error(void);
const char* ct_version(void);

/* ---- dense tensor core ------------------------------------------------ */

/* dense_tensor_dot (src/m_tensor/dense_tensor.mojo:963-1331): contracts
 * ndim_mult leading (axrange=1) or trailing (axrange=0) axes of A and B
 * (aligned, not reversed); C = non-contracted A axes then B axes.  With
 * axrange_A == axrange_B == 0 the B side is inferred exactly as :1141-1159.
 * rankC == 0, or rankC == 1 with shapeC[0] == 1, accepts a scalar result.
 * dtype: element type of A, B and C. */
int ct_tensordot(ct_handle_t h, int prec, int dtype, const void* A, int rankA, const int64_t* shapeA,
                 const void* B, int rankB, const int64_t* shapeB, int ndim_mult, int axrange_A,
                 int axrange_B, void* C, int rankC, const int64_t* shapeC);

/* 3xBF16 building blocks (no reference counterpart: the reference GEMM is
 * float32 only, dense_tensor.mojo:1320-1328).  Operands are "split rows"
 * (uint16 BF16 bit patterns): row r starts at out + r*ld and holds, for each
 * 32-wide block of k, the 32 values hi = bf16(x) followed by the 32 values
 * lo = bf16(x - hi); k is zero-padded to a multiple of 32, and
 * ld >= 2*roundup(K, 32).  ct_split_bf16 writes the split rows of the matrix
 * X (element (r, k) at X[r*s_row + k*s_k], dtype CT_F64/CT_F32).
 * ct_gemm_bf16x3_nt computes C[m][n] = sum_k Ah Bh + Ah Bl + Al Bh over split
 * rows A [M] and B [N] on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; C
 * is CT_F64 or CT_F32 with row stride ldc.  ld must be a multiple of 8 and
 * the bases 16-byte aligned. */
int ct_split_bf16(ct_handle_t h, int dtype, const void* X, int64_t rows, int64_t K, int64_t s_row, int64_t s_k,
                  uint16_t* out, int64_t ld);
int ct_gemm_bf16x3_nt(ct_handle_t h, const uint16_t* A, int64_t lda, const uint16_t* B, int64_t ldb, int dtype_c,
                      void* C, int64_t M, int64_t N, int64_t K, int64_t ldc);
/* 6xBF16 building blocks (CT_PREC_BF16X6; no reference counterpart).  An
 * operand is the split rows above plus an "l plane": element (r, k) at
 * outl[r*ldl + k] holds l = bf16(x - hi - lo), zero for k >= K up to
 * roundup(K, 32); ldl >= roundup(K, 32), a multiple of 8.  ct_split_bf16x6
 * writes both.  ct_gemm_bf16x6_nt computes C[m][n] = sum_k of the six products
 * Al Bh + Ah Bl + Am Bm + Am Bh + Ah Bm + Ah Bh (h, m, l = hi, lo, l plane)
 * with FP32 accumulation: FP32 GEMM accuracy on the BF16 matrix cores. */
int ct_split_bf16x6(ct_handle_t h, int dtype, const void* X, int64_t rows, int64_t K, int64_t s_row, int64_t s_k,
                    uint16_t* out, int64_t ld, uint16_t* outl, int64_t ldl);
int ct_gemm_bf16x6_nt(ct_handle_t h, const uint16_t* A, int64_t lda, const uint16_t* Al, int64_t ldal,
                      const uint16_t* B, int64_t ldb, const uint16_t* Bl, int64_t ldbl, int dtype_c, void* C,
                      int64_t M, int64_t N, int64_t K, int64_t ldc);

/* DenseTensor.transpose (dense_tensor.mojo:460-545): out axis i = in axis perm[i]; rank <= 8. */
int ct_permute(ct_handle_t h, int dtype, const void* in, int rank, const int64_t* shape, const int* perm,
               void* out);

/* BLAS-1 of DenseTensor (dense_tensor.mojo:637-754), FP64 accumulation.
 * ct_norm_sq / ct_dot synchronise and return a host double. */
int ct_norm_sq(ct_handle_t h, int dtype, const void* x, int64_t n, double* result);
int ct_dot(ct_handle_t h, int dtype, const void* x, const void* y, int64_t n, double* result);
int ct_axpy(ct_handle_t h, int dtype, double alpha, const void* x, void* y, int64_t n);
int ct_scale(ct_handle_t h, int dtype, double alpha, void* x, int64_t n);
int ct_convert(ct_handle_t h, int dtype_in, const void* in, int dtype_out, void* out, int64_t n);

/* ---- two-site DMRG contractions ----------------------------------------- */

/* apply_two_site_heff (src/algorithms/dmrg.mojo:254-374):
 *   out[a',s1',s2',b'] = sum L[wl,a,a'] theta[a,s1,s2,b] Wi[wl,s1,s1',wm] Wj[wm,s2,s2',wr] R[wr,b,b']
 * L [wL,Dl,Dlb], R [wR,Dr,Drb], Wi [wL,d,d,wM], Wj [wM,d,d,wR], theta [Dl,d,d,Dr],
 * out [Dlb,d,d,Drb]; all of element type dtype. */
int ct_heff_apply(ct_handle_t h, int prec, int dtype, const void* L, const void* Wi, const void* Wj,
                  const void* R, const void* theta, void* out, int64_t Dl, int64_t Dlb, int64_t d, int64_t Dr,
                  int64_t Drb, int64_t wL, int64_t wM, int64_t wR);

/* Prepared H_eff (no reference counterpart: the prepared form of
 * apply_two_site_heff, dmrg.mojo:254-374, for callers that apply one local
 * H_eff many times, e.g. their own Lanczos-type eigensolver).  ct_heff_create
 * builds W2 = Wi.Wj and the precision-specific copies of L and R (FP32 copies
 * or 3xBF16 split rows) into buffers owned by the operator; with prec FP64 the
 * operator reads L and R in place, so they must stay valid and unchanged.
 * ct_heff_apply_prepared(op, theta, out) is ct_heff_apply with those
 * arguments (theta/out of the operator's dtype), on the handle's stream; the
 * handle must outlive the operator.  ct_heff_destroy synchronises the stream. */
typedef struct ct_heff_s* ct_heff_t;
int ct_heff_create(ct_handle_t h, int prec, int dtype, const void* L, const void* Wi, const void* Wj, const void* R,
                   int64_t Dl, int64_t Dlb, int64_t d, int64_t Dr, int64_t Drb, int64_t wL, int64_t wM, int64_t wR,
                   ct_heff_t* op);
int ct_heff_apply_prepared(ct_heff_t op, const void* theta, void* out);
int ct_heff_destroy(ct_heff_t op);

/* Allocation-free H_eff matvec.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Arguments and
 * result are exactly those of ct_heff_apply, for prec CT_PREC_FP64 or
 * CT_PREC_FP32 and w*d*d <= 64 on both MPO bonds; the call performs no
 * hipMalloc/hipFree and no allocating library routine.  Every intermediate
 * lives in one static device array of CT_NOALLOC_SCRATCH_BYTES per device
 * (held from the first ct_handle_create on that device until the process
 * ends); dimensions whose intermediates exceed it are rejected before any
 * work is enqueued (CT_STATUS_SCRATCH_LIMIT; about D <= 450 at w = 5 and
 * D <= 350 at w = 8 in FP64).
 * All calls on one device share that array: calls that overlap on one GPU
 * (two streams, two threads, or two handles) race and must be serialised by
 * the caller. */
#define CT_NOALLOC_SCRATCH_BYTES (64u << 20)
int ct_heff_apply_noalloc(ct_handle_t h, int prec, int dtype, const void* L, const void* Wi, const void* Wj,
                          const void* R, const void* theta, void* out, int64_t Dl, int64_t Dlb, int64_t d,
                          int64_t Dr, int64_t Drb, int64_t wL, int64_t wM, int64_t wR);

/* update_left_environment (src/state/environments.mojo:99-201):
 *   out[wR,Dr,Dr'] = sum L[wL,Dl,Dl'] A[Dl,s,Dr] W[wL,s,s',wR] A[Dl',s',Dr']
 * L [wL,Dl,Dl], A [Dl,d,Dr], W [wL,d,d,wR], out [wR,Dr,Dr]. */
int ct_env_update_left(ct_handle_t h, int prec, int dtype, const void* L, const void* A, const void* W,
                       void* out, int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wR);

/* update_right_environment (src/state/environments.mojo:309-417):
 *   out[wL,Dl,Dl'] = sum R[wR,Dr,Dr'] A[Dl,s,Dr] W[wL,s,s',wR] A[Dl',s',Dr']
 * R [wR,Dr,Dr], A [Dl,d,Dr], W [wL,d,d,wR], out [wL,Dl,Dl]. */
int ct_env_update_right(ct_handle_t h, int prec, int dtype, const void* R, const void* A, const void* W,
                        void* out, int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wR);

/* build_two_site_theta (src/algorithms/dmrg.mojo:93-140): theta[Dl,d,d,Dr] = A_i[Dl*d,Dm] @ A_j[Dm,d*Dr]. */
int ct_two_site_theta(ct_handle_t h, int prec, int dtype, const void* A_i, const void* A_j, void* theta,
                      int64_t Dl, int64_t d, int64_t Dm, int64_t Dr);

/* lanczos_two_site_optimize (src/algorithms/dmrg.mojo:486-756), FP64 Krylov
 * basis on the device, reference variant (no reorthogonalisation, breakdown
 * at beta < 100*dim*DBL_EPSILON, cyclic-max Jacobi on T with tolerance tol and
 * at most max_rotations rotations - the reference's cap is 200, dmrg.mojo:510 -;
 * max_rotations = -1 instead solves T exactly (Sturm multisection + inverse
 * iteration, the converged Jacobi result to rounding),
 * normalised FP64 Ritz vector).  theta ([Dl,d,d,Dr]) is the start
 * vector on entry and the Ritz vector on return.  dtype is the element type
 * of L, R, Wi, Wj and theta: CT_F64, or CT_F32 (the reference's float32
 * tensors, dmrg.mojo:486-495), which are widened to FP64 on entry and theta
 * rounded back on return; the Krylov basis, the tridiagonal solve and the
 * refinement stay FP64 either way.  prec selects the matvec precision; the FP32 and
 * 3xBF16 forms also stop once beta < 1e-5 times max |alpha|, the noise
 * level of their matvecs.  refine != 0 adds
 * an FP64 Rayleigh quotient of the Ritz vector (E_rq) and, refine == 2, one
 * residual-correction step in span{psi, H psi - E_rq psi}; E0 then returns the
 * refined energy.  With refinement of a mixed-precision solve, the first
 * Krylov matvec runs in FP64 and the start vector (energy alpha_0) is
 * returned whenever the refined energy would lie above it.  Synchronises
 * the stream.  num_iter = Krylov dimension. */
int ct_lanczos_two_site(ct_handle_t h, int prec, int dtype, const void* L, const void* Wi, const void* Wj,
                        const void* R, void* theta, int64_t Dl, int64_t d, int64_t Dr, int64_t wL,
                        int64_t wM, int64_t wR, int max_iter, double tol, int max_rotations, int refine,
                        double* E0, int* num_iter, double* E_lanczos, double* residual_norm);

/* dense_tensor_svd_trunc (src/m_tensor/dense_tensor.mojo:1483-1623) on the
 * device in FP64 (rocSOLVER): A [m,n] (preserved), U [m,chi], S [chi],
 * Vt [chi,n] with chi = smallest chi <= chi_max whose discarded weight
 * (sum s^2 - sum_{<chi} s^2)/sum s^2 <= eps, else min(m,n,chi_max)
 * (:1528-1548).  U/S/Vt must hold min(m,n,chi_max) columns/values/rows.
 * Also returns the discarded weight the reference computes and drops
 * (dmrg.mojo:188-200, :824).  dtype (CT_F64 / CT_F32) is the element type
 * of A, U, S and Vt; float32 operands (the reference's SVD input,
 * dense_tensor.mojo:1500-1502) are factorised in FP64 and the factors
 * rounded back.  Synchronises the stream. */
int ct_svd_trunc(ct_handle_t h, int dtype, const void* A, int64_t m, int64_t n, int64_t chi_max, double eps,
                 void* U, void* S, void* Vt, int64_t* chi, double* discarded_weight);

/* split_two_site_theta_svd (src/algorithms/dmrg.mojo:143-252): theta
 * [Dl,d,d,Dr] (preserved) -> A_i [Dl,d,chi], A_j [chi,d,Dr] with S
 * absorbed into A_j (left_to_right) or into A_i (right-to-left).  A_i / A_j
 * must hold min(Dl*d, d*Dr, chi_max) bond columns / rows.  svd_mode selects
 * CT_SVD_GESVD or CT_SVD_GRAM.  dtype (CT_F64 / CT_F32) is the element type
 * of theta, A_i and A_j; the split itself runs in FP64.  Synchronises the
 * stream. */
int ct_theta_split(ct_handle_t h, int svd_mode, int dtype, const void* theta, int64_t Dl, int64_t d, int64_t Dr,
                   int64_t chi_max, double eps, int left_to_right, void* A_i, void* A_j, int64_t* chi,
                   double* discarded_weight);

/* Symmetric eigensolver of the CT_SVD_GRAM split (no reference counterpart:
 * the reference's SVD is LAPACKE dgesdd through native/svd_shim.c:11-23).
 * ct_sytrd: Householder tridiagonalization of the symmetric n x n FP64 matrix
 * A (device, row stride n, both triangles; 2 <= n <= 2048) in one cooperative
 * kernel.  Output exactly as LAPACK dsytrd with uplo = 'L' on the column-major
 * view of A: d[n], e[n-1], tau[n-1]; row k of A holds d_k at column k, e_k at
 * column k+1 and the reflector v_k at columns k+2..n-1.  Synchronises.
 * ct_syevd: A (destroyed) -> ascending eigenvalues w[n] and eigenvector j in
 * row j of Z[n][n] (LAPACK's column-major eigenvector matrix).  method
 * CT_EIG_AUTO: ct_sytrd + rocSOLVER dstedc + a blocked back-transformation
 * for n <= 2048, rocSOLVER dsyevd above; CT_EIG_ROCSOLVER: rocSOLVER dsyevd.
 * Synchronises. */
int ct_sytrd(ct_handle_t h, int64_t n, double* A, double* d, double* e, double* tau);
int ct_syevd(ct_handle_t h, int64_t n, double* A, double* w, double* Z, int method);

/* ---- sweep driver ------------------------------------------------------
 * dmrg_two_site (src/algorithms/dmrg.mojo:377-483) with DMRGWorkspace
 * (:39-90) and dmrg_local_update_two_site (:759-844).  The MPS and MPO are
 * copied into library-owned device storage at create time; ct_dmrg_run may be
 * called repeatedly (it continues from the current MPS, rebuilding the right
 * environments as the reference does on every dmrg_two_site call). */
typedef struct {
    int num_sweeps;        /* DMRGParams.num_sweeps */
    int64_t chi_max;       /* DMRGParams.chi_max */
    double eps_trunc;      /* DMRGParams.eps_trunc */
    int max_krylov_iter;   /* DMRGParams.max_krylov_iter */
    double krylov_tol;     /* DMRGParams.krylov_tol (Jacobi tolerance) */
    double energy_tol;     /* DMRGParams.energy_tol */
    int jacobi_max_rot;    /* Jacobi rotation cap of the local eigensolve (reference: 200, dmrg.mojo:510);
                              -1: exact tridiagonal eigensolve (converged result) */
    int svd_mode;          /* CT_SVD_GESVD (reference factorisation) or CT_SVD_GRAM */
    int prec;              /* matvec precision CT_PREC_* */
    int refine;            /* 0: reference energies; 1: FP64 Rayleigh quotient; 2: + residual step */
    int env_prec;          /* precision of the environment updates (CT_PREC_*) */
    int verbose;
    int graphs;            /* 1: replay each bond's Lanczos solve (H_eff preparation, Krylov recurrence,
                              tridiagonal solve, Ritz vector, refinement) from a HIP graph cached per bond,
                              captured the second time the bond's shapes repeat; 0: launch every kernel.
                              Same kernels and arguments either way, so the results are identical.
                              No counterpart in the reference (launch overhead at small bond dimensions). */
} ct_dmrg_params;

typedef struct {
    double energy;             /* final energy (last local energy, dmrg.mojo:458-463) */
    int sweeps;                /* sweeps completed */
    int local_updates;
    double trunc_weight_sum;   /* sum over local updates of the discarded weight */
    double trunc_weight_max;
    int64_t max_bond;          /* largest bond dimension reached */
    double t_total;            /* seconds (host wall clock) */
    double t_lanczos;          /* seconds in the Lanczos solves (incl. matvecs) */
    double t_svd;              /* seconds in the SVD/truncation */
    double t_env;              /* seconds in environment updates */
    double t_theta;            /* seconds forming theta */
    int64_t matvecs;           /* number of H_eff applications */
    double matvec_flops;       /* algorithmic FLOPs of those matvecs (reference order) */
    double env_flops;
    int64_t graph_replays;     /* local solves replayed from a cached graph (ct_dmrg_params.graphs) */
    int64_t kept_start;        /* refined mixed-precision solves whose result lay above the start vector's
                                  FP64 energy, so the start vector was kept (no reference counterpart) */
} ct_dmrg_stats;

/* mpo_dims[nsites+1], mpo_sites[i]: device [w_i,d,d,w_{i+1}];
 * mps_dims[nsites+1], mps_sites[i]: device [D_i,d,D_{i+1}];
 * dtype (CT_F64 / CT_F32) is the element type of the given sites (float32 is
 * the reference's storage, dmrg.mojo:377-382).  The driver keeps FP64 copies;
 * params.prec selects the matvec precision. */
int ct_dmrg_create(ct_handle_t h, const ct_dmrg_params* params, int nsites, int64_t d, int dtype,
                   const int64_t* mpo_dims, const void* const* mpo_sites, const int64_t* mps_dims,
                   const void* const* mps_sites, ct_dmrg_t* out);
int ct_dmrg_destroy(ct_dmrg_t s);
/* Runs params.num_sweeps (or num_sweeps if > 0) sweeps; sweep_energies (host,
 * may be NULL) receives one energy per sweep run, and sweeps_run (may be
 * NULL) the number of sweeps this call ran (fewer than asked when
 * params.energy_tol stops it, dmrg.mojo:466-474).  Synchronises. */
int ct_dmrg_run(ct_dmrg_t s, int num_sweeps, double* energy, double* sweep_energies, int* sweeps_run);
int ct_dmrg_get_stats(ct_dmrg_t s, ct_dmrg_stats* stats);
/* current bond dimensions (nsites+1 values) */
int ct_dmrg_get_dims(ct_dmrg_t s, int64_t* dims);
/* copies site i ([D_i,d,D_{i+1}]) into a device buffer of element type dtype */
int ct_dmrg_get_site(ct_dmrg_t s, int i, int dtype, void* dst);

/* ---- adjoint (no reference counterpart: new capability) -----------------
 * dE/dW_i of E = <psi|H|psi> for a two-site wave function theta at sites
 * (i, i+1) with environments L [wL,Dl,Dl], R [wR,Dr,Dr]:
 *   gWi[wl,s1,s1',wm] = sum L[wl,a,a'] theta[a,s1,s2,b] Wj[wm,s2,s2',wr] R[wr,b,b'] theta[a',s1',s2',b']
 *   gWj[wm,s2,s2',wr] = sum L theta Wi R theta   (same network, Wj removed)
 * FP64.  Either gradient pointer may be NULL. */
int ct_heff_adjoint(ct_handle_t h, const double* L, const double* Wi, const double* Wj, const double* R,
                    const double* theta, double* gWi, double* gWj, int64_t Dl, int64_t d, int64_t Dr,
                    int64_t wL, int64_t wM, int64_t wR);

/* Single-site form: dE/dW of E = L . A W A . R for site tensor A [Dl,d,Dr],
 * L [wL,Dl,Dl], R [wR,Dr,Dr]:
 *   gW[wl,s,s',wr] = sum L[wl,a,a'] A[a,s,b] A[a',s',b'] R[wr,b,b']   (FP64) */
int ct_env_adjoint(ct_handle_t h, const double* L, const double* A, const double* R, double* gW, int64_t Dl,
                   int64_t d, int64_t Dr, int64_t wL, int64_t wR);

/* ---- host-side input construction ---------------------------------------
 * create_random_mps_c_compatible (src/tests/benchmarks/rng_c_compat.mojo:174-231):
 * PCG32-seeded bond dimensions and float32 Box-Muller entries (widened to
 * double).  Host buffers: dims[nsites+1]; data may be NULL to query the total
 * number of values (returned in *total). */
int ct_random_mps_c_compatible(int nsites, int64_t d, int64_t chi_max, uint64_t seed, int64_t* dims,
                               double* data, int64_t* total);

#ifdef __cplusplus
}
#endif
#endif /* CHEMTENSOR_HIP_H */
