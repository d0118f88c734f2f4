/*
 * FP64 CPU oracle for the two-site DMRG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is compiled into
 * oracle/liboracle.so and may be loaded only by tests/, by the smoke check and
 * by the CPU-baseline leg of bench.py.  The product (libdmrg_hip.so and the
 * chemtensor_mojo_amd package) never links, loads or calls it.
 *
 * It restates, in plain C and IEEE double precision, the algorithm of the
 * reference Mojo implementation (BatuhanGQskt/chemtensor_mojo).  Every function
 * cites the reference file:line it follows.  The reference itself computes its
 * device tensors in float32; this restatement keeps the same contraction order,
 * the same Lanczos variant (no reorthogonalisation, cyclic-max Jacobi on T),
 * the same truncation rule and the same sweep schedule, in FP64 throughout.
 *
 * BLAS/LAPACK: OpenBLAS dgemm + LAPACKE dgesdd (the scipy-bundled OpenBLAS,
 * exported with a "scipy_" prefix), exactly the library calls the reference
 * makes through native/svd_shim.c:11-23.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ BLAS */
extern void scipy_cblas_dgemm(int layout, int transa, int transb, int m, int n, int k,
                              double alpha, const double* a, int lda, const double* b, int ldb,
                              double beta, double* c, int ldc);
extern int scipy_LAPACKE_dgesdd(int layout, char jobz, int m, int n, double* a, int lda,
                                double* s, double* u, int ldu, double* vt, int ldvt);
extern void scipy_openblas_set_num_threads(int n);
extern int scipy_openblas_get_num_threads(void);

enum { ROW_MAJOR = 101, NO_TRANS = 111, TRANS = 112 };

void ora_set_num_threads(int n) { scipy_openblas_set_num_threads(n); }
int ora_get_num_threads(void) { return scipy_openblas_get_num_threads(); }

/* C[m,n] = A[m,k] @ B[k,n], all row-major contiguous. */
static void gemm_nn(int64_t m, int64_t n, int64_t k, const double* A, const double* B, double* C) {
    if (m == 0 || n == 0) return;
    if (k == 0) { memset(C, 0, sizeof(double) * (size_t)(m * n)); return; }
    scipy_cblas_dgemm(ROW_MAJOR, NO_TRANS, NO_TRANS, (int)m, (int)n, (int)k, 1.0, A, (int)k, B,
                      (int)n, 0.0, C, (int)n);
}

/* ------------------------------------------------------------- PCG32 RNG
 * Restates src/tests/benchmarks/rng_c_compat.mojo:14-149 (itself a clone of
 * chemtensor's pcg_basic.c / rng.c).                                       */
typedef struct { uint64_t state, inc; } ora_pcg32;
typedef struct { ora_pcg32 g0, g1; } ora_rng;

static const uint64_t PCG_MULT = 6364136223846793005ULL;

/* rng_c_compat.mojo:21-30 */
uint32_t ora_pcg32_random(ora_pcg32* r) {
    uint64_t old = r->state;
    r->state = old * PCG_MULT + r->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xorshifted >> rot) | (xorshifted << ((-(int32_t)rot) & 31));
}

/* rng_c_compat.mojo:49-56 (pcg32_srandom_r) */
void ora_pcg32_seed(ora_pcg32* r, uint64_t initstate, uint64_t initseq) {
    r->state = 0;
    r->inc = (initseq << 1) | 1u;
    ora_pcg32_random(r);
    r->state += initstate;
    ora_pcg32_random(r);
}

/* rng_c_compat.mojo:129-149 (seed_rng_state) */
void ora_seed_rng_state(ora_rng* s, uint64_t seed) {
    ora_pcg32 base;
    ora_pcg32_seed(&base, seed, PCG_MULT * seed + 1442695040888963407ULL);
    uint64_t v[8];
    for (int i = 0; i < 8; ++i) v[i] = ora_pcg32_random(&base);
    uint64_t seed1 = (v[0] << 32) | v[1];
    uint64_t seed2 = (v[2] << 32) | v[3];
    uint64_t seq1 = (v[4] << 32) | v[5];
    uint64_t seq2 = (v[6] << 32) | v[7];
    const uint64_t mask = 0x7FFFFFFFFFFFFFFFULL;
    if ((seq1 & mask) == (seq2 & mask)) seq2 = ~seq2;
    ora_pcg32_seed(&s->g0, seed1, seq1);
    ora_pcg32_seed(&s->g1, seed2, seq2);
}

/* rng_c_compat.mojo:65-67 */
uint32_t ora_rand_uint32(ora_rng* s) { return ora_pcg32_random(&s->g0); }

/* rng_c_compat.mojo:69-74 */
uint64_t ora_rand_uint64(ora_rng* s) {
    uint64_t hi = ora_pcg32_random(&s->g0);
    uint64_t lo = ora_pcg32_random(&s->g1);
    return (hi << 32) | lo;
}

/* rng_c_compat.mojo:76-84 (pcg32x2_boundedrand_r) */
uint64_t ora_rand_interval(ora_rng* s, uint64_t bound) {
    uint64_t threshold = (~bound + 1u) % bound;
    for (;;) {
        uint64_t r = ora_rand_uint64(s);
        if (r >= threshold) return r % bound;
    }
}

/* rng_c_compat.mojo:86-96 (RNG consumption of Floyd's rand_choice) */
void ora_rand_choice_advance(ora_rng* s, int64_t bound, int64_t num_samples) {
    for (int64_t i = 0; i < num_samples; ++i)
        (void)ora_rand_interval(s, (uint64_t)(bound - num_samples + 1 + i));
}

/* rng_c_compat.mojo:98-108: uniform floats in float32, Box-Muller in float32. */
float ora_randnf(ora_rng* s) {
    const float inv_2p32 = 2.32830644e-10f;
    float u1 = (float)ora_rand_uint32(s) * inv_2p32;
    float u2 = (float)ora_rand_uint32(s) * inv_2p32;
    if (u1 == 0.0f) u1 = 1.0f;
    return sqrtf(-2.0f * logf(u1)) * sinf(6.2831853071795864769f * u2);
}

/* rng_c_compat.mojo:190-214: bond dims of construct_random_mps, advancing the
 * RNG exactly as rand_choice does.  dims has nsites+1 entries. */
static void random_mps_dims(ora_rng* rng, int nsites, int d, int64_t chi_max, int64_t* dims) {
    for (int i = 0; i <= nsites; ++i) dims[i] = 0;
    dims[0] = 1;
    dims[nsites] = 1;
    int mid = (nsites + 1) / 2;
    for (int l = 1; l < mid; ++l) {
        int64_t full = dims[l - 1] * d;
        dims[l] = full < chi_max ? full : chi_max;
        if (full > chi_max) ora_rand_choice_advance(rng, full, chi_max);
    }
    for (int l = nsites - 1; l >= mid; --l) {
        int64_t full = dims[l + 1] * d;
        dims[l] = full < chi_max ? full : chi_max;
        if (full > chi_max) ora_rand_choice_advance(rng, full, chi_max);
    }
}

int64_t ora_random_mps_dims(int nsites, int d, int64_t chi_max, uint64_t seed, int64_t* dims) {
    ora_rng rng;
    ora_seed_rng_state(&rng, seed);
    random_mps_dims(&rng, nsites, d, chi_max, dims);
    int64_t total = 0;
    for (int i = 0; i < nsites; ++i) total += dims[i] * d * dims[i + 1];
    return total;
}

/* rng_c_compat.mojo:174-231 (create_random_mps_c_compatible): site i holds
 * [dims[i], d, dims[i+1]] entries randnf()*float32(1/sqrt(nelem)), in float32,
 * stored here widened to double.  Returns the number of values written. */
int64_t ora_random_mps_fill(int nsites, int d, int64_t chi_max, uint64_t seed, double* out) {
    ora_rng rng;
    ora_seed_rng_state(&rng, seed);
    int64_t* dims = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nsites + 1));
    random_mps_dims(&rng, nsites, d, chi_max, dims);
    int64_t off = 0;
    for (int i = 0; i < nsites; ++i) {
        int64_t nelem = dims[i] * d * dims[i + 1];
        float scale = (float)(1.0 / sqrt((double)nelem));
        for (int64_t k = 0; k < nelem; ++k) out[off++] = (double)(ora_randnf(&rng) * scale);
    }
    free(dims);
    return off;
}

/* ------------------------------------------------------- tensor utilities */
static int64_t prod(const int64_t* s, int n) {
    int64_t p = 1;
    for (int i = 0; i < n; ++i) p *= s[i];
    return p;
}

/* Physical permute: out[idx permuted] = in[idx]; out axis i = in axis perm[i].
 * Restates DenseTensor.transpose (src/m_tensor/dense_tensor.mojo:460-545). */
void ora_permute(const double* in, int rank, const int64_t* shape, const int* perm, double* out) {
    int64_t stride_in[8], new_shape[8], idx[8];
    int64_t n = prod(shape, rank);
    stride_in[rank - 1] = 1;
    for (int i = rank - 2; i >= 0; --i) stride_in[i] = stride_in[i + 1] * shape[i + 1];
    for (int i = 0; i < rank; ++i) { new_shape[i] = shape[perm[i]]; idx[i] = 0; }
    for (int64_t o = 0; o < n; ++o) {
        int64_t src = 0;
        for (int i = 0; i < rank; ++i) src += idx[i] * stride_in[perm[i]];
        out[o] = in[src];
        for (int i = rank - 1; i >= 0; --i) {
            if (++idx[i] < new_shape[i]) break;
            idx[i] = 0;
        }
    }
}

/* Axis inference of dense_tensor_dot (src/m_tensor/dense_tensor.mojo:1141-1159):
 * returns the effective axrange_B, or -1 on a dimension mismatch
 * (:1161-1167) / invalid ndim_mult (:1126-1127). */
static int tensordot_axes(int rA, const int64_t* sA, int rB, const int64_t* sB, int nm, int axA, int axB) {
    if (nm < 1 || rA < nm || rB < nm) return -1;
    int effB = axB;
    if (!axA && !axB) {
        int lead = 1, trail = 1;
        for (int i = 0; i < nm; ++i) {
            int64_t a = sA[rA - nm + i];
            if (a != sB[i]) lead = 0;
            if (a != sB[rB - nm + i]) trail = 0;
        }
        effB = lead ? 1 : (trail ? 0 : 1);
    }
    for (int i = 0; i < nm; ++i) {
        int64_t a = sA[(axA ? 0 : rA - nm) + i];
        int64_t b = sB[(effB ? 0 : rB - nm) + i];
        if (a != b) return -1;
    }
    return effB;
}

/* Output shape of dense_tensor_dot (dense_tensor.mojo:1170-1182, :1261-1299).
 * Returns output rank (0 for a scalar) or -1. */
int ora_tensordot_shape(int rA, const int64_t* sA, int rB, const int64_t* sB, int nm, int axA, int axB,
                        int64_t* sC) {
    int effB = tensordot_axes(rA, sA, rB, sB, nm, axA, axB);
    if (effB < 0) return -1;
    int r = 0;
    for (int i = axA ? nm : 0; i < (axA ? rA : rA - nm); ++i) sC[r++] = sA[i];
    for (int i = effB ? nm : 0; i < (effB ? rB : rB - nm); ++i) sC[r++] = sB[i];
    return r;
}

/* dense_tensor_dot (src/m_tensor/dense_tensor.mojo:963-1331): flatten to
 * [m,k]/[k,m] and [k,n]/[n,k], physically transpose when the contracted block
 * is leading in A or trailing in B (:1249-1257), then one GEMM (:1322). */
int ora_tensordot(const double* A, int rA, const int64_t* sA, const double* B, int rB, const int64_t* sB,
                  int nm, int axA, int axB, double* C) {
    int effB = tensordot_axes(rA, sA, rB, sB, nm, axA, axB);
    if (effB < 0) return -1;
    int64_t k = axA ? prod(sA, nm) : prod(sA + rA - nm, nm);
    int64_t m = prod(sA, rA) / (k ? k : 1);
    int64_t n = prod(sB, rB) / (k ? k : 1);
    const double* Am = A;
    const double* Bm = B;
    double* At = NULL;
    double* Bt = NULL;
    if (axA) {  /* A is [k, m] -> transpose to [m, k] */
        int64_t sh[2] = {k, m};
        int perm[2] = {1, 0};
        At = (double*)malloc(sizeof(double) * (size_t)(m * k + 1));
        ora_permute(A, 2, sh, perm, At);
        Am = At;
    }
    if (!effB) {  /* B is [n, k] -> transpose to [k, n] */
        int64_t sh[2] = {n, k};
        int perm[2] = {1, 0};
        Bt = (double*)malloc(sizeof(double) * (size_t)(n * k + 1));
        ora_permute(B, 2, sh, perm, Bt);
        Bm = Bt;
    }
    gemm_nn(m, n, k, Am, Bm, C);
    free(At);
    free(Bt);
    return 0;
}

/* Brute-force contraction restated from the reference's test oracle
 * _host_dot_contract_a (src/tests/m_tensor/test_dense_tensor_dot.mojo:25-106):
 * explicit multi-index loops, double accumulation, no BLAS. */
int ora_tensordot_brute(const double* A, int rA, const int64_t* sA, const double* B, int rB,
                        const int64_t* sB, int nm, int axA, int axB, double* C) {
    if (nm < 1 || rA < nm || rB < nm) return -1;
    int offA = axA ? 0 : rA - nm, offB = axB ? 0 : rB - nm;
    for (int i = 0; i < nm; ++i)
        if (sA[offA + i] != sB[offB + i]) return -1;
    int64_t oshape[16];
    int orank = 0;
    for (int i = axA ? nm : 0; i < (axA ? rA : rA - nm); ++i) oshape[orank++] = sA[i];
    for (int i = axB ? nm : 0; i < (axB ? rB : rB - nm); ++i) oshape[orank++] = sB[i];
    int64_t stA[8], stB[8];
    stA[rA - 1] = 1;
    for (int i = rA - 2; i >= 0; --i) stA[i] = stA[i + 1] * sA[i + 1];
    stB[rB - 1] = 1;
    for (int i = rB - 2; i >= 0; --i) stB[i] = stB[i + 1] * sB[i + 1];
    int64_t osize = prod(oshape, orank);
    int64_t ncon = prod(sA + offA, nm);
    int nAn = rA - nm, nBn = rB - nm;
    int64_t om[16], cm[8], ia[8], ib[8];
    for (int64_t o = 0; o < osize; ++o) {
        int64_t rem = o;
        for (int i = orank - 1; i >= 0; --i) { om[i] = rem % oshape[i]; rem /= oshape[i]; }
        double sum = 0.0;
        for (int64_t c = 0; c < ncon; ++c) {
            rem = c;
            for (int i = nm - 1; i >= 0; --i) { cm[i] = rem % sA[offA + i]; rem /= sA[offA + i]; }
            for (int i = 0; i < nAn; ++i) ia[axA ? nm + i : i] = om[i];
            for (int i = 0; i < nm; ++i) ia[offA + i] = cm[i];
            for (int i = 0; i < nBn; ++i) ib[axB ? nm + i : i] = om[nAn + i];
            for (int i = 0; i < nm; ++i) ib[offB + i] = cm[i];
            int64_t la = 0, lb = 0;
            for (int i = 0; i < rA; ++i) la += ia[i] * stA[i];
            for (int i = 0; i < rB; ++i) lb += ib[i] * stB[i];
            sum += A[la] * B[lb];
        }
        C[o] = sum;
    }
    return 0;
}

/* -------------------------------------------------- H_eff and environments */

/* apply_two_site_heff (src/algorithms/dmrg.mojo:254-374), same four steps:
 *  1  L[wL,Dl,Dlb] -> perm(1,0,2) -> [Dl, wL*Dlb]; contract leading axes with
 *     theta[Dl, d*d*Dr]                              -> T1[wL,Dlb,d,d,Dr]
 *  2  T1 perm(1,3,4,0,2) [Dlb*d*Dr, wL*d] x Wi[wL*d, d*wM] -> T2[Dlb,d,Dr,d,wM]
 *  3  T2 perm(0,2,3,4,1) [Dlb*Dr*d, wM*d] x Wj[wM*d, d*wR] -> T3[Dlb,Dr,d,d,wR]
 *  4  T3 perm(0,2,3,4,1) [Dlb*d*d, wR*Dr] x R[wR*Dr, Drb]  -> out[Dlb,d,d,Drb]
 * Dlb/Drb are the bra bond dimensions (L shape[2], R shape[2]). */
int ora_heff_apply(const double* L, const double* Wi, const double* Wj, const double* R,
                   const double* theta, double* out, int64_t Dl, int64_t Dlb, int64_t d, int64_t Dr,
                   int64_t Drb, int64_t wL, int64_t wM, int64_t wR) {
    int64_t n1 = wL * Dlb * d * d * Dr;
    int64_t n2 = Dlb * d * Dr * d * wM;
    int64_t n3 = Dlb * Dr * d * d * wR;
    size_t big = (size_t)(n1 > n2 ? (n1 > n3 ? n1 : n3) : (n2 > n3 ? n2 : n3));
    double* a = (double*)malloc(sizeof(double) * (big + 1));
    double* b = (double*)malloc(sizeof(double) * (big + 1));
    double* Lt = (double*)malloc(sizeof(double) * (size_t)(wL * Dl * Dlb + 1));
    /* step 1 */
    {
        int64_t sh[3] = {wL, Dl, Dlb};
        int p[3] = {1, 0, 2};
        ora_permute(L, 3, sh, p, Lt);
        int64_t sA[2] = {Dl, wL * Dlb}, sB[2] = {Dl, d * d * Dr};
        ora_tensordot(Lt, 2, sA, theta, 2, sB, 1, 1, 1, a); /* a = T1[wL,Dlb,d,d,Dr] */
    }
    /* step 2 */
    {
        int64_t sh[5] = {wL, Dlb, d, d, Dr};
        int p[5] = {1, 3, 4, 0, 2};
        ora_permute(a, 5, sh, p, b);
        gemm_nn(Dlb * d * Dr, d * wM, wL * d, b, Wi, a); /* a = T2[Dlb,d,Dr,d,wM] */
    }
    /* step 3 */
    {
        int64_t sh[5] = {Dlb, d, Dr, d, wM};
        int p[5] = {0, 2, 3, 4, 1};
        ora_permute(a, 5, sh, p, b);
        gemm_nn(Dlb * Dr * d, d * wR, wM * d, b, Wj, a); /* a = T3[Dlb,Dr,d,d,wR] */
    }
    /* step 4 */
    {
        int64_t sh[5] = {Dlb, Dr, d, d, wR};
        int p[5] = {0, 2, 3, 4, 1};
        ora_permute(a, 5, sh, p, b);
        gemm_nn(Dlb * d * d, Drb, wR * Dr, b, R, out);
    }
    free(a);
    free(b);
    free(Lt);
    return 0;
}

/* update_left_environment (src/state/environments.mojo:99-201):
 *  1  L perm(1,0,2) [Dl, wL*Dlb] . A[Dl, s*Dr] (leading x leading) -> T1[wL,Dlb,s,Dr]
 *  2  T1 perm(1,3,0,2) [Dlb*Dr, wL*s] x W[wL*s, s'*wR]             -> T2[Dlb,Dr,s',wR]
 *  3  T2 perm(1,3,0,2) [Dr*wR, Dlb*s'] x A[Dlb*s', Dr]             -> [Dr,wR,Dr]
 *  4  perm(1,0,2)                                                   -> L'[wR,Dr,Dr]
 * The bra tensor is the same A (real arithmetic, no conjugation), so Dlb == Dl. */
int ora_env_update_left(const double* L, const double* A, const double* W, double* out, int64_t Dl,
                        int64_t d, int64_t Dr, int64_t wL, int64_t wR) {
    int64_t n1 = wL * Dl * d * Dr, n2 = Dl * Dr * d * wR;
    size_t big = (size_t)(n1 > n2 ? n1 : n2);
    double* a = (double*)malloc(sizeof(double) * (big + 1));
    double* b = (double*)malloc(sizeof(double) * (big + 1));
    double* Lt = (double*)malloc(sizeof(double) * (size_t)(wL * Dl * Dl + 1));
    {
        int64_t sh[3] = {wL, Dl, Dl};
        int p[3] = {1, 0, 2};
        ora_permute(L, 3, sh, p, Lt);
        int64_t sA[2] = {Dl, wL * Dl}, sB[2] = {Dl, d * Dr};
        ora_tensordot(Lt, 2, sA, A, 2, sB, 1, 1, 1, a);
    }
    {
        int64_t sh[4] = {wL, Dl, d, Dr};
        int p[4] = {1, 3, 0, 2};
        ora_permute(a, 4, sh, p, b);
        gemm_nn(Dl * Dr, d * wR, wL * d, b, W, a);
    }
    {
        int64_t sh[4] = {Dl, Dr, d, wR};
        int p[4] = {1, 3, 0, 2};
        ora_permute(a, 4, sh, p, b);
        gemm_nn(Dr * wR, Dr, Dl * d, b, A, a);
    }
    {
        int64_t sh[3] = {Dr, wR, Dr};
        int p[3] = {1, 0, 2};
        ora_permute(a, 3, sh, p, out);
    }
    free(a);
    free(b);
    free(Lt);
    return 0;
}

/* update_right_environment (src/state/environments.mojo:309-417):
 *  1  R perm(1,0,2) [Dr, wR*Dr] . A perm(2,0,1) [Dr, Dl*s] (leading x leading) -> T1[wR,Dr,Dl,s]
 *  2  T1 perm(1,2,0,3) [Dr*Dl, wR*s] x (W perm(0,2,3,1) [wL*s', wR*s])^T      -> T2[Dr,Dl,wL,s']
 *  3  T2 perm(1,2,0,3) [Dl*wL, Dr*s'] x A perm(2,1,0) [Dr*s', Dl]             -> [Dl,wL,Dl]
 *  4  perm(1,0,2)                                                              -> R'[wL,Dl,Dl] */
int ora_env_update_right(const double* R, const double* A, const double* W, double* out, int64_t Dl,
                         int64_t d, int64_t Dr, int64_t wL, int64_t wR) {
    int64_t n1 = wR * Dr * Dl * d, n2 = Dr * Dl * wL * d;
    size_t big = (size_t)(n1 > n2 ? n1 : n2);
    if ((size_t)(wR * Dr * Dr) > big) big = (size_t)(wR * Dr * Dr);
    if ((size_t)(Dl * d * Dr) > big) big = (size_t)(Dl * d * Dr);
    double* a = (double*)malloc(sizeof(double) * (big + 1));
    double* b = (double*)malloc(sizeof(double) * (big + 1));
    double* c = (double*)malloc(sizeof(double) * (big + 1));
    double* Wp = (double*)malloc(sizeof(double) * (size_t)(wL * d * d * wR + 1));
    double* Wt = (double*)malloc(sizeof(double) * (size_t)(wL * d * d * wR + 1));
    {
        int64_t shR[3] = {wR, Dr, Dr};
        int pR[3] = {1, 0, 2};
        ora_permute(R, 3, shR, pR, b); /* [Dr, wR, Dr'] */
        int64_t shA[3] = {Dl, d, Dr};
        int pA[3] = {2, 0, 1};
        ora_permute(A, 3, shA, pA, c); /* [Dr, Dl, s] */
        int64_t sA[2] = {Dr, wR * Dr}, sB[2] = {Dr, Dl * d};
        ora_tensordot(b, 2, sA, c, 2, sB, 1, 1, 1, a); /* T1[wR,Dr',Dl,s] */
    }
    {
        int64_t sh[4] = {wR, Dr, Dl, d};
        int p[4] = {1, 2, 0, 3};
        ora_permute(a, 4, sh, p, b); /* [Dr', Dl, wR, s] */
        int64_t shW[4] = {wL, d, d, wR};
        int pW[4] = {0, 2, 3, 1};
        ora_permute(W, 4, shW, pW, Wp); /* [wL, s', wR, s] */
        int64_t shW2[2] = {wL * d, wR * d};
        int pW2[2] = {1, 0};
        ora_permute(Wp, 2, shW2, pW2, Wt); /* [wR*s, wL*s'] */
        gemm_nn(Dr * Dl, wL * d, wR * d, b, Wt, a); /* T2[Dr',Dl,wL,s'] */
    }
    {
        int64_t sh[4] = {Dr, Dl, wL, d};
        int p[4] = {1, 2, 0, 3};
        ora_permute(a, 4, sh, p, b); /* [Dl, wL, Dr', s'] */
        int64_t shA[3] = {Dl, d, Dr};
        int pA[3] = {2, 1, 0};
        ora_permute(A, 3, shA, pA, c); /* [Dr', s', Dl'] */
        gemm_nn(Dl * wL, Dl, Dr * d, b, c, a); /* [Dl, wL, Dl'] */
    }
    {
        int64_t sh[3] = {Dl, wL, Dl};
        int p[3] = {1, 0, 2};
        ora_permute(a, 3, sh, p, out);
    }
    free(a);
    free(b);
    free(c);
    free(Wp);
    free(Wt);
    return 0;
}

/* ------------------------------------------------------------ Lanczos */

/* _jacobi_smallest_eigpair (src/algorithms/dmrg.mojo:507-598): at most
 * max_rot single rotations, each at the largest off-diagonal element; stops
 * when that element is below tol.  A is n*n row-major and is destroyed.
 * Returns the smallest diagonal value and its normalised eigenvector. */
double ora_jacobi_smallest(double* A, int n, double tol, int max_rot, double* vec) {
    if (n == 1) { vec[0] = 1.0; return A[0]; }
    double* V = (double*)calloc((size_t)(n * n), sizeof(double));
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int it = 0; it < max_rot; ++it) {
        int p = 0, q = 1;
        double mx = fabs(A[1]);
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                double v = fabs(A[i * n + j]);
                if (v > mx) { mx = v; p = i; q = j; }
            }
        if (mx < tol) break;
        double app = A[p * n + p], aqq = A[q * n + q], apq = A[p * n + q];
        double phi = 0.5 * atan2(2.0 * apq, aqq - app);
        double c = cos(phi), s = sin(phi);
        for (int k = 0; k < n; ++k) {
            if (k == p || k == q) continue;
            double akp = A[k * n + p], akq = A[k * n + q];
            double nkp = c * akp - s * akq, nkq = s * akp + c * akq;
            A[k * n + p] = nkp; A[p * n + k] = nkp;
            A[k * n + q] = nkq; A[q * n + k] = nkq;
        }
        A[p * n + p] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
        A[q * n + q] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
        A[p * n + q] = 0.0;
        A[q * n + p] = 0.0;
        for (int k = 0; k < n; ++k) {
            double vp = V[k * n + p], vq = V[k * n + q];
            V[k * n + p] = c * vp - s * vq;
            V[k * n + q] = s * vp + c * vq;
        }
    }
    int mi = 0;
    double mv = A[0];
    for (int i = 1; i < n; ++i)
        if (A[i * n + i] < mv) { mv = A[i * n + i]; mi = i; }
    double nr = 0.0;
    for (int i = 0; i < n; ++i) { vec[i] = V[i * n + mi]; nr += vec[i] * vec[i]; }
    nr = sqrt(nr);
    if (nr > 0.0)
        for (int i = 0; i < n; ++i) vec[i] /= nr;
    free(V);
    return mv;
}

typedef struct {
    const double *L, *R, *Wi, *Wj;
    int64_t Dl, d, Dr, wL, wM, wR;
} heff_ctx;

static void heff(const heff_ctx* h, const double* v, double* w) {
    ora_heff_apply(h->L, h->Wi, h->Wj, h->R, v, w, h->Dl, h->Dl, h->d, h->Dr, h->Dr, h->wL, h->wM, h->wR);
}

/* lanczos_two_site_optimize (src/algorithms/dmrg.mojo:486-756): normalise
 * the start vector, max_iter-1 full steps without reorthogonalisation
 * (:637-689), breakdown when beta < 100*dim*DBL_EPSILON (:630-632, :671-673),
 * a final alpha-only matvec when no breakdown happened (:691-703), Jacobi on
 * the tridiagonal T (:709-722) with at most max_rot rotations (the reference
 * uses 200, :510), FP64 Ritz vector, normalised (:728-745).
 * theta is overwritten by the Ritz vector.  Returns 0 or -1 (zero start). */
int ora_lanczos_two_site(const double* L, const double* Wi, const double* Wj, const double* R,
                         double* theta, int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wM,
                         int64_t wR, int max_iter, double tol, int max_rot, double* E0_out, int* nK_out,
                         double* alpha_out, double* beta_out) {
    heff_ctx h = {L, R, Wi, Wj, Dl, d, Dr, wL, wM, wR};
    int64_t dim = Dl * d * d * Dr;
    double nrm = 0.0;
    for (int64_t i = 0; i < dim; ++i) nrm += theta[i] * theta[i];
    nrm = sqrt(nrm);
    if (nrm < 1e-14) return -1;
    double* K = (double*)malloc(sizeof(double) * (size_t)(dim * max_iter));
    double* w = (double*)malloc(sizeof(double) * (size_t)dim);
    double* alpha = (double*)calloc((size_t)max_iter, sizeof(double));
    double* beta = (double*)calloc((size_t)max_iter, sizeof(double));
    for (int64_t i = 0; i < dim; ++i) K[i] = theta[i] / nrm;
    const double beta_thr = 100.0 * (double)dim * 2.2204460492503131e-16;
    int numiter = max_iter;
    for (int j = 0; j < max_iter - 1; ++j) {
        const double* vj = K + (size_t)j * dim;
        heff(&h, vj, w);
        double a = 0.0;
        for (int64_t i = 0; i < dim; ++i) a += vj[i] * w[i];
        alpha[j] = a;
        for (int64_t i = 0; i < dim; ++i) w[i] -= a * vj[i];
        if (j > 0) {
            const double* vp = K + (size_t)(j - 1) * dim;
            double bp = beta[j - 1];
            for (int64_t i = 0; i < dim; ++i) w[i] -= bp * vp[i];
        }
        double b = 0.0;
        for (int64_t i = 0; i < dim; ++i) b += w[i] * w[i];
        b = sqrt(b);
        if (b < beta_thr) { numiter = j + 1; break; }
        beta[j] = b;
        double* vn = K + (size_t)(j + 1) * dim;
        for (int64_t i = 0; i < dim; ++i) vn[i] = w[i] / b;
    }
    if (numiter == max_iter) {
        const double* vj = K + (size_t)(max_iter - 1) * dim;
        heff(&h, vj, w);
        double a = 0.0;
        for (int64_t i = 0; i < dim; ++i) a += vj[i] * w[i];
        alpha[max_iter - 1] = a;
    }
    int nK = numiter;
    double* T = (double*)calloc((size_t)(nK * nK), sizeof(double));
    for (int i = 0; i < nK; ++i) T[i * nK + i] = alpha[i];
    for (int i = 0; i < nK - 1; ++i) { T[i * nK + i + 1] = beta[i]; T[(i + 1) * nK + i] = beta[i]; }
    double* c = (double*)malloc(sizeof(double) * (size_t)nK);
    double E0 = ora_jacobi_smallest(T, nK, tol, max_rot, c);
    for (int64_t i = 0; i < dim; ++i) theta[i] = 0.0;
    for (int k = 0; k < nK; ++k) {
        const double* vk = K + (size_t)k * dim;
        for (int64_t i = 0; i < dim; ++i) theta[i] += c[k] * vk[i];
    }
    double on = 0.0;
    for (int64_t i = 0; i < dim; ++i) on += theta[i] * theta[i];
    on = sqrt(on);
    if (on > 0.0)
        for (int64_t i = 0; i < dim; ++i) theta[i] /= on;
    if (E0_out) *E0_out = E0;
    if (nK_out) *nK_out = nK;
    if (alpha_out) memcpy(alpha_out, alpha, sizeof(double) * (size_t)nK);
    if (beta_out && nK > 1) memcpy(beta_out, beta, sizeof(double) * (size_t)(nK - 1));
    free(K); free(w); free(alpha); free(beta); free(T); free(c);
    return 0;
}

/* ------------------------------------------------------ truncated SVD */

/* dense_tensor_svd_trunc_lapack_f64 (src/m_tensor/dense_tensor.mojo:1483-1592):
 * dgesdd jobz='S' on the row-major [m,n] matrix, then chi = smallest
 * chi <= chi_max with (sum s^2 - sum_{<chi} s^2)/sum s^2 <= eps, else
 * min(k, chi_max); eps <= 0 keeps min(k, chi_max) (:1528-1548).
 * U [m,chi], S [chi], Vt [chi,n] are written to caller buffers sized for
 * chi = min(m,n,chi_max).  A is not modified.  Returns chi (>=0) or
 * -(100+info) on a LAPACK failure. */
int ora_svd_trunc(const double* A, int64_t m, int64_t n, int64_t chi_max, double eps, double* U,
                  double* S, double* Vt) {
    int64_t k = m < n ? m : n;
    double* a = (double*)malloc(sizeof(double) * (size_t)(m * n));
    double* s = (double*)malloc(sizeof(double) * (size_t)k);
    double* u = (double*)malloc(sizeof(double) * (size_t)(m * k));
    double* vt = (double*)malloc(sizeof(double) * (size_t)(k * n));
    memcpy(a, A, sizeof(double) * (size_t)(m * n));
    int info = scipy_LAPACKE_dgesdd(ROW_MAJOR, 'S', (int)m, (int)n, a, (int)n, s, u, (int)k, vt, (int)n);
    if (info != 0) { free(a); free(s); free(u); free(vt); return -(100 + info); }
    int64_t chi = k < chi_max ? k : chi_max;
    if (eps > 0.0 && k > 0) {
        double tot = 0.0;
        for (int64_t i = 0; i < k; ++i) tot += s[i] * s[i];
        if (tot <= 0.0) chi = 0;
        else {
            double kept = 0.0;
            int64_t lim = k < chi_max ? k : chi_max;
            chi = lim;
            for (int64_t t = 1; t <= lim; ++t) {
                kept += s[t - 1] * s[t - 1];
                if ((tot - kept) / tot <= eps) { chi = t; break; }
            }
        }
    }
    for (int64_t i = 0; i < chi; ++i) S[i] = s[i];
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < chi; ++j) U[i * chi + j] = u[i * k + j];
    for (int64_t i = 0; i < chi; ++i)
        for (int64_t j = 0; j < n; ++j) Vt[i * n + j] = vt[i * n + j];
    free(a); free(s); free(u); free(vt);
    return (int)chi;
}

/* ------------------------------------------------------- DMRG driver */

typedef struct {
    int num_sweeps;
    int64_t chi_max;
    double eps_trunc;
    int max_krylov_iter;
    double krylov_tol;
    double energy_tol;
    int jacobi_max_rot; /* 200 in the reference (dmrg.mojo:510) */
} ora_dmrg_params; /* DMRGParams, src/algorithms/dmrg.mojo:22-36 (two_site/verbose omitted) */

/* dmrg_two_site (src/algorithms/dmrg.mojo:377-483) with DMRGWorkspace
 * (:39-90), dmrg_local_update_two_site (:759-844), build_two_site_theta
 * (:93-140), split_two_site_theta_svd (:143-252) and
 * build_right_environments (src/state/environments.mojo:420-472).
 *
 *  nsites, d            chain length, physical dimension
 *  mps_dims[nsites+1]   in: initial bond dims; out: final bond dims
 *  mps_in               concatenated initial sites [Dl,d,Dr]
 *  mpo_dims[nsites+1]   MPO bond dims; mpo concatenated sites [wL,d,d,wR]
 *  mps_out              concatenated final sites (capacity mps_out_cap values)
 *  sweep_E[num_sweeps]  last local eigenvalue of each completed sweep
 *  trunc_w              sum over all local updates of the discarded weight
 * Returns the number of sweeps run (>0) or a negative error. */
int ora_dmrg_two_site(int nsites, int64_t d, int64_t* mps_dims, const double* mps_in,
                      const int64_t* mpo_dims, const double* mpo, const ora_dmrg_params* prm,
                      double* mps_out, int64_t mps_out_cap, double* sweep_E, double* trunc_w,
                      double* E_final) {
    int N = nsites;
    if (N < 2) return -2;
    double** site = (double**)calloc((size_t)N, sizeof(double*));
    int64_t* D = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N + 1));
    const double** W = (const double**)malloc(sizeof(double*) * (size_t)N);
    for (int i = 0; i <= N; ++i) D[i] = mps_dims[i];
    {
        int64_t off = 0, woff = 0;
        for (int i = 0; i < N; ++i) {
            int64_t n = D[i] * d * D[i + 1];
            site[i] = (double*)malloc(sizeof(double) * (size_t)n);
            memcpy(site[i], mps_in + off, sizeof(double) * (size_t)n);
            off += n;
            W[i] = mpo + woff;
            woff += mpo_dims[i] * d * d * mpo_dims[i + 1];
        }
    }
    double** Lenv = (double**)calloc((size_t)(N + 1), sizeof(double*));
    double** Renv = (double**)calloc((size_t)(N + 1), sizeof(double*));
    int64_t* LD = (int64_t*)calloc((size_t)(N + 1), sizeof(int64_t));
    int64_t* RD = (int64_t*)calloc((size_t)(N + 1), sizeof(int64_t));
    /* R[N] = identity [wR, Dr, Dr] (environments.mojo:441-461) */
    {
        int64_t w = mpo_dims[N], Dn = D[N];
        Renv[N] = (double*)calloc((size_t)(w * Dn * Dn), sizeof(double));
        for (int64_t a = 0; a < w; ++a)
            for (int64_t b = 0; b < Dn; ++b) Renv[N][a * Dn * Dn + b * Dn + b] = 1.0;
        RD[N] = Dn;
        for (int i = N - 1; i >= 0; --i) {
            Renv[i] = (double*)malloc(sizeof(double) * (size_t)(mpo_dims[i] * D[i] * D[i]));
            ora_env_update_right(Renv[i + 1], site[i], W[i], Renv[i], D[i], d, D[i + 1], mpo_dims[i],
                                 mpo_dims[i + 1]);
            RD[i] = D[i];
        }
    }
    /* L[0] = identity [wL, Dl, Dl] (dmrg.mojo:70-90) */
    {
        int64_t w = mpo_dims[0], D0 = D[0];
        Lenv[0] = (double*)calloc((size_t)(w * D0 * D0), sizeof(double));
        for (int64_t a = 0; a < w; ++a)
            for (int64_t b = 0; b < D0; ++b) Lenv[0][a * D0 * D0 + b * D0 + b] = 1.0;
        LD[0] = D0;
    }
    double energy = 0.0, prev = 0.0, tw = 0.0;
    int sweeps_done = 0;
    int rc = 0;
    for (int sw = 0; sw < prm->num_sweeps && rc == 0; ++sw) {
        for (int dir = 0; dir < 2 && rc == 0; ++dir) {
            int l2r = dir == 0;
            for (int step = 0; step < N - 1; ++step) {
                int i = l2r ? step : N - 2 - step;
                int64_t Dl = D[i], Dm = D[i + 1], Dr = D[i + 2];
                int64_t m = Dl * d, n = d * Dr;
                /* build_two_site_theta: [Dl*d, Dm] @ [Dm, d*Dr] */
                double* th = (double*)malloc(sizeof(double) * (size_t)(m * n));
                gemm_nn(m, n, Dm, site[i], site[i + 1], th);
                double E;
                int nK;
                if (ora_lanczos_two_site(Lenv[i], W[i], W[i + 1], Renv[i + 2], th, Dl, d, Dr, mpo_dims[i],
                                         mpo_dims[i + 1], mpo_dims[i + 2], prm->max_krylov_iter,
                                         prm->krylov_tol, prm->jacobi_max_rot, &E, &nK, NULL, NULL) != 0) {
                    free(th);
                    rc = -3;
                    break;
                }
                /* split_two_site_theta_svd */
                double tot = 0.0;
                for (int64_t q = 0; q < m * n; ++q) tot += th[q] * th[q];
                int64_t kmax = m < n ? m : n;
                if (kmax > prm->chi_max) kmax = prm->chi_max;
                double* U = (double*)malloc(sizeof(double) * (size_t)(m * kmax));
                double* S = (double*)malloc(sizeof(double) * (size_t)kmax);
                double* Vt = (double*)malloc(sizeof(double) * (size_t)(kmax * n));
                int chi = ora_svd_trunc(th, m, n, prm->chi_max, prm->eps_trunc, U, S, Vt);
                free(th);
                if (chi <= 0) { free(U); free(S); free(Vt); rc = -4; break; }
                double kept = 0.0;
                for (int q = 0; q < chi; ++q) kept += S[q] * S[q];
                if (tot > 0.0) tw += (tot - kept) / tot;
                if (l2r) {
                    for (int q = 0; q < chi; ++q)
                        for (int64_t j = 0; j < n; ++j) Vt[q * n + j] *= S[q];
                } else {
                    for (int64_t r = 0; r < m; ++r)
                        for (int q = 0; q < chi; ++q) U[r * chi + q] *= S[q];
                }
                free(site[i]);
                free(site[i + 1]);
                site[i] = U;
                site[i + 1] = Vt;
                free(S);
                D[i + 1] = chi;
                if (l2r) {
                    free(Lenv[i + 1]);
                    Lenv[i + 1] = (double*)malloc(sizeof(double) * (size_t)(mpo_dims[i + 1] * chi * chi));
                    ora_env_update_left(Lenv[i], site[i], W[i], Lenv[i + 1], Dl, d, chi, mpo_dims[i],
                                        mpo_dims[i + 1]);
                    LD[i + 1] = chi;
                } else {
                    free(Renv[i + 1]);
                    Renv[i + 1] = (double*)malloc(sizeof(double) * (size_t)(mpo_dims[i + 1] * chi * chi));
                    ora_env_update_right(Renv[i + 2], site[i + 1], W[i + 1], Renv[i + 1], chi, d, Dr,
                                         mpo_dims[i + 1], mpo_dims[i + 2]);
                    RD[i + 1] = chi;
                }
                energy = E;
            }
        }
        if (rc != 0) break;
        if (sweep_E) sweep_E[sw] = energy;
        sweeps_done = sw + 1;
        if (sw > 0 && fabs(energy - prev) < prm->energy_tol) break;
        prev = energy;
    }
    if (rc == 0) {
        int64_t off = 0;
        for (int i = 0; i < N; ++i) off += D[i] * d * D[i + 1];
        if (off > mps_out_cap) rc = -5;
        else {
            off = 0;
            for (int i = 0; i < N; ++i) {
                int64_t n = D[i] * d * D[i + 1];
                memcpy(mps_out + off, site[i], sizeof(double) * (size_t)n);
                off += n;
            }
            for (int i = 0; i <= N; ++i) mps_dims[i] = D[i];
        }
    }
    if (trunc_w) *trunc_w = tw;
    if (E_final) *E_final = energy;
    for (int i = 0; i < N; ++i) free(site[i]);
    for (int i = 0; i <= N; ++i) { free(Lenv[i]); free(Renv[i]); }
    free(site); free(D); free(W); free(Lenv); free(Renv); free(LD); free(RD);
    return rc == 0 ? sweeps_done : rc;
}
