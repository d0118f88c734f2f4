// Device-resident Lanczos for the two-site local problem.
//
// Restates lanczos_two_site_optimize (reference src/algorithms/dmrg.mojo:486-756)
// with the Krylov basis, the recurrence coefficients and the tridiagonal
// eigensolve on the GPU in FP64.  The host enqueues every kernel up front and
// synchronises once at the end: the breakdown test beta < 100*dim*DBL_EPSILON
// (:630-632, :671-673) clears a device flag that turns all later kernels of
// the recurrence into no-ops, so no per-iteration read-back is needed (the
// reference copies every Krylov vector device->host and back, :640-689).
//
// Mixed-precision matvecs (no reference counterpart) carry a relative error
// far above DBL_EPSILON, so for them the breakdown test also fires once
// beta_j falls below 1e-5 times max_i |alpha_i| (FP32 and 3xBF16; their
// matvec errors reach 5e-6 and 8e-6 at the largest MPO bonds): past that
// point new Krylov directions are matvec noise.
// The FP64 path keeps the reference test unchanged.
//
// Optional refinement for mixed-precision matvecs (no reference counterpart):
//   refine >= 1: E_rq = <psi|H psi> with an FP64 matvec of the normalised
//                Ritz vector psi; r = H psi - E_rq psi.
//   refine >= 2: one Rayleigh-Ritz step in span{psi, r/|r|}, whose 2x2
//                projected matrix [[E_rq, |r|], [|r|, <r^|H r^>]] uses one more
//                matvec at the Krylov precision; psi and the energy are updated.
// With refinement the first Krylov matvec (H v0) also runs in FP64, so
// alpha_0 = <v0|H v0> is the start vector's energy to FP64 rounding.  A
// mixed-precision Krylov space can yield a Ritz vector whose true energy lies
// above its start's (the recurrence coefficients carry the matvec error);
// when the refined energy exceeds alpha_0 the solve returns v0 with energy
// alpha_0 instead, so no local update raises the energy.  When the Krylov
// space breaks down after one vector the Ritz vector is v0 itself and the
// saved FP64 product H v0 replaces the refinement's FP64 matvec.
#include "ct_internal.h"
#include "ct_device.h"

namespace ct {

namespace {

constexpr int MAXK = 256;

struct LzState {
    double alpha[MAXK];
    double beta[MAXK];
    double coef[MAXK];
    double scal[16];  // 0: |theta|^2, 1: E0 (Jacobi), 2: |ritz|^2, 3: E_rq, 4: |r|^2, 5: h_rr, 6: E_ref
    int active;
    int numiter;
    int err;
    int exact_start;  // H v0 ran in FP64 (refinement of a mixed-precision solve)
    int refmv;        // the refinement's FP64 matvec of psi is needed (numiter > 1)
    int kept_start;   // the refined result lay above alpha_0: v0 was returned
    int pad[2];
};

constexpr int NB = 512;  // reduction blocks

__global__ void lz_init_kernel(LzState* s, int max_iter, int exact_start) {
    s->active = 1;
    s->numiter = max_iter;
    s->err = 0;
    s->exact_start = exact_start;
    s->refmv = 1;
    s->kept_start = 0;
    for (int i = 0; i < MAXK; ++i) { s->alpha[i] = 0.0; s->beta[i] = 0.0; s->coef[i] = 0.0; }
    for (int i = 0; i < 16; ++i) s->scal[i] = 0.0;
}

// partial sums of x.y  (y == nullptr: x.x)
__global__ __launch_bounds__(256) void lz_dot_partial(const double* __restrict__ x, const double* __restrict__ y,
                                                      int64_t n, double* __restrict__ part, const LzState* s,
                                                      int gated) {
    if (gated && !s->active) return;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = x[i];
        acc += a * (y ? y[i] : a);
    }
    acc = block_sum256(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Sum of the NB partial sums, in the same order in every block of every
// kernel, so that every block gets the same bits.
__device__ __forceinline__ double lz_sum_partials(const double* __restrict__ part) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < NB; i += blockDim.x) acc += part[i];
    return block_sum256(acc);
}

// what: 0 -> scal[idx] = sum; 1 -> alpha[idx] = sum
__global__ __launch_bounds__(256) void lz_reduce(const double* __restrict__ part, LzState* s, int what, int idx,
                                                 int gated) {
    if (gated && !s->active) return;
    const double acc = lz_sum_partials(part);
    if (threadIdx.x != 0) return;
    if (what == 0)
        s->scal[idx] = acc;
    else
        s->alpha[idx] = acc;
}

// v0 = theta / |theta|; zero start vector -> err (dmrg.mojo:605-613)
__global__ __launch_bounds__(256) void lz_start(const double* __restrict__ theta, double* __restrict__ v0, int64_t n,
                                                LzState* s) {
    const double nrm = sqrt(s->scal[0]);
    if (nrm < 1e-14) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { s->err = 1; s->active = 0; }
        return;
    }
    const double inv = 1.0 / nrm;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v0[i] = theta[i] * inv;
}

// alpha_j = v_j.w from the partial sums of lz_dot_partial (block 0 records
// it), then w -= alpha_j v_j + beta_{j-1} v_{j-1} and partial sums of w.w
// (dmrg.mojo:649-669).  Every block forms alpha_j itself from the partials,
// which saves a reduction launch per Krylov step.
__global__ __launch_bounds__(256) void lz_update_alpha(double* __restrict__ w, const double* __restrict__ vj,
                                                       const double* __restrict__ vjm1, int64_t n, int j, LzState* s,
                                                       const double* __restrict__ part_in,
                                                       double* __restrict__ part_out) {
    if (!s->active) return;
    const double a = lz_sum_partials(part_in);
    if (blockIdx.x == 0 && threadIdx.x == 0) s->alpha[j] = a;
    const double b = j > 0 ? s->beta[j - 1] : 0.0;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double x = w[i] - a * vj[i];
        if (j > 0) x -= b * vjm1[i];
        w[i] = x;
        acc += x * x;
    }
    acc = block_sum256(acc);
    if (threadIdx.x == 0) part_out[blockIdx.x] = acc;
}

// beta_j = |w| from the partial sums of lz_update_alpha, the breakdown test
// (beta_j below thr, or below floor_rel * max|alpha_i| for the mixed-precision
// forms: the Krylov space ends, numiter = j + 1), else v_{j+1} = w / beta_j
// (dmrg.mojo:671-681).  Every block forms beta_j and takes the same decision;
// block 0 records it.  The state's active flag is read and written with
// atomics: block 0 may clear it while other blocks of this launch start.
__global__ __launch_bounds__(256) void lz_next_beta(const double* __restrict__ w, double* __restrict__ vn, int64_t n,
                                                    int j, LzState* s, const double* __restrict__ part, double thr,
                                                    double floor_rel) {
    if (__hip_atomic_load(&s->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    const double b = sqrt(lz_sum_partials(part));
    double t = thr;
    if (floor_rel > 0.0) {
        double scale = 0.0;
        for (int i = 0; i <= j; ++i) scale = fmax(scale, fabs(s->alpha[i]));
        t = fmax(t, floor_rel * scale);
    }
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (b < t) {
        if (lead) {
            s->numiter = j + 1;
            __hip_atomic_store(&s->active, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }
    if (lead) s->beta[j] = b;
    const double inv = 1.0 / b;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        vn[i] = w[i] * inv;
}

// Cyclic-max Jacobi on the tridiagonal T (dmrg.mojo:507-598, :709-722), one
// wavefront.  Each of at most max_rot rotations zeroes the largest
// off-diagonal element (first in row-major order on ties, like the serial
// scan), stopping when it is below tol.  Writes E0 = smallest diagonal entry
// (first on ties) and the normalised eigenvector into coef.
//
// Lane k owns row k.  Before every rotation each lane scans its row's strict
// upper part with NMAX independent (fully unrolled, predicated) LDS loads and
// keeps the first maximum.  The pivot is then the largest row maximum, the
// lowest row on ties (a 64-bit max over the non-negative values' bit patterns
// followed by a ballot) - the same pivot as a serial row-major scan of the
// whole triangle.  Rows are stored with an odd stride (ld = n | 1) so that a
// column access by 32 lanes touches 32 distinct bank pairs.
// The rotation angle phi = atan2(2 apq, aqq - app) / 2 in (-pi/2, pi/2] is
// taken through the half-angle identities: with r = |(aqq - app, 2 apq)|,
// cos 2phi = (aqq - app) / r and sin 2phi = 2 apq / r; c = cos phi >= 0 and
// s = sin phi follow from whichever of (1 + cos 2phi) / 2, (1 - cos 2phi) / 2
// is not a cancellation, and the other from sin 2phi = 2 s c.
// Lane k reads columns k+1 .. k+NMAX-1 of its row with NMAX-1 unconditional
// loads (columns past n re-read column n-1; rows past n read unused LDS) and
// selects branch-free, so the loads issue back to back and are waited for
// once.  The per-lane bound n - k passes through an empty asm each call: the
// column masks are then recomputed as vector compares in every rotation
// instead of being hoisted out of the rotation loop as SGPR masks.
template <int NMAX>
__device__ __forceinline__ void row_scan(const double* A, int n, int ld, int k, double& m, int& arg) {
    int nk = n - k;
    asm volatile("" : "+v"(nk));
    double v[NMAX];
#pragma unroll
    for (int t = 1; t < NMAX; ++t) v[t] = A[k * ld + min(k + t, n - 1)];
    m = -1.0;
    arg = 0x7fffffff;
#pragma unroll
    for (int t = 1; t < NMAX; ++t) {
        const double a = t < nk ? fabs(v[t]) : -1.0;
        const bool g = a > m;
        m = g ? a : m;
        arg = g ? k + t : arg;
    }
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

template <int NMAX>
__device__ void jacobi_solve(double* A, double* V, int n, int ld, double tol, int max_rot) {
    const int lane = threadIdx.x;
    for (int it = 0; it < max_rot; ++it) {
        double m = -1.0;
        int rarg = 0x7fffffff;
        row_scan<NMAX>(A, n, ld, lane, m, rarg);  // lanes >= n: every column masked
        // m >= 0 on rows with an upper part; -1 (sign bit set) elsewhere -> key 0
        const unsigned long long key = m >= 0.0 ? (unsigned long long)__double_as_longlong(m) : 0ull;
        const unsigned long long kmax = wave_max_u64(key);
        const double best = __longlong_as_double((long long)kmax);
        if (kmax == 0ull || best < tol) break;
        const unsigned long long hit = __ballot(key == kmax);
        const int p = __ffsll((long long)hit) - 1;  // lowest row holding the maximum
        const int q = __shfl(rarg, p, 64);
        const double app = A[p * ld + p], aqq = A[q * ld + q], apq = A[p * ld + q];
        const double x = aqq - app, y = 2.0 * apq;
        const double r = sqrt(x * x + y * y);
        double c = 1.0, sn = 0.0;
        if (r > 0.0) {
            const double c2 = x / r, s2 = y / r;
            if (c2 >= 0.0) {
                c = sqrt(0.5 * (1.0 + c2));
                sn = s2 / (2.0 * c);
            } else {
                sn = copysign(sqrt(0.5 * (1.0 - c2)), y);
                c = s2 / (2.0 * sn);
            }
        }
        __syncthreads();
        for (int k = lane; k < n; k += 64) {
            if (k != p && k != q) {
                const double akp = A[k * ld + p], akq = A[k * ld + q];
                const double nkp = c * akp - sn * akq, nkq = sn * akp + c * akq;
                A[k * ld + p] = nkp; A[p * ld + k] = nkp;
                A[k * ld + q] = nkq; A[q * ld + k] = nkq;
            }
            const double vp = V[k * ld + p], vq = V[k * ld + q];
            V[k * ld + p] = c * vp - sn * vq;
            V[k * ld + q] = sn * vp + c * vq;
        }
        __syncthreads();
        if (lane == 0) {
            A[p * ld + p] = c * c * app - 2.0 * sn * c * apq + sn * sn * aqq;
            A[q * ld + q] = sn * sn * app + 2.0 * sn * c * apq + c * c * aqq;
            A[p * ld + q] = 0.0;
            A[q * ld + p] = 0.0;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(64) void lz_jacobi(LzState* s, double tol, int max_rot) {
    __shared__ double A[64 * 65];
    __shared__ double V[64 * 65];
    if (s->err) return;
    const int n = s->numiter;
    const int lane = threadIdx.x;
    if (n > 64) {  // the host limits max_iter; keep the kernel bounded anyway
        if (lane == 0) s->err = 2;
        return;
    }
    const int ld = n | 1;
    for (int i = lane; i < n * n; i += 64) {
        const int r = i / n, c = i % n;
        double v = 0.0;
        if (r == c) v = s->alpha[r];
        else if (c == r + 1) v = s->beta[r];
        else if (r == c + 1) v = s->beta[c];
        A[r * ld + c] = v;
        V[r * ld + c] = r == c ? 1.0 : 0.0;
    }
    __syncthreads();
    if (n == 1) {
        if (lane == 0) { s->scal[1] = A[0]; s->coef[0] = 1.0; }
        return;
    }
    if (n <= 32)
        jacobi_solve<32>(A, V, n, ld, tol, max_rot);
    else
        jacobi_solve<64>(A, V, n, ld, tol, max_rot);
    if (lane == 0) {
        int mi = 0;
        double mv = A[0];
        for (int i = 1; i < n; ++i)
            if (A[i * ld + i] < mv) { mv = A[i * ld + i]; mi = i; }
        double nr = 0.0;
        for (int i = 0; i < n; ++i) nr += V[i * ld + mi] * V[i * ld + mi];
        nr = sqrt(nr);
        for (int i = 0; i < n; ++i) s->coef[i] = nr > 0.0 ? V[i * ld + mi] / nr : V[i * ld + mi];
        s->scal[1] = mv;
    }
}

// Exact lowest eigenpair of the n x n Lanczos tridiagonal T (no reference
// counterpart; selected by a negative rotation cap).  It returns what the
// reference's Jacobi returns once converged, to rounding, in microseconds:
//  1. Sturm-sequence multisection: each of the 64 lanes counts the
//     eigenvalues of T below its own point of the current bracket
//     (LDL^T pivots q_i = (a_i - x) - b_{i-1}^2 / q_{i-1}, negative pivots
//     counted, tiny pivots replaced by -pivmin); the bracket shrinks 65x per
//     step around the smallest eigenvalue, starting from the Gershgorin
//     interval, until it is a few ulps wide.
//  2. Inverse iteration (one lane): three solves with T - lambda I by
//     Gaussian elimination with partial pivoting on the tridiagonal band
//     (U keeps two superdiagonals), normalising between solves; the
//     eigenvector is returned normalised with a non-negative first entry.
__device__ int sturm_count(const double* a, const double* b2, int n, double x, double pivmin) {
    int c = 0;
    double q = a[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
    for (int i = 1; i < n; ++i) {
        q = (a[i] - x) - b2[i - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        c += q < 0.0;
    }
    return c;
}

__global__ __launch_bounds__(64) void lz_tridiag_exact(LzState* s) {
    __shared__ double a[64], b[64], b2[64];
    __shared__ double u0[64], u1[64], u2[64], ml[64], x[64];
    __shared__ int sw[64];
    if (s->err) return;
    const int n = s->numiter;
    const int lane = threadIdx.x;
    if (n > 64) {
        if (lane == 0) s->err = 2;
        return;
    }
    if (lane < n) {
        a[lane] = s->alpha[lane];
        b[lane] = lane + 1 < n ? s->beta[lane] : 0.0;
        b2[lane] = b[lane] * b[lane];
    }
    __syncthreads();
    if (n == 1) {
        if (lane == 0) { s->scal[1] = a[0]; s->coef[0] = 1.0; }
        return;
    }
    // Gershgorin interval and the pivot floor (wavefront min / max)
    double lo = 1e300, hi = -1e300, tn = 0.0;
    if (lane < n) {
        const double r = fabs(b[lane]) + (lane > 0 ? fabs(b[lane - 1]) : 0.0);
        lo = a[lane] - r;
        hi = a[lane] + r;
        tn = fabs(a[lane]) + r;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
        tn = fmax(tn, __shfl_xor(tn, off, 64));
    }
    const double eps = 2.220446049250313e-16;
    const double pivmin = fmax(eps * eps * tn * tn, 1e-290);
    lo -= 2.0 * eps * tn + pivmin;
    hi += 2.0 * eps * tn + pivmin;
    // multisection: keep count(lo) == 0 <= count(hi) - 1
    for (int it = 0; it < 40; ++it) {
        const double w = hi - lo;
        if (w <= 4.0 * eps * fmax(fabs(lo), fabs(hi)) + pivmin) break;
        const double x = lo + w * (double)(lane + 1) / 65.0;
        const int c = sturm_count(a, b2, n, x, pivmin);
        const unsigned long long hit = __ballot(c >= 1);
        if (hit == 0ull) {
            lo = lo + w * 64.0 / 65.0;  // the smallest eigenvalue is above every point
        } else {
            const int f = __ffsll((long long)hit) - 1;
            const double xf = lo + w * (double)(f + 1) / 65.0;
            const double xp = f > 0 ? lo + w * (double)f / 65.0 : lo;
            lo = xp;
            hi = xf;
        }
    }
    const double lam = 0.5 * (lo + hi);
    if (lane == 0) {
        // LU of T - lam I with partial pivoting: row i of U = (u0, u1, u2) at columns i, i+1, i+2
        // (work arrays in LDS; only this lane touches them)
        const double tiny = eps * fmax(tn, 1e-300);
        double c0 = a[0] - lam, c1 = b[0], c2 = 0.0;
        for (int i = 0; i < n - 1; ++i) {
            const double n0 = b[i], n1 = a[i + 1] - lam, n2 = (i + 2 < n) ? b[i + 1] : 0.0;
            double p0, p1, p2, o0, o1, o2;
            if (fabs(n0) > fabs(c0)) { p0 = n0; p1 = n1; p2 = n2; o0 = c0; o1 = c1; o2 = c2; sw[i] = 1; }
            else { p0 = c0; p1 = c1; p2 = c2; o0 = n0; o1 = n1; o2 = n2; sw[i] = 0; }
            if (p0 == 0.0) p0 = tiny;
            const double m = o0 / p0;
            u0[i] = p0; u1[i] = p1; u2[i] = p2; ml[i] = m;
            c0 = o1 - m * p1;
            c1 = o2 - m * p2;
            c2 = 0.0;
        }
        u0[n - 1] = c0 == 0.0 ? tiny : c0;
        u1[n - 1] = 0.0;
        u2[n - 1] = 0.0;
        if (n >= 2) u2[n - 2] = 0.0;
        for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.25 * (double)((i * 37) % 11) / 11.0;  // generic start
        for (int pass = 0; pass < 3; ++pass) {
            for (int i = 0; i < n - 1; ++i) {
                if (sw[i]) { const double t = x[i]; x[i] = x[i + 1]; x[i + 1] = t; }
                x[i + 1] -= ml[i] * x[i];
            }
            x[n - 1] /= u0[n - 1];
            for (int i = n - 2; i >= 0; --i) {
                double v = x[i] - u1[i] * x[i + 1];
                if (i + 2 < n) v -= u2[i] * x[i + 2];
                x[i] = v / u0[i];
            }
            double nr = 0.0;
            for (int i = 0; i < n; ++i) nr = fmax(nr, fabs(x[i]));
            for (int i = 0; i < n; ++i) x[i] /= nr;
        }
        double nr = 0.0;
        for (int i = 0; i < n; ++i) nr += x[i] * x[i];
        nr = sqrt(nr);
        const double sg = x[0] < 0.0 ? -1.0 : 1.0;
        for (int i = 0; i < n; ++i) s->coef[i] = sg * x[i] / nr;
        // Rayleigh quotient of the vector: the eigenvalue to rounding
        double rq = 0.0;
        for (int i = 0; i < n; ++i) {
            double t = a[i] * s->coef[i];
            if (i > 0) t += b[i - 1] * s->coef[i - 1];
            if (i + 1 < n) t += b[i] * s->coef[i + 1];
            rq += s->coef[i] * t;
        }
        s->scal[1] = rq;
    }
}

// theta = sum_k coef[k] K[k] (k < numiter); partial sums of theta.theta (dmrg.mojo:728-742)
__global__ __launch_bounds__(256) void lz_ritz(const double* __restrict__ K, int64_t n, double* __restrict__ theta,
                                               const LzState* s, double* __restrict__ part) {
    if (s->err) return;
    const int nk = s->numiter;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double x = 0.0;
        for (int k = 0; k < nk; ++k) x += s->coef[k] * K[(int64_t)k * n + i];
        theta[i] = x;
        acc += x * x;
    }
    acc = block_sum256(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// x *= 1/sqrt(scal[idx]) when positive (dmrg.mojo:743-745)
__global__ __launch_bounds__(256) void lz_normalize(double* __restrict__ x, int64_t n, const LzState* s, int idx) {
    if (s->err) return;
    const double nr = sqrt(s->scal[idx]);
    if (!(nr > 0.0)) return;
    const double inv = 1.0 / nr;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= inv;
}

// r = w - E_rq psi (in place in w); partial sums of r.r
__global__ __launch_bounds__(256) void lz_residual(double* __restrict__ w, const double* __restrict__ psi, int64_t n,
                                                   const LzState* s, double* __restrict__ part) {
    if (s->err) return;
    const double e = s->scal[3];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double r = w[i] - e * psi[i];
        w[i] = r;
        acc += r * r;
    }
    acc = block_sum256(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// rhat = r / |r| into out (zero when |r| == 0)
__global__ __launch_bounds__(256) void lz_rhat(const double* __restrict__ r, double* __restrict__ out, int64_t n,
                                               const LzState* s) {
    if (s->err) return;
    const double nr = sqrt(s->scal[4]);
    const double inv = nr > 0.0 ? 1.0 / nr : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = r[i] * inv;
}

// 2x2 Rayleigh-Ritz in span{psi, rhat}: [[E, rho], [rho, h_rr]]; lowest
// eigenpair (c0, c1) -> psi = c0 psi + c1 rhat (normalised by construction).
__global__ __launch_bounds__(256) void lz_correct(double* __restrict__ psi, const double* __restrict__ rhat, int64_t n,
                                                  LzState* s) {
    if (s->err) return;
    const double E = s->scal[3], rho = sqrt(s->scal[4]), hrr = s->scal[5];
    double c0 = 1.0, c1 = 0.0, lam = E;
    if (rho > 0.0) {
        const double mid = 0.5 * (E + hrr), half = 0.5 * (hrr - E);
        const double rad = sqrt(half * half + rho * rho);
        lam = mid - rad;
        // eigenvector of [[E - lam, rho], [rho, hrr - lam]]: (rho, lam - E) up to scale
        double x0 = rho, x1 = lam - E;
        const double nr = sqrt(x0 * x0 + x1 * x1);
        c0 = x0 / nr;
        c1 = x1 / nr;
        if (c0 < 0.0) { c0 = -c0; c1 = -c1; }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) s->scal[6] = lam;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        psi[i] = c0 * psi[i] + c1 * rhat[i];
}

// dst = src (unless the start vector was zero)
__global__ __launch_bounds__(256) void lz_copy(const double* __restrict__ src, double* __restrict__ dst, int64_t n,
                                               const LzState* s) {
    if (s->err) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// One Krylov vector (breakdown at the first beta): the Ritz vector is v0, and
// H psi is the saved FP64 product H v0; otherwise the refinement's FP64 matvec
// of psi runs (refmv).
__global__ __launch_bounds__(256) void lz_take_start(const double* __restrict__ v0, const double* __restrict__ hv0,
                                                     double* __restrict__ psi, double* __restrict__ w, int64_t n,
                                                     LzState* s) {
    if (s->err) return;
    const bool one = s->numiter == 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) s->refmv = one ? 0 : 1;
    if (!one) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        psi[i] = v0[i];
        w[i] = hv0[i];
    }
}

// Refined energy above the start vector's: keep v0 and alpha_0 (= <v0|H v0>, FP64).
__global__ void lz_guard(LzState* s, int refine) {
    if (s->err) return;
    const double e = refine >= 2 ? s->scal[6] : s->scal[3];
    if (e > s->alpha[0]) {
        s->kept_start = 1;
        s->scal[3] = s->alpha[0];
        s->scal[6] = s->alpha[0];
    }
}

// psi = v0 when kept_start; partial sums of |H v0 - alpha_0 v0|^2 (its residual)
__global__ __launch_bounds__(256) void lz_restore_start(const double* __restrict__ v0, const double* __restrict__ hv0,
                                                        double* __restrict__ psi, int64_t n, const LzState* s,
                                                        double* __restrict__ part) {
    if (s->err || !s->kept_start) return;
    const double a = s->alpha[0];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        psi[i] = v0[i];
        const double r = hv0[i] - a * v0[i];
        acc += r * r;
    }
    acc = block_sum256(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void lz_reduce_kept(const double* __restrict__ part, LzState* s) {
    if (s->err || !s->kept_start) return;
    double acc = 0.0;
    for (int i = threadIdx.x; i < NB; i += blockDim.x) acc += part[i];
    acc = block_sum256(acc);
    if (threadIdx.x == 0) s->scal[4] = acc;
}

inline int grid_n(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

}  // namespace

int lanczos_enqueue(Handle* h, const HeffEnv& env, const HeffEnv* env64, double* d_theta, int max_iter, double tol,
                    int max_rot, int refine) {
    const HeffDims& D = env.dims;
    const int64_t dim = D.Dl * D.d * D.d * D.Dr;
    CT_ARG(max_iter >= 1 && max_iter <= 64, 5, "max_iter must be in [1, 64]");
    void *kb, *wb, *pb, *sb;
    CT_CHECK(ws_get(h, WS_KRYLOV, (size_t)dim * max_iter * 8, &kb));
    CT_CHECK(ws_get(h, WS_W, (size_t)dim * 3 * 8, &wb));
    CT_CHECK(ws_get(h, WS_PARTIAL, 2 * NB * 8, &pb));
    CT_CHECK(ws_get(h, WS_LZ, sizeof(LzState), &sb));
    double* K = (double*)kb;
    double* w = (double*)wb;
    double* w2 = w + dim;
    double* hv0 = w + 2 * dim;  // FP64 H v0, kept for the refinement (exact_start)
    double* part = (double*)pb;
    double* part2 = part + NB;  // w.w partials of a Krylov step (part holds its v_j.w partials)
    LzState* s = (LzState*)sb;
    const int* act = &s->active;
    const double thr = 100.0 * (double)dim * 2.2204460492503131e-16;
    // Noise floor of the mixed-precision recurrences (beta_j < floor_rel * max|alpha|
    // ends the Krylov space), about 2-10x the relative matvec error of the form:
    // FP32 and 3xBF16 1e-5 (matvec errors 1e-6 - 8e-6), 6xBF16 1e-6 (3e-7 - 1.2e-6).
    const double floor_rel = env.prec == CT_PREC_FP64 ? 0.0 : env.prec == CT_PREC_BF16X6 ? 1e-6 : 1e-5;
    const bool exact_start = env.prec != CT_PREC_FP64 && refine >= 1 && env64 != nullptr;
    hipStream_t st = h->stream;
    const int G = grid_n(dim);
    // Krylov matvec of v_j: the first one in FP64 when exact_start (and saved)
    auto krylov_matvec = [&](int j, const double* vj) -> int {
        if (j == 0 && exact_start) {
            CT_CHECK(heff_apply(h, *env64, CT_F64, vj, w, act));
            hipLaunchKernelGGL(lz_copy, dim3(G), dim3(256), 0, st, (const double*)w, hv0, dim, (const LzState*)s);
            return 0;
        }
        return heff_apply(h, env, CT_F64, vj, w, act);
    };

    hipLaunchKernelGGL(lz_init_kernel, dim3(1), dim3(1), 0, st, s, max_iter, exact_start ? 1 : 0);
    hipLaunchKernelGGL(lz_dot_partial, dim3(NB), dim3(256), 0, st, (const double*)d_theta, (const double*)nullptr, dim,
                       part, (const LzState*)s, 0);
    hipLaunchKernelGGL(lz_reduce, dim3(1), dim3(256), 0, st, (const double*)part, s, 0, 0, 0);
    hipLaunchKernelGGL(lz_start, dim3(G), dim3(256), 0, st, (const double*)d_theta, K, dim, s);
    for (int j = 0; j < max_iter - 1; ++j) {
        double* vj = K + (int64_t)j * dim;
        CT_CHECK(krylov_matvec(j, vj));
        hipLaunchKernelGGL(lz_dot_partial, dim3(NB), dim3(256), 0, st, (const double*)vj, (const double*)w, dim, part,
                           (const LzState*)s, 1);
        hipLaunchKernelGGL(lz_update_alpha, dim3(NB), dim3(256), 0, st, w, (const double*)vj,
                           (const double*)(j > 0 ? vj - dim : vj), dim, j, s, (const double*)part, part2);
        hipLaunchKernelGGL(lz_next_beta, dim3(G), dim3(256), 0, st, (const double*)w, vj + dim, dim, j, s,
                           (const double*)part2, thr, floor_rel);
    }
    {   // final alpha-only step when no breakdown happened (dmrg.mojo:691-703)
        double* vj = K + (int64_t)(max_iter - 1) * dim;
        CT_CHECK(krylov_matvec(max_iter - 1, vj));
        hipLaunchKernelGGL(lz_dot_partial, dim3(NB), dim3(256), 0, st, (const double*)vj, (const double*)w, dim, part,
                           (const LzState*)s, 1);
        hipLaunchKernelGGL(lz_reduce, dim3(1), dim3(256), 0, st, (const double*)part, s, 1, max_iter - 1, 1);
    }
    if (max_rot < 0)
        hipLaunchKernelGGL(lz_tridiag_exact, dim3(1), dim3(64), 0, st, s);
    else
        hipLaunchKernelGGL(lz_jacobi, dim3(1), dim3(64), 0, st, s, tol, max_rot);
    hipLaunchKernelGGL(lz_ritz, dim3(NB), dim3(256), 0, st, (const double*)K, dim, d_theta, (const LzState*)s, part);
    hipLaunchKernelGGL(lz_reduce, dim3(1), dim3(256), 0, st, (const double*)part, s, 0, 2, 0);
    hipLaunchKernelGGL(lz_normalize, dim3(G), dim3(256), 0, st, d_theta, dim, (const LzState*)s, 2);
    CT_HIP(hipGetLastError());

    if (refine >= 1 && env64) {
        // E_rq = <psi|H psi> with FP64 environments (psi = v0 and the saved
        // H v0 when the Krylov space has one vector)
        if (exact_start)
            hipLaunchKernelGGL(lz_take_start, dim3(G), dim3(256), 0, st, (const double*)K, (const double*)hv0, d_theta,
                               w, dim, s);
        CT_CHECK(heff_apply(h, *env64, CT_F64, d_theta, w, exact_start ? &s->refmv : nullptr));
        hipLaunchKernelGGL(lz_dot_partial, dim3(NB), dim3(256), 0, st, (const double*)d_theta, (const double*)w, dim,
                           part, (const LzState*)s, 0);
        hipLaunchKernelGGL(lz_reduce, dim3(1), dim3(256), 0, st, (const double*)part, s, 0, 3, 0);
        hipLaunchKernelGGL(lz_residual, dim3(NB), dim3(256), 0, st, w, (const double*)d_theta, dim, (const LzState*)s,
                           part);
        hipLaunchKernelGGL(lz_reduce, dim3(1), dim3(256), 0, st, (const double*)part, s, 0, 4, 0);
        if (refine >= 2) {
            hipLaunchKernelGGL(lz_rhat, dim3(G), dim3(256), 0, st, (const double*)w, w2, dim, (const LzState*)s);
            // h_rr = <rhat|H rhat> at the Krylov precision (second-order effect on the energy)
            CT_CHECK(heff_apply(h, env, CT_F64, w2, w, nullptr));
            hipLaunchKernelGGL(lz_dot_partial, dim3(NB), dim3(256), 0, st, (const double*)w2, (const double*)w, dim,
                               part, (const LzState*)s, 0);
            hipLaunchKernelGGL(lz_reduce, dim3(1), dim3(256), 0, st, (const double*)part, s, 0, 5, 0);
            hipLaunchKernelGGL(lz_correct, dim3(G), dim3(256), 0, st, d_theta, (const double*)w2, dim, s);
        }
        if (exact_start) {
            hipLaunchKernelGGL(lz_guard, dim3(1), dim3(1), 0, st, s, refine);
            hipLaunchKernelGGL(lz_restore_start, dim3(NB), dim3(256), 0, st, (const double*)K, (const double*)hv0,
                               d_theta, dim, (const LzState*)s, part);
            hipLaunchKernelGGL(lz_reduce_kept, dim3(1), dim3(256), 0, st, (const double*)part, s);
        }
        CT_HIP(hipGetLastError());
    }
    return 0;
}

int lanczos_finish(Handle* h, int refine, bool have_env64, LanczosResult* res) {
    void* sb;
    CT_CHECK(ws_get(h, WS_LZ, sizeof(LzState), &sb));
    const LzState* s = (const LzState*)sb;
    LzState* hs = (LzState*)h->h_pinned;
    CT_HIP(hipMemcpyAsync(hs, s, sizeof(LzState), hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipStreamSynchronize(h->stream));
    if (hs->err == 1) {
        set_last_error("Initial vector has zero norm");
        return CT_STATUS_ZERO_NORM;
    }
    if (hs->err != 0) {
        set_last_error("Lanczos produced an invalid Krylov basis");
        return CT_STATUS_EMPTY_KRYLOV;
    }
    res->num_iter = hs->numiter;
    res->E_lanczos = hs->scal[1];
    res->E_rq = hs->scal[3];
    res->res_norm = sqrt(hs->scal[4]);
    res->E0 = res->E_lanczos;
    res->kept_start = hs->kept_start;
    res->matvecs = hs->numiter;
    if (refine >= 1 && have_env64) {
        res->E0 = refine >= 2 ? hs->scal[6] : hs->scal[3];
        res->matvecs += (hs->refmv ? 1 : 0) + (refine >= 2 ? 1 : 0);
    }
    return 0;
}

int lanczos(Handle* h, const HeffEnv& env, const HeffEnv* env64, double* d_theta, int max_iter, double tol,
            int max_rot, int refine, LanczosResult* res) {
    CT_CHECK(lanczos_enqueue(h, env, env64, d_theta, max_iter, tol, max_rot, refine));
    return lanczos_finish(h, refine, env64 != nullptr, res);
}

}  // namespace ct
