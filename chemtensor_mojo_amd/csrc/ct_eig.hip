// Symmetric eigensolver behind the Gram-matrix two-site split.
//
// rocSOLVER's dsyevd spends about 70% of its time in the Householder
// tridiagonalization (dsytrd), which issues several small kernels per column
// (about 9700 launches at n = 2048) and is bound by their latency.  Here the
// tridiagonalization is one cooperative kernel: the matrix stays in the
// registers of P = ceil(n/R) workgroups (R = 8 or 16 rows each, one
// workgroup per CU)
// and each column costs one grid-wide exchange.  The tridiagonal
// eigenproblem (dstedc) and the back-transformation (dormtr) stay with
// rocSOLVER: the kernel writes exactly dsytrd's lower-triangle output.
//
// Algorithm: the unblocked Householder reduction (LAPACK dsytd2, lower).
// For k = 0 .. n-2, with x = A[k, k+1:] (= A[k+1:, k]):
//   (beta, tau, v) = dlarfg(x)          v[k+1] = 1, v[c] = 0 for c <= k
//   p = tau A22 v,  alpha = -tau/2 (p.v),  w = p + alpha v
//   A22 -= v w^T + w v^T                 A22 = A[k+1:, k+1:]
// Distribution: row r lives in workgroup r mod P; thread t of a workgroup of
// NT threads holds columns t, t+NT, ... of its R rows.  Step k:
//   1. each workgroup forms p for its rows and publishes it with its share of
//      p.v; the owner of row k+1 also publishes that row as it stands before
//      step k;
//   2. grid-wide arrival: one flag per workgroup, polled by every workgroup;
//   3. every workgroup reads p and row k+1, applies step k to row k+1 and
//      forms reflector k+1 itself.  All workgroups run the same arithmetic on
//      the same data, so every copy of v, tau and w is bitwise identical and
//      no second exchange is needed.  Then each workgroup updates its rows.
// Exchange memory is write-once: step k has its own region (one 128-byte line
// per workgroup with its R p values and its share of p.v, then row k+1),
// never reused within a launch, and every line is written by one workgroup.
// Writers store with agent-scope (write-through) stores, each wave waits for
// its stores to complete, and only then is the arrival flag set (agent-scope
// store; readers poll the flags with agent-scope loads).  A reader loads a
// region only after it has seen every flag of that step.  No cache can hold
// an earlier copy of a region's line (the line was never read before in this
// launch, and caches are invalidated at kernel start), so plain loads return
// the published data and the 32 workgroups of an XCD share one L2 fill.  This
// avoids both the L2 write-back of an agent-scope release fence and the L2
// invalidation of an acquire fence, which measured 2.0 us per step together
// at n = 2048.
#include "ct_internal.h"
#include "ct_device.h"

namespace ct {
namespace {

constexpr int TR_MAXN = 2048;

struct TrArgs {
    double* A;        // [n][lda]: symmetric input; output rows k hold d_k, e_k, v_k (see header comment)
    int64_t lda;
    int n;
    double *d, *e, *tau;
    double* xb;       // n-1 exchange regions of xs doubles: P lines of LW (p[R], share) | row[np]
    int64_t xs;
    unsigned* flags;  // arrival counts, flag g at flags[g * fstride], zero at launch
    int fstride;
    int* err;         // set to 1 when an arrival wait exceeds spin_limit polls
    int spin_limit;
};

// a - (vr wc + wr vc): one rounding order everywhere, so that the owner of a
// row and every other workgroup get bitwise the same values.
__device__ __forceinline__ double upd2(double a, double vr, double wr, double vc, double wc) {
    return fma(-wr, vc, fma(-vr, wc, a));
}

// Sum of one value per thread over the NT threads of the workgroup, in a
// fixed order (every workgroup summing the same inputs gets the same bits).
template <int NT>
__device__ __forceinline__ double block_sum_fixed(double v, double* red) {
    constexpr int W = NT / 64;
    v = wave_sum_dpp(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int w = 1; w < W; ++w) s += red[w];
    return s;
}

// Sums of NV values per thread over the NT threads of the workgroup: a DPP
// butterfly inside each row of 16 lanes, then one LDS pass in which thread v
// (v < NV, all in wave 0) adds the NT/16 row sums of value v.  The sum of
// value v is returned in lane v of wave 0 (other lanes: unspecified).
template <int NT, int NV>
__device__ __forceinline__ double block_sums_to_lanes(double (&x)[NV], double* part) {
    constexpr int G = NT / 16;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        x[v] += dpp_mov64<0xB1>(x[v]);   // quad_perm [1,0,3,2]
        x[v] += dpp_mov64<0x4E>(x[v]);   // quad_perm [2,3,0,1]
        x[v] += dpp_mov64<0x141>(x[v]);  // row_half_mirror
        x[v] += dpp_mov64<0x140>(x[v]);  // row_mirror
    }
    if ((threadIdx.x & 15) == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) part[v * G + (threadIdx.x >> 4)] = x[v];
    }
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x < NV) {
#pragma unroll
        for (int k = 0; k < G; ++k) s += part[threadIdx.x * G + k];
    }
    return s;
}

// LAPACK dlarfg on the row slice x[c], c >= k+1 (alpha = x[k+1]): v[k+1] = 1,
// v[c] = x[c] / (alpha - beta) for c >= k+2, zero elsewhere.  Every thread of
// every workgroup gets the same tau, beta, v.
template <int NT, int NC>
__device__ __forceinline__ void make_reflector(int k, int n, const int (&c)[NC], const double (&x)[NC],
                                               double (&v)[NC], double& tau, double& beta, double* red,
                                               double* bcast) {
    double s2 = 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        if (c[j] >= k + 2 && c[j] < n) s2 = fma(x[j], x[j], s2);
        if (c[j] == k + 1) *bcast = x[j];
    }
    s2 = block_sum_fixed<NT>(s2, red);
    const double alpha = *bcast;
    if (s2 == 0.0) {
        tau = 0.0;
        beta = alpha;
#pragma unroll
        for (int j = 0; j < NC; ++j) v[j] = c[j] == k + 1 ? 1.0 : 0.0;
    } else {
        const double nrm = sqrt(fma(alpha, alpha, s2));
        beta = alpha >= 0.0 ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
#pragma unroll
        for (int j = 0; j < NC; ++j)
            v[j] = c[j] == k + 1 ? 1.0 : ((c[j] >= k + 2 && c[j] < n) ? x[j] * scal : 0.0);
    }
}

// NT threads per workgroup, R matrix rows per workgroup, NC columns per thread
// (n <= NT*NC, P*R >= n).
template <int NT, int R, int NC>
__global__ __launch_bounds__(NT, 1) void sytrd_coop_kernel(TrArgs q) {
    constexpr int W = NT / 64;
    constexpr int LW = R + 1 <= 16 ? 16 : 32;  // doubles per exchange line (p[R], share)
    const int n = q.n, P = gridDim.x, g = blockIdx.x, t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    __shared__ double part[(R + 2) * (NT / 16)];
    __shared__ double redw[W];
    __shared__ double bcast;
    __shared__ double ploc[R], vr[2][R], vrp[R], wrp[R];
    __shared__ double redp[W];
    bool dead = false;  // this thread's arrival wait timed out once: stop waiting

    // Column c[j] of this thread; where its p value sits in a step's
    // exchange region; which of this workgroup's rows it is (-1: none).
    int c[NC], poff[NC], lin[NC], lrow[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        c[j] = t + NT * j;
        lin[j] = c[j] % P;
        poff[j] = LW * lin[j] + c[j] / P;
        const int dl = c[j] - g;
        lrow[j] = (c[j] < n && dl >= 0 && dl % P == 0 && dl / P < R) ? dl / P : -1;
    }
    double a[R][NC];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int r = g + P * s;
#pragma unroll
        for (int j = 0; j < NC; ++j) a[s][j] = (r < n && c[j] < n) ? q.A[(int64_t)r * q.lda + c[j]] : 0.0;
    }
    if (t < R) { vr[0][t] = 0.0; vr[1][t] = 0.0; vrp[t] = 0.0; wrp[t] = 0.0; }

    // reflector 0 from row 0 of the input
    double v[NC], x[NC], w[NC];
    // The rank-2 update of step i is applied to the registers during step i+1,
    // after the arrival flag is set, so it overlaps the wait instead of
    // delaying the next p: "pending" holds (v, w) of that update at this
    // thread's columns (vp, wp) and at this workgroup's rows (vrp, wrp; zero
    // for rows that are finished, so those stay untouched).
    double vp[NC], wp[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) { vp[j] = 0.0; wp[j] = 0.0; }
#pragma unroll
    for (int j = 0; j < NC; ++j) x[j] = c[j] < n ? q.A[c[j]] : 0.0;
    double tau = 0.0, beta = 0.0;
    if (n >= 2) {
        make_reflector<NT, NC>(0, n, c, x, v, tau, beta, redw, &bcast);
    } else {
#pragma unroll
        for (int j = 0; j < NC; ++j) v[j] = 0.0;
    }
    if (g == 0) {
        if (t == 0) {
            q.d[0] = x[0];
            if (n >= 2) { q.e[0] = beta; q.tau[0] = tau; }
        }
        // row 0 (slot 0 of workgroup 0) in output form
#pragma unroll
        for (int j = 0; j < NC; ++j)
            if (c[j] >= 1 && c[j] < n) a[0][j] = c[j] == 1 ? beta : v[j];
    }
    int cur = 0;
#pragma unroll
    for (int j = 0; j < NC; ++j)
        if (lrow[j] >= 0) vr[0][lrow[j]] = v[j];
    __syncthreads();

    int own1 = 1 % P, s1 = 1 / P;  // owner and slot of row i+1
    for (int i = 0; i + 1 < n; ++i) {
        double* xl = q.xb + (int64_t)i * q.xs;  // line g: p at rows g + P*s (s < R), share of p.v
        double* xr = xl + LW * P;               // row i+1 before step i

        // ---- 1. p_i = tau A^(i) v at this workgroup's rows, published.  The
        // registers hold A^(i-1) (step i-1 pending), so
        //   A^(i) v = A^(i-1) v - vp (wp.v) - wp (vp.v).
        double ps[R + 2];
#pragma unroll
        for (int s = 0; s < R; ++s) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < NC; ++j) acc = fma(a[s][j], v[j], acc);
            ps[s] = acc;
        }
        ps[R] = 0.0;
        ps[R + 1] = 0.0;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            ps[R] = fma(wp[j], v[j], ps[R]);
            ps[R + 1] = fma(vp[j], v[j], ps[R + 1]);
        }
        if (g == own1) {  // row i+1 of A^(i)
#pragma unroll
            for (int s = 0; s < R; ++s)
                if (s == s1) {
                    const double vs = vrp[s], ws = wrp[s];
#pragma unroll
                    for (int j = 0; j < NC; ++j)
                        if (c[j] > i && c[j] < n)
                            __hip_atomic_store(&xr[c[j]], upd2(a[s][j], vs, ws, vp[j], wp[j]), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                }
        }
        const double sums = block_sums_to_lanes<NT, R + 2>(ps, part);  // lane s of wave 0: row s, R: wp.v, R+1: vp.v
        if (wave == 0) {
            const double wv = readlane64(sums, R), vv = readlane64(sums, R + 1);
            double share = 0.0;
            if (lane < R) {
                const int r = g + P * lane;
                double pv = fma(-wrp[lane], vv, fma(-vrp[lane], wv, sums));
                pv = (r > i && r < n) ? tau * pv : 0.0;
                __hip_atomic_store(&xl[LW * g + lane], pv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ploc[lane] = pv;
                share = pv * vr[cur][lane];
            }
            share = wave_sum_dpp(share);
            if (lane == 0) __hip_atomic_store(&xl[LW * g + R], share, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // Every wave waits until its own exchange stores have completed (the
        // barrier alone orders only LDS traffic), then the flag is set.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (t == 0)
            __hip_atomic_store(&q.flags[g * q.fstride], (unsigned)(i + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // the pending update of step i-1, while the flags travel
#pragma unroll
        for (int s = 0; s < R; ++s) {
            const double vs = vrp[s], ws = wrp[s];
#pragma unroll
            for (int j = 0; j < NC; ++j) a[s][j] = upd2(a[s][j], vs, ws, vp[j], wp[j]);
        }

        // ---- 2. every thread waits only for the writers of the data it reads
        // and loads each piece as soon as its flag is seen: p of its columns
        // (line c mod P), the share of line t (t < P), row i+1 (its owner).
        // The lines are write-once, so a load issued after the flag returns
        // the published value; the workgroup-scope acquire after each wait
        // keeps the compiler from hoisting the load above the poll.
        double pc[NC], rc[NC];
        double shr = 0.0;
        const unsigned target = (unsigned)(i + 1);
        auto wait_flag = [&](int m) {
            int polls = 0;
            while (!dead && __hip_atomic_load(&q.flags[m * q.fstride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                                target) {
                if (++polls > q.spin_limit) {
                    dead = true;
                    atomicExch(q.err, 1);
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        };
        wait_flag(own1);
        const double p1 = xl[LW * own1 + s1];
#pragma unroll
        for (int j = 0; j < NC; ++j) rc[j] = (c[j] > i && c[j] < n) ? xr[c[j]] : 0.0;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            pc[j] = 0.0;
            if (c[j] > i && c[j] < n) {
                wait_flag(lin[j]);
                pc[j] = xl[poff[j]];
            }
        }
        if (t < P) {
            wait_flag(t);
            shr = xl[LW * t + R];
        }

        // ---- 3. w, row i+1 after step i, reflector i+1, local update
        const double pdotv = block_sum_fixed<NT>(shr, redp);  // same order in every workgroup
        const double alpha = -0.5 * tau * pdotv;
#pragma unroll
        for (int j = 0; j < NC; ++j) w[j] = fma(alpha, v[j], pc[j]);
        const double w1 = fma(alpha, 1.0, p1);
#pragma unroll
        for (int j = 0; j < NC; ++j) x[j] = (c[j] > i && c[j] < n) ? upd2(rc[j], 1.0, w1, v[j], w[j]) : 0.0;
        if (g == 0) {
#pragma unroll
            for (int j = 0; j < NC; ++j)
                if (c[j] == i + 1) q.d[i + 1] = x[j];
        }
        double vn[NC];
        double taun = 0.0, betan = 0.0;
        if (i + 2 < n) {
            make_reflector<NT, NC>(i + 1, n, c, x, vn, taun, betan, redw, &bcast);
            if (g == 0 && t == 0) { q.e[i + 1] = betan; q.tau[i + 1] = taun; }
        } else {
#pragma unroll
            for (int j = 0; j < NC; ++j) vn[j] = 0.0;
            __syncthreads();
        }
        // step i's update becomes pending; row i+1 is finished (output form)
        if (t < R) {
            const int r = g + P * t;
            vrp[t] = r > i + 1 ? vr[cur][t] : 0.0;
            wrp[t] = r > i + 1 ? fma(alpha, vr[cur][t], ploc[t]) : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            vp[j] = v[j];
            wp[j] = w[j];
        }
        if (g == own1) {
#pragma unroll
            for (int s = 0; s < R; ++s)
                if (s == s1) {
#pragma unroll
                    for (int j = 0; j < NC; ++j)
                        if (c[j] > i && c[j] < n) a[s][j] = c[j] == i + 1 ? x[j] : (c[j] == i + 2 ? betan : vn[j]);
                }
        }
        const int nxt = cur ^ 1;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            if (lrow[j] >= 0) vr[nxt][lrow[j]] = vn[j];
            v[j] = vn[j];
        }
        tau = taun;
        cur = nxt;
        if (++own1 == P) { own1 = 0; ++s1; }
        __syncthreads();
    }

    // rows back in dsytrd's lower format (column-major view): A[r][c], c >= r
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int r = g + P * s;
        if (r < n) {
#pragma unroll
            for (int j = 0; j < NC; ++j)
                if (c[j] >= r && c[j] < n) q.A[(int64_t)r * q.lda + c[j]] = a[s][j];
        }
    }
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ------------------------------------------------- back-transformation
// Eigenvectors of A from those of T: u = Q z with Q = H_0 H_1 ... H_{n-2}.
// With eigenvectors as rows (E = [z_j^T]), E' = E Q^T = E H_{n-2} ... H_0.
// Blocks of BT_B reflectors: H_k0 ... H_{k0+B-1} = I - V T V^T (LAPACK dlarft,
// forward, columnwise; V = Vt_m^T), so a block applies as
//   E <- E - (E Vt_m^T) (T^T Vt_m),
// last block first.  All Vt_m, T_m and M_m = T_m^T Vt_m are built up front
// (batched GEMMs and one small kernel); the sequential part is two GEMMs per
// block on the rows actually needed (the chi kept eigenvectors).
constexpr int BT_B = 64;

// Vt[k][c] = v_k[c]: 0 for c <= k, 1 at c = k+1, A[k][c] beyond; rows k >= n-1
// (padding up to a whole block) are zero with tau 0.
__global__ void build_vt_kernel(const double* __restrict__ A, const double* __restrict__ tau, double* __restrict__ Vt,
                                double* __restrict__ taup, int n, int nr) {
    const int64_t total = (int64_t)nr * n;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(t / n), c = (int)(t - (int64_t)k * n);
        double v = 0.0;
        if (k < n - 1) v = c <= k ? 0.0 : (c == k + 1 ? 1.0 : A[(int64_t)k * n + c]);
        Vt[t] = v;
        if (c == 0) taup[k] = k < n - 1 ? tau[k] : 0.0;
    }
}

// T (upper triangular) of one block from S = V^T V: T(i,i) = tau_i,
// T(0:i, i) = -tau_i T(0:i, 0:i) S(0:i, i).  One workgroup per block; row r of
// column i is summed by the 4 lanes of a quad (q = r + l, r + l + 4, ...).
__global__ __launch_bounds__(256) void larft_kernel(const double* __restrict__ S, const double* __restrict__ taup,
                                                    double* __restrict__ T) {
    constexpr int B = BT_B;
    __shared__ double s[B][B + 1], tt[B][B + 1];
    const int m = blockIdx.x;
    const int r = threadIdx.x >> 2, l = threadIdx.x & 3;
    const double* Sm = S + (int64_t)m * B * B;
    for (int t = threadIdx.x; t < B * B; t += blockDim.x) {
        s[t / B][t % B] = Sm[t];
        tt[t / B][t % B] = 0.0;
    }
    __syncthreads();
    for (int i = 0; i < B; ++i) {
        const double ti = taup[m * B + i];
        double acc = 0.0;  // reads columns < i of tt, writes column i
        for (int q = r + l; q < i; q += 4) acc = fma(tt[r][q], s[q][i], acc);
        acc += dpp_mov64<0xB1>(acc);
        acc += dpp_mov64<0x4E>(acc);
        if (l == 0) {
            if (r < i) tt[r][i] = -ti * acc;
            if (r == i) tt[i][i] = ti;
        }
        __syncthreads();
    }
    double* Tm = T + (int64_t)m * B * B;
    for (int t = threadIdx.x; t < B * B; t += blockDim.x) Tm[t] = tt[t / B][t % B];
}

template <int NT, int R, int NC>
int launch_sytrd(Handle* h, TrArgs a, bool* ok) {
    constexpr int LW = R + 1 <= 16 ? 16 : 32;
    const int n = a.n, P = (n + R - 1) / R;
    static int per_cu = -1;  // co-resident workgroups per CU (same for every MI355X)
    if (per_cu < 0) CT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sytrd_coop_kernel<NT, R, NC>, NT, 0));
    if ((int64_t)per_cu * h->num_cu < P) {
        *ok = false;
        return 0;
    }
    const int64_t np = (n + 63) / 64 * 64;
    a.xs = (int64_t)LW * P + np;  // multiple of 16 doubles: every line 128-byte aligned
    void* ws;
    CT_CHECK(ws_get(h, WS_EIG, (size_t)(n * a.xs) * 8 + (size_t)P * a.fstride * 4, &ws));
    a.xb = (double*)ws;
    a.flags = (unsigned*)((double*)ws + n * a.xs);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(P), dim3(64), 0, h->stream, a.flags, P * a.fstride);
    CT_HIP(hipGetLastError());
    void* kargs[] = {&a};
    CT_HIP(hipLaunchCooperativeKernel((const void*)sytrd_coop_kernel<NT, R, NC>, dim3(P), dim3(NT), kargs, 0,
                                      h->stream));
    *ok = true;
    return 0;
}

}  // namespace

bool sytrd_coop_supported(int64_t n) { return n >= 2 && n <= TR_MAXN; }

int sytrd_coop(Handle* h, double* A, int64_t n, double* d, double* e, double* tau, int* d_err, bool* ok) {
    *ok = false;
    if (!sytrd_coop_supported(n)) return 0;
    TrArgs a;
    a.A = A; a.lda = n; a.n = (int)n;
    a.d = d; a.e = e; a.tau = tau;
    a.err = d_err;
    a.spin_limit = 1 << 22;
    a.fstride = 16;  // one flag per 64 bytes: measured 2x shorter arrival waits than packed flags
    // 8 rows per workgroup (64 doubles of the matrix per thread at n = 2048);
    // measured faster than 16-row workgroups of 512 threads, whose per-step
    // reductions cost more than the smaller grid saves
    if (n <= 256) return launch_sytrd<256, 8, 1>(h, a, ok);
    if (n <= 512) return launch_sytrd<256, 8, 2>(h, a, ok);
    if (n <= 1024) return launch_sytrd<256, 8, 4>(h, a, ok);
    return launch_sytrd<256, 8, 8>(h, a, ok);
}

namespace {

int bt_blocks(int64_t n) { return n >= 2 ? (int)((n - 1 + BT_B - 1) / BT_B) : 0; }

// Builds Vt, T and M = T^T Vt for every block (see the back-transformation notes above).
int bt_prepare(Handle* h, const double* A, const double* tau, int64_t n) {
    const int nb = bt_blocks(n);
    if (nb == 0) return 0;
    const int64_t nr = (int64_t)nb * BT_B;
    void *wv, *wm, *wst;
    CT_CHECK(ws_get(h, WS_EIG3, (size_t)(nr * n) * 8, &wv));
    CT_CHECK(ws_get(h, WS_EIG4, (size_t)(nr * n) * 8, &wm));
    CT_CHECK(ws_get(h, WS_EIG5, (size_t)(2 * nb * BT_B * BT_B + nr) * 8, &wst));
    double* Vt = (double*)wv;
    double* M = (double*)wm;
    double* S = (double*)wst;
    double* T = S + (int64_t)nb * BT_B * BT_B;
    double* taup = T + (int64_t)nb * BT_B * BT_B;
    hipLaunchKernelGGL(build_vt_kernel, dim3(1024), dim3(256), 0, h->stream, A, tau, Vt, taup, (int)n, (int)nr);
    CT_HIP(hipGetLastError());
    {   // S_m = Vt_m Vt_m^T
        GemmDesc g;
        g.M = BT_B; g.N = BT_B; g.K = (int)n; g.batch = nb;
        g.a_kc = 1; g.sAm = n; g.sAk = 1; g.bsA = BT_B * n;
        g.b_kc = 1; g.sBn = n; g.sBk = 1; g.bsB = BT_B * n;
        g.ldc = BT_B; g.bsC = BT_B * BT_B;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, Vt, CT_F64, Vt, CT_F64, S, g));
    }
    hipLaunchKernelGGL(larft_kernel, dim3(nb), dim3(256), 0, h->stream, (const double*)S, (const double*)taup, T);
    CT_HIP(hipGetLastError());
    {   // M_m = T_m^T Vt_m
        GemmDesc g;
        g.M = BT_B; g.N = (int)n; g.K = BT_B; g.batch = nb;
        g.a_kc = 0; g.sAm = 1; g.sAk = BT_B; g.bsA = BT_B * BT_B;
        g.b_kc = 0; g.sBk = n; g.sBn = 1; g.bsB = BT_B * n;
        g.ldc = n; g.bsC = BT_B * n;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, T, CT_F64, Vt, CT_F64, M, g));
    }
    return 0;
}

}  // namespace

// Symmetric eigendecomposition in two stages.  Stage 1: A (n x n, symmetric,
// destroyed) -> ascending eigenvalues w[n] and, in row j of Z[n][n], the
// eigenvector j of the tridiagonal matrix (LAPACK's column-major eigenvector
// matrix read row-major); it also builds the back-transformation factors.
// Device status words: info[0] from dstedc / dsyevd, info[1] set when the
// tridiagonalization's arrival wait timed out.  Above n = 2048 (or when the
// cooperative grid cannot be resident) it runs rocSOLVER dsyevd instead and
// sets *staged = false: Z then already holds the eigenvectors of A.
int syevd_stage1(Handle* h, double* A, int64_t n, double* w, double* Z, int* info, bool* staged,
                 bool force_rocsolver) {
    void* wsb;
    CT_CHECK(ws_get(h, WS_EIG2, (size_t)(2 * n + 8) * 8, &wsb));
    double* e = (double*)wsb;
    double* tau = e + n;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, h->stream, (unsigned*)info, 2);
    CT_HIP(hipGetLastError());
    bool ok = false;
    if (!force_rocsolver) CT_CHECK(sytrd_coop(h, A, n, w, e, tau, info + 1, &ok));
    rocblas_status st;
    if (ok) {
        st = rocsolver_dstedc(h->blas, rocblas_evect_tridiagonal, (rocblas_int)n, w, e, Z, (rocblas_int)n,
                              (rocblas_int*)info);
        if (st == rocblas_status_success) CT_CHECK(bt_prepare(h, A, tau, n));
    } else {
        CT_HIP(hipMemcpyAsync(Z, A, (size_t)(n * n) * 8, hipMemcpyDeviceToDevice, h->stream));
        st = rocsolver_dsyevd(h->blas, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, Z, (rocblas_int)n,
                              w, e, (rocblas_int*)info);
    }
    if (st != rocblas_status_success) {
        set_last_error(std::string("rocSOLVER eigensolver: ") + rocblas_status_to_string(st));
        return CT_STATUS_ROCBLAS_ERROR_BASE + (int)st;
    }
    *staged = ok;
    return 0;
}

// Stage 2: rows n-rows .. n-1 of Z (the eigenvectors of the `rows` largest
// eigenvalues) become eigenvectors of A.  Only after a staged stage 1 on the
// same handle with no other eigensolver call in between.
int syevd_stage2(Handle* h, int64_t n, double* Z, int64_t rows) {
    const int nb = bt_blocks(n);
    if (nb == 0 || rows <= 0) return 0;
    void *wv, *wm, *wst, *ww;
    CT_CHECK(ws_get(h, WS_EIG3, 0, &wv));
    CT_CHECK(ws_get(h, WS_EIG4, 0, &wm));
    CT_CHECK(ws_get(h, WS_EIG5, 0, &wst));
    CT_CHECK(ws_get(h, WS_TMP2, (size_t)(rows * BT_B) * 8, &ww));
    const double* Vt = (const double*)wv;
    const double* M = (const double*)wm;
    double* W = (double*)ww;
    double* E = Z + (n - rows) * n;
    for (int m = nb - 1; m >= 0; --m) {
        // Vt_m and M_m are zero in columns < m*B + 1; starting at m*B keeps the
        // operand rows 16-byte aligned and adds only zero products
        const int64_t off = (int64_t)m * BT_B;
        const int64_t K = n - off;
        GemmDesc g1;  // W[rows][B] = E[:, off:] Vt_m[:, off:]^T (E: the kept rows of Z)
        g1.M = (int)rows; g1.N = BT_B; g1.K = (int)K;
        g1.a_kc = 1; g1.sAm = n; g1.sAk = 1;
        g1.b_kc = 1; g1.sBn = n; g1.sBk = 1;
        g1.ldc = BT_B;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, E + off, CT_F64, Vt + (int64_t)m * BT_B * n + off, CT_F64, W, g1));
        GemmDesc g2;  // E[:, off:] -= W M_m[:, off:]
        g2.M = (int)rows; g2.N = (int)K; g2.K = BT_B;
        g2.a_kc = 1; g2.sAm = BT_B; g2.sAk = 1;
        g2.b_kc = 0; g2.sBk = n; g2.sBn = 1;
        g2.ldc = n; g2.alpha = -1.0; g2.beta = 1.0;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, W, CT_F64, M + (int64_t)m * BT_B * n + off, CT_F64, E + off, g2));
    }
    return 0;
}

int syevd(Handle* h, double* A, int64_t n, double* w, double* Z, int* info, bool force_rocsolver) {
    bool staged = false;
    CT_CHECK(syevd_stage1(h, A, n, w, Z, info, &staged, force_rocsolver));
    if (staged) CT_CHECK(syevd_stage2(h, n, Z, n));
    return 0;
}

}  // namespace ct
