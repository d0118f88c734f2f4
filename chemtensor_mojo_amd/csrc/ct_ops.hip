// Memory-bound kernels of the DMRG hot path on gfx950: the MPO mixing
// (the K = w*d contractions of apply_two_site_heff steps 2-3 and of the
// environment updates), generic permutes, conversions, fills and FP64
// wavefront reductions.
#include <algorithm>

#include "ct_internal.h"
#include "ct_device.h"

namespace ct {

namespace {

// ------------------------------------------------------------------ mixing
// Y[co, p, eo, q] = sum_{ci, ei} coef[(ci*EIN + ei)*NOUT + co*EOUT + eo] * X[ci, p, ei, q]
// One thread per pixel (p, q); q is the contiguous index on both sides, so
// every load and store of a wavefront is 64 consecutive elements.  All K
// inputs of a pixel are loaded before any arithmetic (K loads in flight per
// thread), and the coefficients are wave-uniform, so they are read through
// the scalar cache straight into SGPR operands of the FMAs.
constexpr int MIX_KMAX = 64;
constexpr int MIX_NMAX = 64;

struct MixArgs {
    int CIN, EIN, COUT, EOUT;
    int64_t P, Q;
    int64_t sXc, sXp, sXe, sYc, sYp, sYe;
    const int* active;
    uint16_t* yl;  // OUT 3: l planes (half the split rows' row stride)
};

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename T, int N>
struct alignas(sizeof(T) * N) VecN {
    T v[N];
};

// OUT: 0 = double, 1 = float, 2 = split BF16 hi/lo planes (uint16), 3 = the
// same plus the l planes of the 6xBF16 form (a.yl).
// KM: compile-time size (16, 20, 32 or 64) of the zero-padded KM x KM coefficient
// block `coef` (row k = input, column n = output).  Inputs k >= K re-read
// input K-1 and meet zero coefficients, so the unrolled loops carry no
// per-element guards; every coefficient address is a compile-time offset
// from a kernel argument, so the FMAs take them from SGPRs loaded through
// the scalar cache.  VQ: consecutive pixels q per thread, loaded and stored
// as one VQ-element vector; the launch guarantees Q % VQ == 0 and VQ-aligned
// bases and strides.
template <typename TX, typename TACC, int OUT, int KM, bool FULL, int VQ>
__global__ __launch_bounds__(256) void mix_kernel(const TX* __restrict__ X, void* __restrict__ Yv,
                                                  const TACC* __restrict__ coef, MixArgs a) {
    if (a.active && *a.active == 0) return;
    __shared__ int64_t in_off[KM];
    __shared__ int64_t out_off[KM];  // OUT 2: row part of the output offset
    __shared__ int64_t out_k[KM];    // OUT 2: k index of output channel co
    const int K = a.CIN * a.EIN, NO = a.COUT * a.EOUT;
    for (int k = threadIdx.x; k < KM; k += blockDim.x) {
        const int kk = k < K ? k : K - 1;
        const int ci = kk / a.EIN, ei = kk - ci * a.EIN;
        in_off[k] = ci * a.sXc + ei * a.sXe;
        const int nn = k < NO ? k : NO - 1;
        const int co = nn / a.EOUT, eo = nn - co * a.EOUT;
        if (OUT >= 2) {
            out_off[k] = eo * a.sYe;
            out_k[k] = co * a.sYc;
        } else {
            out_off[k] = co * a.sYc + eo * a.sYe;
        }
    }
    __syncthreads();
    const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * VQ;
    if (q >= a.Q) return;
    for (int64_t p = blockIdx.y; p < a.P; p += gridDim.y) {
        // Keeps the coefficient loads inside the loop: hoisting all KM*KM of
        // them out of it would spill.
        asm volatile("" ::: "memory");
        const TX* xp = X + p * a.sXp + q;
        TACC x[KM][VQ];
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            const VecN<TX, VQ> t = *(const VecN<TX, VQ>*)(xp + in_off[k]);
#pragma unroll
            for (int j = 0; j < VQ; ++j) x[k][j] = (TACC)t.v[j];
        }
        TACC acc[KM][VQ];
#pragma unroll
        for (int n = 0; n < KM; ++n)
#pragma unroll
            for (int j = 0; j < VQ; ++j) acc[n][j] = 0;
#pragma unroll
        for (int k = 0; k < KM; ++k)
#pragma unroll
            for (int n = 0; n < KM; ++n) {
                const TACC cf = coef[k * KM + n];
#pragma unroll
                for (int j = 0; j < VQ; ++j) acc[n][j] += cf * x[k][j];
            }
        // FULL (NO == KM): no guard, so the loop stays unrolled
#pragma unroll
        for (int n = 0; n < KM; ++n) {
            if (!FULL && n >= NO) break;
            if (OUT == 0) {
                VecN<double, VQ> o;
#pragma unroll
                for (int j = 0; j < VQ; ++j) o.v[j] = (double)acc[n][j];
                *(VecN<double, VQ>*)((double*)Yv + out_off[n] + p * a.sYp + q) = o;
            } else if (OUT == 1) {
                VecN<float, VQ> o;
#pragma unroll
                for (int j = 0; j < VQ; ++j) o.v[j] = (float)acc[n][j];
                *(VecN<float, VQ>*)((float*)Yv + out_off[n] + p * a.sYp + q) = o;
            } else {
                // split rows: hi at 2*(k & ~31) + (k & 31), lo 32 elements later
                // (q .. q+VQ-1 stay inside one 32-wide block: Q % 32 == 0 when VQ > 1)
                const int64_t k = out_k[n] + q;
                const int64_t o = out_off[n] + p * a.sYp + 2 * (k & ~(int64_t)31) + (k & 31);
                VecN<uint16_t, VQ> hv, lv, mv;
#pragma unroll
                for (int j = 0; j < VQ; ++j) {
                    const float f = (float)acc[n][j];
                    hv.v[j] = f2bf(f);
                    const float r = f - __uint_as_float(((uint32_t)hv.v[j]) << 16);
                    lv.v[j] = f2bf(r);
                    if (OUT == 3) mv.v[j] = f2bf(r - __uint_as_float(((uint32_t)lv.v[j]) << 16));
                }
                *(VecN<uint16_t, VQ>*)((uint16_t*)Yv + o) = hv;
                *(VecN<uint16_t, VQ>*)((uint16_t*)Yv + o + 32) = lv;
                if (OUT == 3) *(VecN<uint16_t, VQ>*)(a.yl + ((out_off[n] + p * a.sYp) >> 1) + k) = mv;
            }
        }
    }
}

// Zero-padded KM x KM copy of the [K][NO] FP64 coefficients in the
// accumulation type.
template <typename TACC>
__global__ void pad_coef_kernel(const double* __restrict__ in, TACC* __restrict__ out, int K, int NO, int KM) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= KM * KM) return;
    const int k = i / KM, n = i - k * KM;
    out[i] = (k < K && n < NO) ? (TACC)in[k * NO + n] : (TACC)0;
}

// ------------------------------------------------------------------ permute
struct PermArgs {
    int rank;
    int64_t n;
    int64_t out_shape[8];
    int64_t in_stride_perm[8];  // stride of input axis perm[i]
};

template <typename T>
__global__ __launch_bounds__(256) void permute_kernel(const T* __restrict__ in, T* __restrict__ out, PermArgs a) {
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < a.n; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t rem = o, src = 0;
        for (int i = a.rank - 1; i >= 0; --i) {
            const int64_t c = rem % a.out_shape[i];
            rem /= a.out_shape[i];
            src += c * a.in_stride_perm[i];
        }
        out[o] = in[src];
    }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void convert_kernel(const TI* __restrict__ in, TO* __restrict__ out, int64_t n,
                                                      const int* active) {
    if (active && *active == 0) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (TO)in[i];
}

// FP64 -> FP32 four elements per thread (two 16-byte loads, one 16-byte
// store); n % 4 == 0 and both pointers 16-byte aligned
__global__ __launch_bounds__(256) void convert4_kernel(const double2* __restrict__ in, float4* __restrict__ out,
                                                       int64_t n4, const int* active) {
    if (active && *active == 0) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = in[2 * i], b = in[2 * i + 1];
        out[i] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T* __restrict__ x, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void identity_env_kernel(T* __restrict__ x, int64_t w, int64_t D) {
    const int64_t n = w * D * D;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i % (D * D);
        x[i] = (r / D == r % D) ? (T)1 : (T)0;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void axpy_kernel(T* __restrict__ y, const T* __restrict__ x, double alpha, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (T)((double)y[i] + alpha * (double)x[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(T* __restrict__ x, double alpha, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (T)(alpha * (double)x[i]);
}

}  // namespace

// -------------------------------------------------------- FP64 reductions
// Two-pass deterministic dot product: RED_BLOCKS partial sums in a fixed
// order, then one block reduces them.  Shared with the Lanczos kernels.
template <typename T>
__global__ __launch_bounds__(256) void dot_partial_kernel(const T* __restrict__ x, const T* __restrict__ y, int64_t n,
                                                          double* __restrict__ partial, const int* active) {
    if (active && *active == 0) return;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += (double)x[i] * (double)y[i];
    s = block_sum256(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void reduce_final_kernel(const double* __restrict__ partial, int np,
                                                           double* __restrict__ out, const int* active) {
    if (active && *active == 0) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += partial[i];
    s = block_sum256(s);
    if (threadIdx.x == 0) *out = s;
}

template __global__ void dot_partial_kernel<double>(const double*, const double*, int64_t, double*, const int*);
template __global__ void dot_partial_kernel<float>(const float*, const float*, int64_t, double*, const int*);

static inline int grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (int)g;
}

namespace {

int mix_args(const MixDesc& m, MixArgs* a) {
    a->CIN = m.CIN; a->EIN = m.EIN; a->COUT = m.COUT; a->EOUT = m.EOUT;
    a->P = m.P; a->Q = m.Q;
    a->sXc = m.sXc; a->sXp = m.sXp; a->sXe = m.sXe;
    a->sYc = m.sYc; a->sYp = m.sYp; a->sYe = m.sYe;
    a->active = m.active;
    a->yl = nullptr;
    return 0;
}

// Padded block size: 20 fits the reference's w = 5 MPOs at d = 2 exactly (K = N_O = 20).
static int mix_km(int K, int NO) {
    const int m = K > NO ? K : NO;
    return m <= 16 ? 16 : m <= 20 ? 20 : m <= 32 ? 32 : 64;
}
static int mix_km(const MixArgs& a) { return mix_km(a.CIN * a.EIN, a.COUT * a.EOUT); }

template <typename TX, typename TACC, int OUT, int VQ>
int launch_mix_vq(Handle* h, const TX* X, void* Y, const TACC* c, const MixArgs& a, int NO, int KM) {
    const int64_t threads = a.Q / VQ;
    dim3 grid((unsigned)((threads + 255) / 256), (unsigned)(a.P < 65535 ? a.P : 65535)), block(256);
    // Every output channel real (NO == KM): the unguarded store loop keeps acc[]
    // in registers (a guarded loop is rolled and, in FP64 or with VQ > 1, puts
    // acc[] in scratch).  FP32 with VQ = 1 keeps the guarded loop in registers
    // at higher occupancy and measures faster.
    const bool full = NO == KM && (sizeof(TACC) == 8 || VQ > 1);
    if (full) {
        if (KM == 16)
            hipLaunchKernelGGL((mix_kernel<TX, TACC, OUT, 16, true, VQ>), grid, block, 0, h->stream, X, Y, c, a);
        else if (KM == 20)
            hipLaunchKernelGGL((mix_kernel<TX, TACC, OUT, 20, true, VQ>), grid, block, 0, h->stream, X, Y, c, a);
        else if (KM == 32)
            hipLaunchKernelGGL((mix_kernel<TX, TACC, OUT, 32, true, VQ>), grid, block, 0, h->stream, X, Y, c, a);
        else if constexpr (VQ == 1)
            hipLaunchKernelGGL((mix_kernel<TX, TACC, OUT, 64, true, 1>), grid, block, 0, h->stream, X, Y, c, a);
        return 0;
    }
    if constexpr (VQ == 1) {
        if (KM == 16)
            hipLaunchKernelGGL((mix_kernel<TX, TACC, OUT, 16, false, 1>), grid, block, 0, h->stream, X, Y, c, a);
        else if (KM == 20)
            hipLaunchKernelGGL((mix_kernel<TX, TACC, OUT, 20, false, 1>), grid, block, 0, h->stream, X, Y, c, a);
        else if (KM == 32)
            hipLaunchKernelGGL((mix_kernel<TX, TACC, OUT, 32, false, 1>), grid, block, 0, h->stream, X, Y, c, a);
        else
            hipLaunchKernelGGL((mix_kernel<TX, TACC, OUT, 64, false, 1>), grid, block, 0, h->stream, X, Y, c, a);
    }
    return 0;
}

// Padded coefficient block (workspace, unless `padded` already holds it) in
// the accumulation type, then the mixing launch.
template <typename TX, typename TACC, int OUT>
int launch_mix(Handle* h, const TX* X, void* Y, const double* d_coef, const MixArgs& a, const void* padded) {
    const int K = a.CIN * a.EIN, NO = a.COUT * a.EOUT, KM = mix_km(a);
    const TACC* c = (const TACC*)padded;
    if (!c) {
        void* w;
        CT_CHECK(ws_get(h, WS_MIX, (size_t)KM * KM * sizeof(TACC), &w));
        hipLaunchKernelGGL(pad_coef_kernel<TACC>, dim3((KM * KM + 255) / 256), dim3(256), 0, h->stream, d_coef,
                           (TACC*)w, K, NO, KM);
        c = (const TACC*)w;
    }
    // FP64 inputs: two consecutive pixels per thread (16-byte accesses) when
    // every output channel is real and every base, stride and Q allow it
    // (D = 1024, w = 8: 133 -> 127 us).  FP32 inputs keep one pixel per thread:
    // two or four per thread measured the same or slower (the split-row output
    // 63 -> 79 / 98 us), as the register use of KM * VQ inputs and sums cuts
    // the occupancy from 7 waves per SIMD to 3 or 1.
    constexpr int VQV = sizeof(TX) == 8 ? 2 : 1;
    const auto div = [](int64_t v) { return v % VQV == 0; };
    const bool vec = VQV > 1 && KM <= 32 && NO == KM && div(a.Q) && div(a.sXc) && div(a.sXp) && div(a.sXe) && div(a.sYc) && div(a.sYp) &&
                     div(a.sYe) && (uintptr_t)X % 16 == 0 &&
                     (OUT >= 2 ? (a.Q % 32 == 0 && a.sYc % 32 == 0 && (uintptr_t)Y % (size_t)(VQV * 2) == 0)
                               : (uintptr_t)Y % (size_t)(VQV * (OUT == 0 ? 8 : 4)) == 0);
    if (vec) return launch_mix_vq<TX, TACC, OUT, VQV>(h, X, Y, c, a, NO, KM);
    return launch_mix_vq<TX, TACC, OUT, 1>(h, X, Y, c, a, NO, KM);
}

// ------------------------------------------------ large MPO bond dimension
// When w*d (environments) or w*d^2 (H_eff) exceeds 64 the per-pixel mixing
// is a real matrix product: per pixel row p, Y_p[n][q] = sum_k coef[k][n] X_p[k][q]
// with K, NO up to thousands.  X is packed to [P][K][Q], one batched MFMA
// GEMM (FP64, or FP32 for FP32 contractions) forms [P][NO][Q], and an unpack
// kernel writes Y's layout.  The 3xBF16 form (mix_large_split) splits both
// operands and runs the three-product BF16 GEMM instead.
template <typename TX, typename TT>
__global__ void mix_pack_kernel(const TX* __restrict__ X, TT* __restrict__ Xt, MixArgs a) {
    if (a.active && *a.active == 0) return;
    const int K = a.CIN * a.EIN;
    const int64_t total = a.P * (int64_t)K * a.Q;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = t % a.Q, pk = t / a.Q;
        const int k = (int)(pk % K);
        const int64_t p = pk / K;
        const int ci = k / a.EIN, ei = k - ci * a.EIN;
        Xt[t] = (TT)X[ci * a.sXc + p * a.sXp + ei * a.sXe + q];
    }
}

template <typename TT, int OUT>
__global__ void mix_unpack_kernel(const TT* __restrict__ Yt, void* __restrict__ Yv, MixArgs a) {
    if (a.active && *a.active == 0) return;
    const int NO = a.COUT * a.EOUT;
    const int64_t total = a.P * (int64_t)NO * a.Q;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = t % a.Q, pn = t / a.Q;
        const int n = (int)(pn % NO);
        const int64_t p = pn / NO;
        const int co = n / a.EOUT, eo = n - co * a.EOUT;
        const double v = (double)Yt[t];
        if (OUT == 0) {
            ((double*)Yv)[co * a.sYc + p * a.sYp + eo * a.sYe + q] = v;
        } else if (OUT == 1) {
            ((float*)Yv)[co * a.sYc + p * a.sYp + eo * a.sYe + q] = (float)v;
        } else {
            // split rows (OUT 2), plus the l plane at half the row stride (OUT 3)
            const int64_t k = co * a.sYc + q;
            const int64_t row = eo * a.sYe + p * a.sYp;
            const int64_t o = row + 2 * (k & ~(int64_t)31) + (k & 31);
            const float f = (float)v;
            const uint16_t hb = f2bf(f);
            const float r = f - __uint_as_float(((uint32_t)hb) << 16);
            const uint16_t lb = f2bf(r);
            ((uint16_t*)Yv)[o] = hb;
            ((uint16_t*)Yv)[o + 32] = lb;
            if (OUT == 3) a.yl[(row >> 1) + k] = f2bf(r - __uint_as_float(((uint32_t)lb) << 16));
        }
    }
}

// Split-rows unpack with four neighbouring q per thread (Q, sYc, sYe, sYp
// multiples of 4): each thread writes its four hi and four lo BF16 values as
// one 8-byte store each.
template <bool L3>
__global__ void mix_unpack_split4_kernel(const float* __restrict__ Yt, uint16_t* __restrict__ Y, MixArgs a) {
    if (a.active && *a.active == 0) return;
    const int NO = a.COUT * a.EOUT;
    const int64_t Q4 = a.Q >> 2;
    const int64_t total = a.P * (int64_t)NO * Q4;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = (t % Q4) << 2, pn = t / Q4;
        const int n = (int)(pn % NO);
        const int64_t p = pn / NO;
        const int co = n / a.EOUT, eo = n - co * a.EOUT;
        const float4 v = *(const float4*)(Yt + pn * a.Q + q);
        const int64_t k = co * a.sYc + q;  // k % 4 == 0: the four k share a 32-block
        const int64_t row = eo * a.sYe + p * a.sYp;
        const int64_t o = row + 2 * (k & ~(int64_t)31) + (k & 31);
        const float f[4] = {v.x, v.y, v.z, v.w};
        uint16_t hi[4], lo[4], l3[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            hi[i] = f2bf(f[i]);
            const float r = f[i] - __uint_as_float(((uint32_t)hi[i]) << 16);
            lo[i] = f2bf(r);
            l3[i] = f2bf(r - __uint_as_float(((uint32_t)lo[i]) << 16));
        }
        *(uint2*)(Y + o) = make_uint2((uint32_t)hi[0] | ((uint32_t)hi[1] << 16), (uint32_t)hi[2] | ((uint32_t)hi[3] << 16));
        *(uint2*)(Y + o + 32) =
            make_uint2((uint32_t)lo[0] | ((uint32_t)lo[1] << 16), (uint32_t)lo[2] | ((uint32_t)lo[3] << 16));
        if (L3)  // the l plane: half the row stride, the same k
            *(uint2*)(a.yl + (row >> 1) + k) =
                make_uint2((uint32_t)l3[0] | ((uint32_t)l3[1] << 16), (uint32_t)l3[2] | ((uint32_t)l3[3] << 16));
    }
}

template <typename TX, typename TT, int OUT>
int mix_large(Handle* h, const TX* X, void* Y, const double* d_coef, const MixArgs& a) {
    const int K = a.CIN * a.EIN, NO = a.COUT * a.EOUT;
    void *wx, *wy;
    CT_CHECK(ws_get(h, WS_MIXA, (size_t)(a.P * K * a.Q) * sizeof(TT), &wx));
    CT_CHECK(ws_get(h, WS_MIXB, (size_t)(a.P * NO * a.Q) * sizeof(TT), &wy));
    const int64_t nin = a.P * (int64_t)K * a.Q, nout = a.P * (int64_t)NO * a.Q;
    hipLaunchKernelGGL((mix_pack_kernel<TX, TT>), dim3((unsigned)std::min<int64_t>((nin + 255) / 256, 16384)),
                       dim3(256), 0, h->stream, X, (TT*)wx, a);
    CT_HIP(hipGetLastError());
    GemmDesc g;  // Yt[p][n][q] = sum_k coef[k][n] Xt[p][k][q]
    g.M = NO; g.N = (int)a.Q; g.K = K; g.batch = (int)a.P;
    g.a_kc = 0; g.sAm = 1; g.sAk = NO; g.bsA = 0;
    g.b_kc = 0; g.sBk = a.Q; g.sBn = 1; g.bsB = (int64_t)K * a.Q;
    g.ldc = a.Q; g.bsC = (int64_t)NO * a.Q;
    g.active = a.active;
    const int tt = sizeof(TT) == 8 ? CT_F64 : CT_F32;
    CT_CHECK(gemm(h, tt == CT_F64 ? CT_PREC_FP64 : CT_PREC_FP32, CT_F64, d_coef, tt, wx, tt, wy, g));
    hipLaunchKernelGGL((mix_unpack_kernel<TT, OUT>), dim3((unsigned)std::min<int64_t>((nout + 255) / 256, 16384)),
                       dim3(256), 0, h->stream, (const TT*)wy, Y, a);
    CT_HIP(hipGetLastError());
    return 0;
}

// 3xBF16 / 6xBF16 forms with split-rows output: the coefficients (rows n) and
// X (rows q of each pixel row p, k = (ci, ei)) are split once into bf16 rows
// (6xBF16: with their l planes) and the three- or six-product NT GEMM forms
// Yt[p][n][q] in FP32 before the unpack (6xBF16: also into the l planes, Yl).
int mix_large_split(Handle* h, const float* X, uint16_t* Y, const double* d_coef, const MixArgs& a, uint16_t* Yl) {
    const int K = a.CIN * a.EIN, NO = a.COUT * a.EOUT;
    const int64_t K32 = split_k32(K), ld = 2 * K32;
    void *wx, *wy, *wc, *wxl = nullptr, *wcl = nullptr;
    CT_CHECK(ws_get(h, WS_MIXA, (size_t)(a.P * a.Q * ld) * 2, &wx));
    CT_CHECK(ws_get(h, WS_MIXB, (size_t)(a.P * NO * a.Q) * 4, &wy));
    CT_CHECK(ws_get(h, WS_MIXC, (size_t)(NO * ld) * 2, &wc));
    if (Yl) {
        CT_CHECK(ws_get(h, WS_MIXA_L, (size_t)(a.P * a.Q * K32) * 2, &wxl));
        CT_CHECK(ws_get(h, WS_MIXC_L, (size_t)(NO * K32) * 2, &wcl));
    }
    CT_CHECK(split_bf16(h, CT_F64, d_coef, NO, K, 1, NO, 1, 0, (uint16_t*)wc, ld, 0, a.active, (uint16_t*)wcl, K32,
                        0));
    CT_CHECK(split_bf16_kk(h, CT_F32, X, (int)a.Q, K, a.EIN, a.sXe, a.sXc, (int)a.P, a.sXp, (uint16_t*)wx, ld,
                           a.Q * ld, a.active, (uint16_t*)wxl, K32, a.Q * K32));
    SplitOp A, B;
    A.p = (const uint16_t*)wc; A.ld = ld; A.bs = 0;
    B.p = (const uint16_t*)wx; B.ld = ld; B.bs = a.Q * ld;
    if (Yl) {
        A.l = (const uint16_t*)wcl; A.ldl = K32; A.bsl = 0;
        B.l = (const uint16_t*)wxl; B.ldl = K32; B.bsl = a.Q * K32;
        CT_CHECK(gemm_bf16x6_nt(h, A, B, CT_F32, wy, NO, (int)a.Q, K, a.Q, (int64_t)NO * a.Q, (int)a.P, 1.0, 0.0,
                                a.active));
    } else {
        CT_CHECK(gemm_bf16x3_nt(h, A, B, CT_F32, wy, NO, (int)a.Q, K, a.Q, (int64_t)NO * a.Q, (int)a.P, 1.0, 0.0,
                                a.active));
    }
    MixArgs au = a;
    au.yl = Yl;
    const int64_t nout = a.P * (int64_t)NO * a.Q;
    if (a.Q % 4 == 0 && a.sYc % 4 == 0 && a.sYe % 4 == 0 && a.sYp % 4 == 0 && ((uintptr_t)Y % 8) == 0 &&
        ((uintptr_t)Yl % 8) == 0) {
        const dim3 grid((unsigned)std::min<int64_t>((nout / 4 + 255) / 256, 16384));
        if (Yl)
            hipLaunchKernelGGL(mix_unpack_split4_kernel<true>, grid, dim3(256), 0, h->stream, (const float*)wy, Y, au);
        else
            hipLaunchKernelGGL(mix_unpack_split4_kernel<false>, grid, dim3(256), 0, h->stream, (const float*)wy, Y, au);
    } else {
        const dim3 grid((unsigned)std::min<int64_t>((nout + 255) / 256, 16384));
        if (Yl)
            hipLaunchKernelGGL((mix_unpack_kernel<float, 3>), grid, dim3(256), 0, h->stream, (const float*)wy, (void*)Y,
                               au);
        else
            hipLaunchKernelGGL((mix_unpack_kernel<float, 2>), grid, dim3(256), 0, h->stream, (const float*)wy, (void*)Y,
                               au);
    }
    CT_HIP(hipGetLastError());
    return 0;
}

bool mix_is_large(const MixArgs& a) { return a.CIN * a.EIN > MIX_KMAX || a.COUT * a.EOUT > MIX_NMAX; }

}  // namespace

size_t mix_pad_bytes(int K, int NO, int tAcc) {
    const int KM = mix_km(K, NO);
    return (size_t)KM * KM * (tAcc == CT_F64 ? 8 : 4);
}

int mix_pad_coef(Handle* h, const double* d_coef, int K, int NO, int tAcc, void* dst) {
    if (K > MIX_KMAX || NO > MIX_NMAX) return 0;  // the large-w path reads the FP64 coefficients directly
    const int KM = mix_km(K, NO);
    if (tAcc == CT_F64)
        hipLaunchKernelGGL(pad_coef_kernel<double>, dim3((KM * KM + 255) / 256), dim3(256), 0, h->stream, d_coef,
                           (double*)dst, K, NO, KM);
    else
        hipLaunchKernelGGL(pad_coef_kernel<float>, dim3((KM * KM + 255) / 256), dim3(256), 0, h->stream, d_coef,
                           (float*)dst, K, NO, KM);
    CT_HIP(hipGetLastError());
    return 0;
}

int mix(Handle* h, int tX, const void* X, int tY, void* Y, const double* d_coef, const MixDesc& m,
        const void* padded) {
    if (m.P <= 0 || m.Q <= 0) return 0;
    MixArgs a;
    CT_CHECK(mix_args(m, &a));
    if (mix_is_large(a)) {
        if (tX == CT_F64 && tY == CT_F64) return mix_large<double, double, 0>(h, (const double*)X, Y, d_coef, a);
        if (tX == CT_F64 && tY == CT_F32) return mix_large<double, double, 1>(h, (const double*)X, Y, d_coef, a);
        if (tX == CT_F32 && tY == CT_F32) return mix_large<float, float, 1>(h, (const float*)X, Y, d_coef, a);
        if (tX == CT_F32 && tY == CT_F64) return mix_large<float, float, 0>(h, (const float*)X, Y, d_coef, a);
        set_last_error("mix: unsupported types");
        return CT_STATUS_NOT_SUPPORTED;
    }
    if (tX == CT_F64 && tY == CT_F64) {
        CT_CHECK((launch_mix<double, double, 0>(h, (const double*)X, Y, d_coef, a, padded)));
    } else if (tX == CT_F64 && tY == CT_F32) {
        CT_CHECK((launch_mix<double, double, 1>(h, (const double*)X, Y, d_coef, a, padded)));
    } else if (tX == CT_F32 && tY == CT_F32) {
        CT_CHECK((launch_mix<float, float, 1>(h, (const float*)X, Y, d_coef, a, padded)));
    } else if (tX == CT_F32 && tY == CT_F64) {
        CT_CHECK((launch_mix<float, float, 0>(h, (const float*)X, Y, d_coef, a, padded)));
    } else {
        set_last_error("mix: unsupported types");
        return CT_STATUS_NOT_SUPPORTED;
    }
    CT_HIP(hipGetLastError());
    return 0;
}

// Mixing whose output goes straight into 3xBF16 split rows (and, with Yl, the
// l planes of the 6xBF16 form); X is FP32.
int mix_split(Handle* h, const float* X, uint16_t* Y, const double* d_coef, const MixDesc& m,
              const void* padded, uint16_t* Yl) {
    if (m.P <= 0 || m.Q <= 0) return 0;
    MixArgs a;
    CT_CHECK(mix_args(m, &a));
    if (mix_is_large(a)) return mix_large_split(h, X, Y, d_coef, a, Yl);
    if (Yl) {
        a.yl = Yl;
        CT_CHECK((launch_mix<float, float, 3>(h, X, (void*)Y, d_coef, a, padded)));
    } else {
        CT_CHECK((launch_mix<float, float, 2>(h, X, (void*)Y, d_coef, a, padded)));
    }
    CT_HIP(hipGetLastError());
    return 0;
}

__global__ void split_zero_pad_kernel(uint16_t* __restrict__ out, int64_t rows, int64_t K, int64_t npad,
                                      int64_t ld, uint16_t* __restrict__ outl, int64_t ldl) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * npad) return;
    const int64_t r = e / npad, k = K + (e - r * npad);
    uint16_t* o = out + r * ld + 2 * (k & ~(int64_t)31) + (k & 31);
    o[0] = 0;
    o[32] = 0;
    if (outl) outl[r * ldl + k] = 0;
}

int split_zero_pad(Handle* h, uint16_t* out, int64_t rows, int64_t K, int64_t ld, uint16_t* outl, int64_t ldl) {
    const int64_t npad = split_k32(K) - K;
    if (npad == 0 || rows <= 0) return 0;
    const int64_t n = rows * npad;
    hipLaunchKernelGGL(split_zero_pad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, out, rows,
                       K, npad, ld, outl, ldl);
    CT_HIP(hipGetLastError());
    return 0;
}

int permute(Handle* h, int dtype, const void* in, int rank, const int64_t* shape, const int* perm, void* out) {
    PermArgs a;
    if (rank < 1 || rank > 8) {
        set_last_error("permute: rank must be 1..8");
        return CT_STATUS_NOT_SUPPORTED;
    }
    int64_t stride[8];
    stride[rank - 1] = 1;
    for (int i = rank - 2; i >= 0; --i) stride[i] = stride[i + 1] * shape[i + 1];
    a.rank = rank;
    a.n = 1;
    for (int i = 0; i < rank; ++i) {
        a.out_shape[i] = shape[perm[i]];
        a.in_stride_perm[i] = stride[perm[i]];
        a.n *= shape[i];
    }
    if (a.n == 0) return 0;
    if (dtype == CT_F64)
        hipLaunchKernelGGL(permute_kernel<double>, dim3(grid_for(a.n)), dim3(256), 0, h->stream, (const double*)in,
                           (double*)out, a);
    else
        hipLaunchKernelGGL(permute_kernel<float>, dim3(grid_for(a.n)), dim3(256), 0, h->stream, (const float*)in,
                           (float*)out, a);
    CT_HIP(hipGetLastError());
    return 0;
}

int convert(Handle* h, int tin, const void* in, int tout, void* out, int64_t n, const int* active) {
    if (n <= 0) return 0;
    dim3 g(grid_for(n)), b(256);
    if (tin == CT_F64 && tout == CT_F32 && n % 4 == 0 && (uintptr_t)in % 16 == 0 && (uintptr_t)out % 16 == 0)
        hipLaunchKernelGGL(convert4_kernel, dim3(grid_for(n / 4)), b, 0, h->stream, (const double2*)in, (float4*)out,
                           n / 4, active);
    else if (tin == CT_F64 && tout == CT_F32)
        hipLaunchKernelGGL((convert_kernel<double, float>), g, b, 0, h->stream, (const double*)in, (float*)out, n,
                           active);
    else if (tin == CT_F32 && tout == CT_F64)
        hipLaunchKernelGGL((convert_kernel<float, double>), g, b, 0, h->stream, (const float*)in, (double*)out, n,
                           active);
    else if (tin == tout && active) {
        set_last_error("convert: a gated copy needs different types");
        return CT_STATUS_NOT_SUPPORTED;
    } else if (tin == tout) {
        CT_HIP(hipMemcpyAsync(out, in, (size_t)n * (tin == CT_F64 ? 8 : 4), hipMemcpyDeviceToDevice, h->stream));
        return 0;
    } else {
        set_last_error("convert: unsupported types");
        return CT_STATUS_NOT_SUPPORTED;
    }
    CT_HIP(hipGetLastError());
    return 0;
}

int fill(Handle* h, int dtype, void* x, int64_t n, double value) {
    if (n <= 0) return 0;
    if (dtype == CT_F64)
        hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, (double*)x, n, value);
    else
        hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(256), 0, h->stream, (float*)x, n, (float)value);
    CT_HIP(hipGetLastError());
    return 0;
}

int set_identity_env(Handle* h, int dtype, void* x, int64_t w, int64_t D) {
    const int64_t n = w * D * D;
    if (n <= 0) return 0;
    if (dtype == CT_F64)
        hipLaunchKernelGGL(identity_env_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, (double*)x, w, D);
    else
        hipLaunchKernelGGL(identity_env_kernel<float>, dim3(grid_for(n)), dim3(256), 0, h->stream, (float*)x, w, D);
    CT_HIP(hipGetLastError());
    return 0;
}

constexpr int RED_BLOCKS = 512;

int dot_dev(Handle* h, int dtype, const void* x, const void* y, int64_t n, double* d_out, const int* active) {
    void* part;
    CT_CHECK(ws_get(h, WS_PARTIAL, RED_BLOCKS * sizeof(double), &part));
    if (dtype == CT_F64)
        hipLaunchKernelGGL(dot_partial_kernel<double>, dim3(RED_BLOCKS), dim3(256), 0, h->stream, (const double*)x,
                           (const double*)y, n, (double*)part, active);
    else
        hipLaunchKernelGGL(dot_partial_kernel<float>, dim3(RED_BLOCKS), dim3(256), 0, h->stream, (const float*)x,
                           (const float*)y, n, (double*)part, active);
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(256), 0, h->stream, (const double*)part, RED_BLOCKS, d_out,
                       active);
    CT_HIP(hipGetLastError());
    return 0;
}

int dot_host(Handle* h, int dtype, const void* x, const void* y, int64_t n, double* out) {
    CT_CHECK(dot_dev(h, dtype, x, y, n, h->d_scalars, nullptr));
    CT_HIP(hipMemcpyAsync(h->h_pinned, h->d_scalars, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    CT_HIP(hipStreamSynchronize(h->stream));
    *out = h->h_pinned[0];
    return 0;
}

int axpy(Handle* h, int dtype, double alpha, const void* x, void* y, int64_t n) {
    if (n <= 0) return 0;
    if (dtype == CT_F64)
        hipLaunchKernelGGL(axpy_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, (double*)y,
                           (const double*)x, alpha, n);
    else
        hipLaunchKernelGGL(axpy_kernel<float>, dim3(grid_for(n)), dim3(256), 0, h->stream, (float*)y,
                           (const float*)x, alpha, n);
    CT_HIP(hipGetLastError());
    return 0;
}

int scale(Handle* h, int dtype, double alpha, void* x, int64_t n) {
    if (n <= 0) return 0;
    if (dtype == CT_F64)
        hipLaunchKernelGGL(scale_kernel<double>, dim3(grid_for(n)), dim3(256), 0, h->stream, (double*)x, alpha, n);
    else
        hipLaunchKernelGGL(scale_kernel<float>, dim3(grid_for(n)), dim3(256), 0, h->stream, (float*)x, alpha, n);
    CT_HIP(hipGetLastError());
    return 0;
}

}  // namespace ct
