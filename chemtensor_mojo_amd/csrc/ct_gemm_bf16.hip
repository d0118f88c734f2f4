// 3xBF16 split-precision GEMM on gfx950 MFMA (v_mfma_f32_16x16x32_bf16).
//
// Each FP32/FP64 operand x is split once into hi = bf16(x) and
// lo = bf16(x - hi).  The product a*b is then formed as
//   ah*bh + ah*bl + al*bh
// with every BF16xBF16 product exact in FP32 and accumulated in FP32; the
// dropped al*bl term and the rounding of lo bound the relative error of one
// product by about 2^-16.
//
// The split planes are stored k-contiguous ([row][k]) so an MFMA fragment
// (8 consecutive k of one row per lane) is one 16-byte read, in LDS and in
// HBM.  The kernel therefore has a single "NT" form C = A * B^T; callers that
// hold m- or n-contiguous operands split them with a transposing split.
//
// Two tilings share the fragment code:
//  * 256x256x32 blocks, 512 threads = 8 waves (2 x 4), 128x64 per wave, one
//    80 KiB LDS stage refilled from registers prefetched during the MFMAs;
//    optional split-K over blockIdx.z with FP32 partial sums reduced by a
//    second kernel.  Each staged byte feeds twice the MFMA work of the small
//    tiling, which halves the L2/HBM operand traffic per FLOP.
//  * 128x128x32 blocks, 256 threads = 4 waves (2x2), 64x64 per wave, LDS
//    double-buffered (2 x 40 KiB), for problems too small to fill the chip
//    with 256x256 tiles.
#include <algorithm>

#include "ct_internal.h"
#include "ct_device.h"

namespace ct {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan: truncate
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

template <typename T>
__device__ __forceinline__ void split_one(T v, uint16_t& hi, uint16_t& lo) {
    float x = (float)v;
    hi = f32_to_bf16_rne(x);
    lo = f32_to_bf16_rne(x - bf16_to_f32(hi));
}

// ---------------------------------------------------------------- split
// Generic strided split.  One thread per (row, k) output element; blocks are
// 32 (k) x 8 (rows) and stage through LDS so both the strided read and the
// k-contiguous write are coalesced whichever of s_row / s_k is 1.
template <typename T>
__global__ __launch_bounds__(256) void split_kernel(const T* __restrict__ X, int rows, int K, int64_t s_row,
                                                    int64_t s_k, int64_t bsX, uint16_t* __restrict__ hi,
                                                    uint16_t* __restrict__ lo, int64_t ld, int64_t bs,
                                                    const int* active) {
    if (active && *active == 0) return;
    __shared__ float tile[32][33];
    const int b = blockIdx.z;
    const int r0 = blockIdx.y * 32, k0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    X += (int64_t)b * bsX;
    if (s_row == 1) {
        // rows contiguous: read with tx along rows
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int k = k0 + ty + 8 * i, r = r0 + tx;
            tile[ty + 8 * i][tx] = (k < K && r < rows) ? (float)X[(int64_t)r + (int64_t)k * s_k] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int r = r0 + ty + 8 * i, k = k0 + tx;
            if (r < rows && k < K) {
                uint16_t h, l;
                split_one(tile[tx][ty + 8 * i], h, l);
                int64_t o = (int64_t)b * bs + (int64_t)r * ld + k;
                hi[o] = h;
                lo[o] = l;
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int r = r0 + ty + 8 * i, k = k0 + tx;
            if (r < rows && k < K) {
                uint16_t h, l;
                split_one(X[(int64_t)r * s_row + (int64_t)k * s_k], h, l);
                int64_t o = (int64_t)b * bs + (int64_t)r * ld + k;
                hi[o] = h;
                lo[o] = l;
            }
        }
    }
}

// ------------------------------------------------------------------ GEMM
constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;  // LDS row = 40 bf16 = 80 B

struct NtArgs {
    const uint16_t *Ah, *Al, *Bh, *Bl;
    void* C;
    int M, N, K;
    int tilesM, tilesN;
    int vec;  // 16-byte loads legal (ld % 8 == 0, 16B aligned, K-tail handled separately)
    int64_t lda, ldb, bsA, bsB, ldc, bsC;
    double alpha, beta;
    const int* active;
    int GM;           // tile rows per L2 group (see gemm_tile_coords)
    int splits;       // split-K factor (256x256 tiling)
    int kchunk;       // K range per split, multiple of BK
    float* part;      // [batch*splits][M][N] FP32 partial sums when splits > 1
};

// Each thread copies one 32-byte half row (16 bf16) of each of the four
// planes: row = tid % 128, k-half = tid / 128.
__device__ __forceinline__ void nt_load(const uint16_t* P, int64_t ld, int R0, int RN, int k0, int K, bool vec,
                                        v4u* r) {
    const int row = R0 + (threadIdx.x & 127), kh = threadIdx.x >> 7;
    const int kb = k0 + 16 * kh;
    if (vec && row < RN && kb + 15 < K) {
        const v4u* p = (const v4u*)(P + (int64_t)row * ld + kb);
        r[0] = p[0];
        r[1] = p[1];
    } else {
        uint16_t t[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) t[e] = (row < RN && kb + e < K) ? P[(int64_t)row * ld + kb + e] : (uint16_t)0;
#pragma unroll
        for (int q = 0; q < 2; ++q)
            r[q] = v4u{(uint32_t)t[8 * q + 0] | ((uint32_t)t[8 * q + 1] << 16),
                       (uint32_t)t[8 * q + 2] | ((uint32_t)t[8 * q + 3] << 16),
                       (uint32_t)t[8 * q + 4] | ((uint32_t)t[8 * q + 5] << 16),
                       (uint32_t)t[8 * q + 6] | ((uint32_t)t[8 * q + 7] << 16)};
    }
}

__device__ __forceinline__ void nt_store(uint16_t (*S)[LDK], const v4u* r) {
    const int row = threadIdx.x & 127, kh = threadIdx.x >> 7;
    v4u* p = (v4u*)&S[row][16 * kh];
    p[0] = r[0];
    p[1] = r[1];
}

template <typename TO>
__global__ __launch_bounds__(256) void gemm_bf16x3_nt_kernel(NtArgs p) {
    if (p.active && *p.active == 0) return;
    __shared__ __attribute__((aligned(16))) uint16_t sAh[2][BM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[2][BM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[2][BN][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[2][BN][LDK];

    int tm, tn;
    gemm_tile_coords(blockIdx.x, gridDim.x, p.tilesM, p.tilesN, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int z = blockIdx.z;
    const uint16_t* Ah = p.Ah + (int64_t)z * p.bsA;
    const uint16_t* Al = p.Al + (int64_t)z * p.bsA;
    const uint16_t* Bh = p.Bh + (int64_t)z * p.bsB;
    const uint16_t* Bl = p.Bl + (int64_t)z * p.bsB;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int li = lane & 15, lk = lane >> 4;

    v4f acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    v4u rah[2], ral[2], rbh[2], rbl[2];
    const bool vec = p.vec != 0;
    nt_load(Ah, p.lda, m0, p.M, 0, p.K, vec, rah);
    nt_load(Al, p.lda, m0, p.M, 0, p.K, vec, ral);
    nt_load(Bh, p.ldb, n0, p.N, 0, p.K, vec, rbh);
    nt_load(Bl, p.ldb, n0, p.N, 0, p.K, vec, rbl);
    nt_store(sAh[0], rah);
    nt_store(sAl[0], ral);
    nt_store(sBh[0], rbh);
    nt_store(sBl[0], rbl);
    __syncthreads();

    int buf = 0;
    for (int k0 = 0; k0 < p.K; k0 += BK) {
        const bool more = k0 + BK < p.K;
        if (more) {
            nt_load(Ah, p.lda, m0, p.M, k0 + BK, p.K, vec, rah);
            nt_load(Al, p.lda, m0, p.M, k0 + BK, p.K, vec, ral);
            nt_load(Bh, p.ldb, n0, p.N, k0 + BK, p.K, vec, rbh);
            nt_load(Bl, p.ldb, n0, p.N, k0 + BK, p.K, vec, rbl);
        }
        v8bf ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = wm * 64 + i * 16 + li;
            ah[i] = __builtin_bit_cast(v8bf, *(const v4u*)&sAh[buf][row][8 * lk]);
            al[i] = __builtin_bit_cast(v8bf, *(const v4u*)&sAl[buf][row][8 * lk]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = wn * 64 + j * 16 + li;
            bh[j] = __builtin_bit_cast(v8bf, *(const v4u*)&sBh[buf][col][8 * lk]);
            bl[j] = __builtin_bit_cast(v8bf, *(const v4u*)&sBl[buf][col][8 * lk]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        if (more) {
            nt_store(sAh[buf ^ 1], rah);
            nt_store(sAl[buf ^ 1], ral);
            nt_store(sBh[buf ^ 1], rbh);
            nt_store(sBl[buf ^ 1], rbl);
        }
        __syncthreads();
        buf ^= 1;
    }

    TO* C = (TO*)p.C + (int64_t)z * p.bsC;
    const double alpha = p.alpha, beta = p.beta;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + wn * 64 + j * 16 + li;
            if (n >= p.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm * 64 + i * 16 + lk * 4 + r;
                if (m < p.M) {
                    TO* c = C + (int64_t)m * p.ldc + n;
                    double v = alpha * (double)acc[i][j][r];
                    if (beta != 0.0) v += beta * (double)(*c);
                    *c = (TO)v;
                }
            }
        }
}

// ------------------------------------------------------- 256x256 tiling
constexpr int BB = 256;  // block rows of A and of B

// Thread t copies 32 bytes (16 bf16) of row t % 256, k-half t / 256, of one
// plane: two 16-byte loads.
__device__ __forceinline__ void nt_load256(const uint16_t* P, int64_t ld, int R0, int RN, int k0, int kend, bool vec,
                                           v4u* r) {
    const int row = R0 + (threadIdx.x & 255), kh = threadIdx.x >> 8;
    const int kb = k0 + 16 * kh;
    if (vec && row < RN && kb + 15 < kend) {
        const v4u* q = (const v4u*)(P + (int64_t)row * ld + kb);
        r[0] = q[0];
        r[1] = q[1];
    } else {
        uint16_t t[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) t[e] = (row < RN && kb + e < kend) ? P[(int64_t)row * ld + kb + e] : (uint16_t)0;
#pragma unroll
        for (int q = 0; q < 2; ++q)
            r[q] = v4u{(uint32_t)t[8 * q + 0] | ((uint32_t)t[8 * q + 1] << 16),
                       (uint32_t)t[8 * q + 2] | ((uint32_t)t[8 * q + 3] << 16),
                       (uint32_t)t[8 * q + 4] | ((uint32_t)t[8 * q + 5] << 16),
                       (uint32_t)t[8 * q + 6] | ((uint32_t)t[8 * q + 7] << 16)};
    }
}

__device__ __forceinline__ void nt_store256(uint16_t (*S)[LDK], const v4u* r) {
    const int row = threadIdx.x & 255, kh = threadIdx.x >> 8;
    v4u* q = (v4u*)&S[row][16 * kh];
    q[0] = r[0];
    q[1] = r[1];
}

// SPLIT: write FP32 partial sums of this block's K range to p.part instead
// of alpha*acc + beta*C.
template <typename TO, bool SPLIT>
__global__ __launch_bounds__(512) void gemm_bf16x3_nt256_kernel(NtArgs p) {
    if (p.active && *p.active == 0) return;
    __shared__ __attribute__((aligned(16))) uint16_t sAh[BB][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[BB][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[BB][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[BB][LDK];

    int tm, tn;
    gemm_tile_coords(blockIdx.x, gridDim.x, p.tilesM, p.tilesN, p.GM, tm, tn);
    const int m0 = tm * BB, n0 = tn * BB;
    const int z = blockIdx.z;
    const int b = z / p.splits, sk = z - b * p.splits;
    const int kbeg = sk * p.kchunk;
    const int kend = min(p.K, kbeg + p.kchunk);
    const uint16_t* Ah = p.Ah + (int64_t)b * p.bsA;
    const uint16_t* Al = p.Al + (int64_t)b * p.bsA;
    const uint16_t* Bh = p.Bh + (int64_t)b * p.bsB;
    const uint16_t* Bl = p.Bl + (int64_t)b * p.bsB;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 2, wn = wave & 3;  // 2 x 4 waves, 128 x 64 each
    const int li = lane & 15, lk = lane >> 4;

    v4f acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    v4u rah[2], ral[2], rbh[2], rbl[2];
    const bool vec = p.vec != 0;
    if (kbeg < kend) {
        nt_load256(Ah, p.lda, m0, p.M, kbeg, kend, vec, rah);
        nt_load256(Al, p.lda, m0, p.M, kbeg, kend, vec, ral);
        nt_load256(Bh, p.ldb, n0, p.N, kbeg, kend, vec, rbh);
        nt_load256(Bl, p.ldb, n0, p.N, kbeg, kend, vec, rbl);
    }
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        nt_store256(sAh, rah);
        nt_store256(sAl, ral);
        nt_store256(sBh, rbh);
        nt_store256(sBl, rbl);
        __syncthreads();
        if (k0 + BK < kend) {  // prefetch the next K step into registers
            nt_load256(Ah, p.lda, m0, p.M, k0 + BK, kend, vec, rah);
            nt_load256(Al, p.lda, m0, p.M, k0 + BK, kend, vec, ral);
            nt_load256(Bh, p.ldb, n0, p.N, k0 + BK, kend, vec, rbh);
            nt_load256(Bl, p.ldb, n0, p.N, k0 + BK, kend, vec, rbl);
        }
        v8bf bh[4], bl[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = wn * 64 + j * 16 + li;
            bh[j] = __builtin_bit_cast(v8bf, *(const v4u*)&sBh[col][8 * lk]);
            bl[j] = __builtin_bit_cast(v8bf, *(const v4u*)&sBl[col][8 * lk]);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row = wm * 128 + i * 16 + li;
            const v8bf ah = __builtin_bit_cast(v8bf, *(const v4u*)&sAh[row][8 * lk]);
            const v8bf al = __builtin_bit_cast(v8bf, *(const v4u*)&sAl[row][8 * lk]);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    if (SPLIT) {
        float* P = p.part + (int64_t)z * p.M * p.N;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = n0 + wn * 64 + j * 16 + li;
                if (n >= p.N) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * 128 + i * 16 + lk * 4 + r;
                    if (m < p.M) P[(int64_t)m * p.N + n] = acc[i][j][r];
                }
            }
    } else {
        TO* C = (TO*)p.C + (int64_t)b * p.bsC;
        const double alpha = p.alpha, beta = p.beta;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = n0 + wn * 64 + j * 16 + li;
                if (n >= p.N) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * 128 + i * 16 + lk * 4 + r;
                    if (m < p.M) {
                        TO* c = C + (int64_t)m * p.ldc + n;
                        double v = alpha * (double)acc[i][j][r];
                        if (beta != 0.0) v += beta * (double)(*c);
                        *c = (TO)v;
                    }
                }
            }
    }
}

// C[b][m][n] = alpha * sum_s part[b*splits + s][m][n] + beta * C[b][m][n]
template <typename TO>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const float* __restrict__ part, TO* __restrict__ C, int M,
                                                            int N, int splits, int batch, int64_t ldc, int64_t bsC,
                                                            double alpha, double beta, const int* active) {
    if (active && *active == 0) return;
    const int64_t MN = (int64_t)M * N;
    const int64_t total = MN * batch;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = e / MN, r = e - b * MN;
        const int64_t m = r / N, n = r - m * N;
        const float* q = part + b * splits * MN + r;
        float s = 0.f;
        for (int k = 0; k < splits; ++k) s += q[k * MN];
        TO* c = C + b * bsC + m * ldc + n;
        double v = alpha * (double)s;
        if (beta != 0.0) v += beta * (double)(*c);
        *c = (TO)v;
    }
}

}  // namespace

int split_bf16(Handle* h, int tX, const void* X, int rows, int K, int64_t s_row, int64_t s_k, int batch,
               int64_t bsX, uint16_t* hi, uint16_t* lo, int64_t ld_out, int64_t bs_out, const int* active) {
    if (rows <= 0 || K <= 0 || batch <= 0) return 0;
    dim3 grid((K + 31) / 32, (rows + 31) / 32, batch), block(256);
    if (tX == CT_F64)
        hipLaunchKernelGGL(split_kernel<double>, grid, block, 0, h->stream, (const double*)X, rows, K, s_row, s_k,
                           bsX, hi, lo, ld_out, bs_out, active);
    else
        hipLaunchKernelGGL(split_kernel<float>, grid, block, 0, h->stream, (const float*)X, rows, K, s_row, s_k,
                           bsX, hi, lo, ld_out, bs_out, active);
    CT_HIP(hipGetLastError());
    return 0;
}

int gemm_bf16x3_nt(Handle* h, const SplitOp& A, const SplitOp& B, int tC, void* C, int M, int N, int K,
                   int64_t ldc, int64_t bsC, int batch, double alpha, double beta, const int* active) {
    if (M <= 0 || N <= 0 || batch <= 0) return 0;
    NtArgs p;
    p.Ah = A.hi; p.Al = A.lo; p.Bh = B.hi; p.Bl = B.lo;
    p.C = C;
    p.M = M; p.N = N; p.K = K;
    p.tilesM = (M + BM - 1) / BM;
    p.lda = A.ld; p.ldb = B.ld; p.bsA = A.bs; p.bsB = B.bs;
    p.ldc = ldc; p.bsC = bsC;
    p.alpha = alpha; p.beta = beta; p.active = active;
    auto al16 = [](const void* q) { return ((uintptr_t)q % 16) == 0; };
    p.vec = (A.ld % 8 == 0) && (B.ld % 8 == 0) && (A.bs % 8 == 0) && (B.bs % 8 == 0) && al16(A.hi) &&
            al16(A.lo) && al16(B.hi) && al16(B.lo);
    p.tilesN = (N + BN - 1) / BN;
    p.GM = 8;
    p.splits = 1;
    p.kchunk = K;
    p.part = nullptr;

    // 256x256 tiling when it still gives at least one block per CU (after
    // splitting K where K is long), else the 128x128 tiling.
    const int64_t t256 = (int64_t)((M + BB - 1) / BB) * ((N + BB - 1) / BB) * batch;
    int splits = 1;
    while (t256 * splits < 256 && K / (splits * 2) >= 1024) splits *= 2;
    if (t256 * splits >= 256) {
        p.tilesM = (M + BB - 1) / BB;
        p.tilesN = (N + BB - 1) / BB;
        // Per-XCD share of one z-plane; group tile rows so that share spans
        // whole rows of tiles (each A panel then leaves HBM once).
        const int per_xcd = (p.tilesM * p.tilesN) / 8;
        p.GM = per_xcd >= p.tilesN ? per_xcd / p.tilesN : 1;
        if (p.GM < 1) p.GM = 1;
        p.splits = splits;
        p.kchunk = (int)(((int64_t)K / splits + BK - 1) / BK * BK);
        dim3 grid(p.tilesM * p.tilesN, 1, batch * splits), block(512);
        if (splits > 1) {
            void* w;
            CT_CHECK(ws_get(h, WS_SPLITK, (size_t)batch * splits * M * N * 4, &w));
            p.part = (float*)w;
            hipLaunchKernelGGL((gemm_bf16x3_nt256_kernel<float, true>), grid, block, 0, h->stream, p);
            const int64_t total = (int64_t)M * N * batch;
            const int g = (int)std::min<int64_t>((total + 255) / 256, 8192);
            if (tC == CT_F64)
                hipLaunchKernelGGL(splitk_reduce_kernel<double>, dim3(g), dim3(256), 0, h->stream,
                                   (const float*)p.part, (double*)C, M, N, splits, batch, ldc, bsC, alpha, beta,
                                   active);
            else
                hipLaunchKernelGGL(splitk_reduce_kernel<float>, dim3(g), dim3(256), 0, h->stream,
                                   (const float*)p.part, (float*)C, M, N, splits, batch, ldc, bsC, alpha, beta,
                                   active);
        } else if (tC == CT_F64) {
            hipLaunchKernelGGL((gemm_bf16x3_nt256_kernel<double, false>), grid, block, 0, h->stream, p);
        } else {
            hipLaunchKernelGGL((gemm_bf16x3_nt256_kernel<float, false>), grid, block, 0, h->stream, p);
        }
        CT_HIP(hipGetLastError());
        return 0;
    }
    dim3 grid(p.tilesM * p.tilesN, 1, batch), block(256);
    if (tC == CT_F64)
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<double>, grid, block, 0, h->stream, p);
    else
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<float>, grid, block, 0, h->stream, p);
    CT_HIP(hipGetLastError());
    return 0;
}

// Generic entry used by ct_tensordot / the contractions when no pre-split
// operand is available: split both operands into workspace, then run the NT
// kernel.  A(m,k) and B(k,n) follow GemmDesc.
int gemm_bf16x3(Handle* h, int tA, const void* A, int tB, const void* B, int tC, void* C, const GemmDesc& g) {
    const int64_t K8 = ((int64_t)g.K + 7) / 8 * 8;  // padded row length keeps 16-byte rows
    const int64_t bsA = (int64_t)g.M * K8, bsB = (int64_t)g.N * K8;
    const int64_t nA = bsA * g.batch, nB = bsB * g.batch;
    void* wa;
    void* wb;
    CT_CHECK(ws_get(h, WS_CONV_A, (size_t)nA * 4, &wa));
    CT_CHECK(ws_get(h, WS_CONV_B, (size_t)nB * 4, &wb));
    uint16_t* ah = (uint16_t*)wa;
    uint16_t* al = ah + nA;
    uint16_t* bh = (uint16_t*)wb;
    uint16_t* bl = bh + nB;
    // A(m, k): row = m, stride sAm; k stride sAk
    CT_CHECK(split_bf16(h, tA, A, g.M, g.K, g.sAm, g.sAk, g.batch, g.bsA, ah, al, K8, bsA, g.active));
    // B(k, n) -> row = n: stride sBn; k stride sBk
    CT_CHECK(split_bf16(h, tB, B, g.N, g.K, g.sBn, g.sBk, g.batch, g.bsB, bh, bl, K8, bsB, g.active));
    SplitOp sa, sb;
    sa.hi = ah; sa.lo = al; sa.ld = K8; sa.bs = bsA;
    sb.hi = bh; sb.lo = bl; sb.ld = K8; sb.bs = bsB;
    return gemm_bf16x3_nt(h, sa, sb, tC, C, g.M, g.N, g.K, g.ldc, g.bsC, g.batch, g.alpha, g.beta, g.active);
}

}  // namespace ct
