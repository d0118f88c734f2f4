// 3xBF16 split-precision GEMM on gfx950 MFMA (v_mfma_f32_16x16x32_bf16).
//
// Each FP32/FP64 operand x is split once into hi = bf16(x) and
// lo = bf16(x - hi).  The product a*b is then formed as
//   ah*bh + ah*bl + al*bh
// with every BF16xBF16 product exact in FP32 and accumulated in FP32; the
// dropped al*bl term and the rounding of lo bound the relative error of one
// product by about 2^-16.
//
// Operand format ("split rows"): row r of an operand holds, for every 32-wide
// block kb of the contraction index, 32 hi values followed by the 32 lo values
// of the same k (128 bytes: one cache line).  K is padded with zeros to a
// multiple of 32, so the kernels need no K-tail code, and an MFMA fragment
// (8 consecutive k of one row per lane) is one 16-byte read.  The kernels have
// a single "NT" form C = A * B^T over rows of A (m) and rows of B (n);
// callers holding m- or n-contiguous data split them with a transposing split.
//
// Two tilings:
//  * 256x256x32 blocks, 256 threads = 4 waves (2x2) of 128x128, accumulators
//    in AGPRs (one wave per SIMD).  Operand tiles travel HBM/L2 -> LDS by
//    direct-to-LDS loads (global_load_lds_dwordx4) into two 64 KiB stages: the
//    next stage is in flight while the MFMAs read the current one, with no
//    VGPR staging and no LDS write instructions.  In each K step all
//    fragment reads are issued between the first MFMAs, before the copy of
//    the next stage (see the main loop).  Optional split-K over blockIdx.z
//    with FP32 partial sums reduced by a second kernel; split-K launches give
//    each XCD one K range and a compact block of tiles.
//  * 128x128x32 blocks, 256 threads = 4 waves (2x2) of 64x64, LDS
//    double-buffered through registers, for problems too small to give every
//    CU a 256x256 tile.
#include <algorithm>

#include "ct_internal.h"
#include "ct_device.h"

namespace ct {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

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

// Three-way split of the FP32 value: hi + lo + l carries all of its 24
// significand bits (x - hi and x - hi - lo are exact in FP32).
template <typename T>
__device__ __forceinline__ void split_three(T v, uint16_t& hi, uint16_t& lo, uint16_t& l) {
    const float x = (float)v;
    hi = f32_to_bf16_rne(x);
    const float r = x - bf16_to_f32(hi);
    lo = f32_to_bf16_rne(r);
    l = f32_to_bf16_rne(r - bf16_to_f32(lo));
}

// ---------------------------------------------------------------- split
// Strided split into split rows.  One thread per (row, k) of a 32 (k) x 32
// (rows) block, 256 threads as 32 x 8; the block stages through LDS so both
// the strided read and the k-contiguous write are coalesced whichever of
// s_row / s_k is 1.  k in [K, K32) is written as zero.
template <typename T>
__global__ __launch_bounds__(256) void split_kernel(const T* __restrict__ X, int rows, int K, int64_t s_row,
                                                    int64_t s_k, int64_t bsX, uint16_t* __restrict__ out, int64_t ld,
                                                    int64_t bs, const int* active, uint16_t* __restrict__ outl,
                                                    int64_t ldl, int64_t bsl) {
    if (active && *active == 0) return;
    __shared__ float tile[32][33];
    const int b = blockIdx.z;
    const int r0 = blockIdx.y * 32, k0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    X += (int64_t)b * bsX;
    out += (int64_t)b * bs;
    if (outl) outl += (int64_t)b * bsl;
    if (s_row == 1) {
        // rows contiguous: read with tx along rows, write with tx along k
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = k0 + ty + 8 * i, r = r0 + tx;
            tile[ty + 8 * i][tx] = (k < K && r < rows) ? (float)X[(int64_t)r + (int64_t)k * s_k] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = r0 + ty + 8 * i;
            if (r < rows) {
                uint16_t h, l, l2;
                split_three(tile[tx][ty + 8 * i], h, l, l2);
                uint16_t* o = out + (int64_t)r * ld + 2 * k0 + tx;
                o[0] = h;
                o[32] = l;
                if (outl) outl[(int64_t)r * ldl + k0 + tx] = l2;
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = r0 + ty + 8 * i, k = k0 + tx;
            if (r < rows) {
                uint16_t h = 0, l = 0, l2 = 0;
                if (k < K) split_three(X[(int64_t)r * s_row + (int64_t)k * s_k], h, l, l2);
                uint16_t* o = out + (int64_t)r * ld + 2 * k0 + tx;
                o[0] = h;
                o[32] = l;
                if (outl) outl[(int64_t)r * ldl + k] = l2;
            }
        }
    }
}

// Transposing split (rows contiguous, s_row == 1) with 16-byte stores: a
// 32 (k) x 64 (rows) tile goes through LDS; each thread then splits 8
// consecutive k of one row and writes their hi and lo halves as one 16-byte
// store each (4 threads fill a row's 128-byte line).  Needs ld % 8 == 0 and a
// 16-byte aligned output base.  Element k of a row sits at k * s_k, or with
// TWO_LEVEL at (k / kin) * s_kc + (k % kin) * s_k.
template <typename T, bool TWO_LEVEL>
__global__ __launch_bounds__(256) void split_t_kernel(const T* __restrict__ X, int rows, int K, int kin, int64_t s_k,
                                                      int64_t s_kc, int64_t bsX, uint16_t* __restrict__ out,
                                                      int64_t ld, int64_t bs, const int* active,
                                                      uint16_t* __restrict__ outl, int64_t ldl, int64_t bsl) {
    if (active && *active == 0) return;
    __shared__ float tile[32][65];
    const int b = blockIdx.z;
    const int r0 = blockIdx.x * 64, k0 = blockIdx.y * 32;  // neighbouring blocks read neighbouring rows
    const int t = threadIdx.x;
    X += (int64_t)b * bsX;
    out += (int64_t)b * bs;
    if (outl) outl += (int64_t)b * bsl;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int kk = (t >> 6) + 4 * i, k = k0 + kk, r = r0 + (t & 63);
        const int64_t ko = TWO_LEVEL ? (int64_t)(k / kin) * s_kc + (int64_t)(k % kin) * s_k : (int64_t)k * s_k;
        tile[kk][t & 63] = (k < K && r < rows) ? (float)X[(int64_t)r + ko] : 0.f;
    }
    __syncthreads();
    const int row = t >> 2, grp = t & 3;
    const int r = r0 + row;
    if (r < rows) {
        uint32_t hp[4], lp[4], mp[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint16_t h0, l0, m0, h1, l1, m1;
            split_three(tile[grp * 8 + 2 * q][row], h0, l0, m0);
            split_three(tile[grp * 8 + 2 * q + 1][row], h1, l1, m1);
            hp[q] = (uint32_t)h0 | ((uint32_t)h1 << 16);
            lp[q] = (uint32_t)l0 | ((uint32_t)l1 << 16);
            mp[q] = (uint32_t)m0 | ((uint32_t)m1 << 16);
        }
        uint16_t* o = out + (int64_t)r * ld + 2 * k0 + grp * 8;
        *(v4u*)o = v4u{hp[0], hp[1], hp[2], hp[3]};
        *(v4u*)(o + 32) = v4u{lp[0], lp[1], lp[2], lp[3]};
        if (outl) *(v4u*)(outl + (int64_t)r * ldl + k0 + grp * 8) = v4u{mp[0], mp[1], mp[2], mp[3]};
    }
}

struct NtArgs {
    const uint16_t *A, *B;  // split rows
    void* C;
    int M, N, K32;          // K32: padded contraction length (multiple of 32)
    int tilesM, tilesN;
    int64_t lda, ldb, bsA, bsB, ldc, bsC;
    double alpha, beta;
    const int* active;
    int GM;                 // tile rows per L2 group (see gemm_tile_coords)
    int splits;             // split-K factor
    int batch_xcd;          // 256x256 kernel: 1 = whole batches per XCD (gemm_batch_per_xcd), 2 = one K range per XCD
    int kchunk;             // K range per split, multiple of 32
    float* part;            // [batch*splits][M][N] FP32 partial sums when splits > 1
};

template <typename TO>
__device__ __forceinline__ void store_out(const NtArgs& p, TO* C, int m, int n, float a) {
    if (m < p.M && n < p.N) {
        TO* c = C + (int64_t)m * p.ldc + n;
        double v = p.alpha * (double)a;
        if (p.beta != 0.0) v += p.beta * (double)(*c);
        *c = (TO)v;
    }
}

// ------------------------------------------------------- 256x256 tiling
constexpr int BB = 256, BK = 32;
constexpr int SROW = 128;             // bytes per row per stage: 32 hi + 32 lo
constexpr int SOPND = BB * SROW;      // 32 KiB per operand per stage
constexpr int STAGE = 2 * SOPND;      // 64 KiB

// SPLIT: write FP32 partial sums of this block's K range to p.part instead of
// alpha*acc + beta*C.
template <typename TO, bool SPLIT>
__global__ __launch_bounds__(256) void gemm_bf16x3_dma_kernel(NtArgs p) {
    if (p.active && *p.active == 0) return;
    __shared__ __attribute__((aligned(1024))) char smem[2 * STAGE];

    int tm, tn, b, sk, z;
    if (p.batch_xcd == 2) {
        // Split-K launches (grid = tiles x 1 x batch*splits, a multiple of 8
        // workgroups): workgroup f runs on XCD f mod 8, so give each XCD a
        // contiguous run of (z, tile) pairs.  The workgroups of one XCD then
        // share one K range and a compact block of tiles, and their A and B
        // panels are fetched into that XCD's L2 once.  A bijection.
        const int nt = gridDim.x, total = nt * gridDim.z;
        const int f = blockIdx.z * nt + blockIdx.x;
        const int v = (f & 7) * (total >> 3) + (f >> 3);
        z = v / nt;
        gemm_group_coords(v - z * nt, p.tilesM, p.tilesN, p.GM, tm, tn);
        b = z / p.splits;
        sk = z - b * p.splits;
    } else if (p.batch_xcd) {  // splits == 1
        int tile;
        gemm_batch_per_xcd(gridDim.x, b, tile);
        gemm_group_coords(tile, p.tilesM, p.tilesN, p.GM, tm, tn);
        sk = 0;
        z = b;
    } else {
        gemm_tile_coords(blockIdx.x, gridDim.x, p.tilesM, p.tilesN, p.GM, tm, tn);
        z = blockIdx.z;
        b = z / p.splits;
        sk = z - b * p.splits;
    }
    const int m0 = tm * BB, n0 = tn * BB;
    const int kbeg = sk * p.kchunk;
    const int kend = min(p.K32, kbeg + p.kchunk);

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int li = lane & 15, lk = lane >> 4;

    // Direct-to-LDS copy: waves 0/1 copy rows 0-127/128-255 of the A tile,
    // waves 2/3 those of the B tile; one instruction moves 8 rows x 128 B.
    // Rows past the matrix edge re-read the last row (their results are
    // never stored).  Inside a row, 16-byte chunk c sits at slot c ^ (r & 7),
    // so the fragment reads of 16 consecutive rows hit distinct banks.
    const bool isA = wave < 2;
    const uint16_t* P = isA ? p.A + (int64_t)b * p.bsA : p.B + (int64_t)b * p.bsB;
    const int64_t ldp = isA ? p.lda : p.ldb;
    const int rbase = (isA ? m0 : n0) + 128 * (wave & 1);
    const int rlast = (isA ? p.M : p.N) - 1;
    const int lrow = lane >> 3, lslot = lane & 7;
    const int64_t soff = (int64_t)(isA ? 0 : SOPND) + (wave & 1) * (128 * SROW);
    auto issue = [&](int stage, int k0) {
        char* base = smem + stage * STAGE + soff;
#pragma unroll
        for (int blk = 0; blk < 16; ++blk) {
            const int row = blk * 8 + lrow;  // 0..127 within this wave's half
            const int rr = (wave & 1) * 128 + row;
            const int chunk = lslot ^ (rr & 7);
            const int grow = min(rbase + row, rlast);
            const uint16_t* src = P + (int64_t)grow * ldp + 2 * k0 + chunk * 8;
            __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)(base + blk * 1024), 16, 0, 0);
        }
    };

    v4f acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    // One K step: every LDS fragment read of the step is issued first (the lo
    // B fragments between the first eight MFMAs, the A fragments of row
    // blocks 1-7 between the next sixteen), and only then the direct-to-LDS
    // copy of the next stage, between the MFMAs of row block 1.  While such a
    // copy is in flight the compiler can only wait for LDS reads with
    // lgkmcnt(0), so this order is what lets the fragment reads overlap the
    // MFMAs.  Every accumulator still sees l*bh, h*bl, h*bh in that order.
    if (kbeg < kend) issue(0, kbeg);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    int stage = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        const char* sa = smem + stage * STAGE;
        const char* sb = sa + SOPND;
        auto rdA = [&](int i, v8bf& h, v8bf& l) {
            const int r = wm * 128 + i * 16 + li;
            h = __builtin_bit_cast(v8bf, *(const v4u*)(sa + r * SROW + ((lk ^ (r & 7)) * 16)));
            l = __builtin_bit_cast(v8bf, *(const v4u*)(sa + r * SROW + (((lk + 4) ^ (r & 7)) * 16)));
        };
        auto rdB = [&](int j, int hl) -> v8bf {
            const int r = wn * 128 + j * 16 + li;
            return __builtin_bit_cast(v8bf, *(const v4u*)(sb + r * SROW + (((lk + 4 * hl) ^ (r & 7)) * 16)));
        };
        // the last step copies its own block again (in bounds, never read)
        const int knext = min(k0 + BK, kend - BK);
        v8bf aH[8], aL[8], bh[8], bl[8];
        rdA(0, aH[0], aL[0]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 8; ++j) bh[j] = rdB(j, 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aL[0], bh[j], acc[0][j], 0, 0, 0);
            bl[j] = rdB(j, 1);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 1; i < 8; ++i) rdA(i, aH[i], aL[i]);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH[0], bl[j], acc[0][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH[0], bh[j], acc[0][j], 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 14; ++q) {
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 1; i < 8; ++i) {
            if (i == 1) issue(stage ^ 1, knext);
            const v8bf h = aH[i], l = aL[i];
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l, bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h, bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h, bh[j], acc[i][j], 0, 0, 0);
            if (i == 1) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // one global read
                }
                __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // the next stage has landed in LDS (this wave's copies), then every wave's
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        stage ^= 1;
    }

    if (SPLIT) {
        float* Pp = p.part + (int64_t)z * p.M * p.N;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int n = n0 + wn * 128 + j * 16 + li;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int m = m0 + wm * 128 + i * 16 + lk * 4 + q;
                    if (m < p.M && n < p.N) Pp[(int64_t)m * p.N + n] = acc[i][j][q];
                }
            }
    } else {
        TO* C = (TO*)p.C + (int64_t)b * p.bsC;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int n = n0 + wn * 128 + j * 16 + li;
#pragma unroll
                for (int q = 0; q < 4; ++q) store_out(p, C, m0 + wm * 128 + i * 16 + lk * 4 + q, n, acc[i][j][q]);
            }
    }
}

// ------------------------------------------------- 6xBF16, 256x128 tiling
// FP32-accurate products on the BF16 matrix cores.  With x = h + m + l
// (h, m: the split rows' hi and lo; l: the l plane), the six products
//   l*H + h*L + m*M + m*H + h*M + h*H        (smallest terms first)
// leave out only m*L + l*M + l*L, below 2^-24 |a b| (the bf16x6 scheme of
// Henry, Tang and Heinecke, "Leveraging the bfloat16 artificial intelligence
// datatype for higher-precision computations", ARITH 2019), so with FP32
// accumulation the result carries FP32 GEMM accuracy.
// 256 (m) x 128 (n) x 32 (k) blocks, 256 threads = 4 waves (2 x 2) of 128 x 64
// with AGPR accumulators (h*H and the five smaller products in separate
// accumulators, summed once at the end).  Per stage the split rows (128 B per row) and the l
// planes (64 B per row) of both tiles travel by direct-to-LDS loads: 72 KiB,
// two stages.  Inside an l-plane row, 16-byte chunk c sits at slot
// c ^ lswz(r), which spreads the fragment reads of 16 consecutive rows over
// distinct banks.
constexpr int X6M = 256, X6N = 128;
constexpr int X6_AS = 0;                    // A split rows: 256 x 128 B
constexpr int X6_BS = X6_AS + X6M * SROW;   // B split rows: 128 x 128 B
constexpr int X6_AL = X6_BS + X6N * SROW;   // A l plane: 256 x 64 B
constexpr int X6_BL = X6_AL + X6M * 64;     // B l plane: 128 x 64 B
constexpr int X6_STAGE = X6_BL + X6N * 64;  // 72 KiB

struct X6Args {
    NtArgs nt;  // split rows, output and tiling (tilesM, tilesN count 256 x 128 or 128 x 128 tiles)
    const uint16_t *Al, *Bl;
    int64_t ldal, ldbl, bsAl, bsBl;
};

__device__ __forceinline__ int lswz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

template <typename TO, bool SPLIT>
__global__ __launch_bounds__(256) void gemm_bf16x6_dma_kernel(X6Args q) {
    const NtArgs& p = q.nt;
    if (p.active && *p.active == 0) return;
    __shared__ __attribute__((aligned(1024))) char smem[2 * X6_STAGE];

    int tm, tn, b, sk, z;
    if (p.batch_xcd == 2) {  // split K: one K range and a compact block of tiles per XCD (see the 3xBF16 kernel)
        const int nt = gridDim.x, total = nt * gridDim.z;
        const int f = blockIdx.z * nt + blockIdx.x;
        const int v = (f & 7) * (total >> 3) + (f >> 3);
        z = v / nt;
        gemm_group_coords(v - z * nt, p.tilesM, p.tilesN, p.GM, tm, tn);
        b = z / p.splits;
        sk = z - b * p.splits;
    } else if (p.batch_xcd) {  // splits == 1, whole batches per XCD
        int tile;
        gemm_batch_per_xcd(gridDim.x, b, tile);
        gemm_group_coords(tile, p.tilesM, p.tilesN, p.GM, tm, tn);
        sk = 0;
        z = b;
    } else {
        gemm_tile_coords(blockIdx.x, gridDim.x, p.tilesM, p.tilesN, p.GM, tm, tn);
        z = blockIdx.z;
        b = z / p.splits;
        sk = z - b * p.splits;
    }
    const int m0 = tm * X6M, n0 = tn * X6N;
    const int kbeg = sk * p.kchunk;
    const int kend = min(p.K32, kbeg + p.kchunk);

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int li = lane & 15, lk = lane >> 4;

    const uint16_t* PA = p.A + (int64_t)b * p.bsA;
    const uint16_t* PB = p.B + (int64_t)b * p.bsB;
    const uint16_t* LA = q.Al + (int64_t)b * q.bsAl;
    const uint16_t* LB = q.Bl + (int64_t)b * q.bsBl;
    // Copies of wave w: split-row rows [64w, 64w+64) of A and [32w, 32w+32) of
    // B (one instruction = 8 rows x 128 B), and the same rows of the l planes
    // (one instruction = 16 rows x 64 B): 18 instructions per wave and stage.
    // Rows past the matrix edge re-read the last row (never stored).
    const int r8 = lane >> 3, s8 = lane & 7, r4 = lane >> 2, s4 = lane & 3;
    auto issue = [&](int stage, int k0) {
        char* base = smem + stage * X6_STAGE;
#pragma unroll
        for (int blk = 0; blk < 8; ++blk) {
            const int row = 64 * wave + 8 * blk + r8;
            const int grow = min(m0 + row, p.M - 1);
            const uint16_t* src = PA + (int64_t)grow * p.lda + 2 * k0 + (s8 ^ (row & 7)) * 8;
            __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)(base + X6_AS + (64 * wave + 8 * blk) * SROW),
                                             16, 0, 0);
        }
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
            const int row = 32 * wave + 8 * blk + r8;
            const int grow = min(n0 + row, p.N - 1);
            const uint16_t* src = PB + (int64_t)grow * p.ldb + 2 * k0 + (s8 ^ (row & 7)) * 8;
            __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)(base + X6_BS + (32 * wave + 8 * blk) * SROW),
                                             16, 0, 0);
        }
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
            const int row = 64 * wave + 16 * blk + r4;
            const int grow = min(m0 + row, p.M - 1);
            const uint16_t* src = LA + (int64_t)grow * q.ldal + k0 + (s4 ^ lswz(row)) * 8;
            __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)(base + X6_AL + (64 * wave + 16 * blk) * 64),
                                             16, 0, 0);
        }
#pragma unroll
        for (int blk = 0; blk < 2; ++blk) {
            const int row = 32 * wave + 16 * blk + r4;
            const int grow = min(n0 + row, p.N - 1);
            const uint16_t* src = LB + (int64_t)grow * q.ldbl + k0 + (s4 ^ lswz(row)) * 8;
            __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)(base + X6_BL + (32 * wave + 16 * blk) * 64),
                                             16, 0, 0);
        }
    };

    // acc: the h*H products; acs: the five smaller ones, so that the large
    // running sum is rounded once per K step instead of six times
    v4f acc[8][4], acs[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
            acs[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
        }

    // One K step: the B fragments and row block 0 of A first, the A fragments
    // of row blocks 1-7 between the MFMAs of row block 0, and only then (with
    // every LDS read of the step issued) the direct-to-LDS copy of the next
    // stage, between the MFMAs of row block 1.
    if (kbeg < kend) issue(0, kbeg);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    int stage = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        const char* st = smem + stage * X6_STAGE;
        auto rdA = [&](int i, v8bf& h, v8bf& m, v8bf& l) {
            const int r = wm * 128 + i * 16 + li;
            h = __builtin_bit_cast(v8bf, *(const v4u*)(st + X6_AS + r * SROW + ((lk ^ (r & 7)) * 16)));
            m = __builtin_bit_cast(v8bf, *(const v4u*)(st + X6_AS + r * SROW + (((lk + 4) ^ (r & 7)) * 16)));
            l = __builtin_bit_cast(v8bf, *(const v4u*)(st + X6_AL + r * 64 + ((lk ^ lswz(r)) * 16)));
        };
        auto rdB = [&](int j, v8bf& h, v8bf& m, v8bf& l) {
            const int r = wn * 64 + j * 16 + li;
            h = __builtin_bit_cast(v8bf, *(const v4u*)(st + X6_BS + r * SROW + ((lk ^ (r & 7)) * 16)));
            m = __builtin_bit_cast(v8bf, *(const v4u*)(st + X6_BS + r * SROW + (((lk + 4) ^ (r & 7)) * 16)));
            l = __builtin_bit_cast(v8bf, *(const v4u*)(st + X6_BL + r * 64 + ((lk ^ lswz(r)) * 16)));
        };
        // the last step copies its own block again (in bounds, never read)
        const int knext = min(k0 + BK, kend - BK);
        v8bf aH[8], aM[8], aL[8], bH[4], bM[4], bL[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) rdB(j, bH[j], bM[j], bL[j]);
        rdA(0, aH[0], aM[0], aL[0]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acs[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aL[0], bH[j], acs[0][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acs[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH[0], bL[j], acs[0][j], 0, 0, 0);
#pragma unroll
        for (int i = 1; i < 8; ++i) rdA(i, aH[i], aM[i], aL[i]);
#pragma unroll
        for (int j = 0; j < 4; ++j) acs[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aM[0], bM[j], acs[0][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acs[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aM[0], bH[j], acs[0][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acs[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH[0], bM[j], acs[0][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH[0], bH[j], acc[0][j], 0, 0, 0);
        // 24 MFMAs of row block 0 with the 21 fragment reads of row blocks 1-7
#pragma unroll
        for (int qq = 0; qq < 21; ++qq) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 1; i < 8; ++i) {
            if (i == 1) issue(stage ^ 1, knext);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aL[i], bH[j], acs[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH[i], bL[j], acs[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aM[i], bM[j], acs[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aM[i], bH[j], acs[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH[i], bM[j], acs[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aH[i], bH[j], acc[i][j], 0, 0, 0);
            if (i == 1) {
                // the 18 copy instructions between the first 18 MFMAs of row block 1
#pragma unroll
                for (int qq = 0; qq < 18; ++qq) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // one global read
                }
                __builtin_amdgcn_sched_group_barrier(0x008, 6, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // the next stage has landed in LDS (this wave's copies), then every wave's
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        stage ^= 1;
    }

#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += acs[i][j];
    if (SPLIT) {
        float* Pp = p.part + (int64_t)z * p.M * p.N;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = n0 + wn * 64 + j * 16 + li;
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) {
                    const int m = m0 + wm * 128 + i * 16 + lk * 4 + qq;
                    if (m < p.M && n < p.N) Pp[(int64_t)m * p.N + n] = acc[i][j][qq];
                }
            }
    } else {
        TO* C = (TO*)p.C + (int64_t)b * p.bsC;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = n0 + wn * 64 + j * 16 + li;
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) store_out(p, C, m0 + wm * 128 + i * 16 + lk * 4 + qq, n, acc[i][j][qq]);
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

// The same reduction, four consecutive n per thread (16-byte partial loads,
// 16-byte stores): needs N % 4 == 0 and C rows 16-byte aligned (the host
// checks).  Each element sums its partials in the same order as above.
template <typename TO>
__global__ __launch_bounds__(256) void splitk_reduce4_kernel(const float* __restrict__ part, TO* __restrict__ C, int M,
                                                             int N, int splits, int batch, int64_t ldc, int64_t bsC,
                                                             double alpha, double beta, const int* active) {
    if (active && *active == 0) return;
    const int64_t MN = (int64_t)M * N;
    const int64_t total4 = MN * batch / 4;
    for (int64_t e4 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e4 < total4;
         e4 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = 4 * e4;
        const int64_t b = e / MN, r = e - b * MN;
        const int64_t m = r / N, n = r - m * N;
        const float* q = part + b * splits * MN + r;
        v4f s = v4f{0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < splits; ++k) s += *(const v4f*)(q + k * MN);
        TO* c = C + b * bsC + m * ldc + n;
        double v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = alpha * (double)s[t];
        if (beta != 0.0) {
#pragma unroll
            for (int t = 0; t < 4; ++t) v[t] += beta * (double)c[t];
        }
        if constexpr (sizeof(TO) == 8) {
            typedef double v2d __attribute__((ext_vector_type(2)));
            *(v2d*)c = v2d{v[0], v[1]};
            *(v2d*)(c + 2) = v2d{v[2], v[3]};
        } else {
            *(v4f*)c = v4f{(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
        }
    }
}

// C = alpha * sum of the split partials + beta * C, vectorized when the
// shapes and alignment allow.
template <typename TO>
void splitk_reduce(hipStream_t st, const float* part, TO* C, int M, int N, int splits, int batch, int64_t ldc,
                   int64_t bsC, double alpha, double beta, const int* active) {
    const int64_t total = (int64_t)M * N * batch;
    const int64_t vecw = 16 / (int64_t)sizeof(TO);  // elements per 16 bytes
    if (N % 4 == 0 && ldc % vecw == 0 && bsC % vecw == 0 && (uintptr_t)C % 16 == 0) {
        const int g = (int)std::min<int64_t>((total / 4 + 255) / 256, 8192);
        hipLaunchKernelGGL(splitk_reduce4_kernel<TO>, dim3(g), dim3(256), 0, st, part, C, M, N, splits, batch, ldc,
                           bsC, alpha, beta, active);
    } else {
        const int g = (int)std::min<int64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL(splitk_reduce_kernel<TO>, dim3(g), dim3(256), 0, st, part, C, M, N, splits, batch, ldc, bsC,
                           alpha, beta, active);
    }
}

// ------------------------------------------------------- 128x128 tiling
constexpr int SM = 128, LDK = BK + 8;  // LDS row = 40 bf16 = 80 B (conflict-free padding)

// Thread t copies the 16 hi and the 16 lo values of k-half t / 128 of row
// t % 128 (four 16-byte loads); rows past the edge re-read the last row.
__device__ __forceinline__ void nt_load(const uint16_t* P, int64_t ld, int R0, int RN, int k0, v4u* r) {
    const int row = min(R0 + (int)(threadIdx.x & 127), RN - 1), kh = threadIdx.x >> 7;
    const v4u* q = (const v4u*)(P + (int64_t)row * ld + 2 * k0 + 16 * kh);
    r[0] = q[0];
    r[1] = q[1];
    r[2] = q[4];  // lo values sit 32 elements (64 B) after the hi ones
    r[3] = q[5];
}

__device__ __forceinline__ void nt_store(uint16_t (*Sh)[LDK], uint16_t (*Sl)[LDK], const v4u* r) {
    const int row = threadIdx.x & 127, kh = threadIdx.x >> 7;
    v4u* h = (v4u*)&Sh[row][16 * kh];
    v4u* l = (v4u*)&Sl[row][16 * kh];
    h[0] = r[0];
    h[1] = r[1];
    l[0] = r[2];
    l[1] = r[3];
}

template <typename TO>
__global__ __launch_bounds__(256) void gemm_bf16x3_nt_kernel(NtArgs p) {
    if (p.active && *p.active == 0) return;
    __shared__ __attribute__((aligned(16))) uint16_t sAh[2][SM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[2][SM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[2][SM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[2][SM][LDK];

    int tm, tn;
    gemm_tile_coords(blockIdx.x, gridDim.x, p.tilesM, p.tilesN, tm, tn);
    const int m0 = tm * SM, n0 = tn * SM;
    const int zz = blockIdx.z;  // batch * splits + split
    const int z = zz / p.splits;
    const int kbeg = (zz - z * p.splits) * p.kchunk;
    const int kend = min(p.K32, kbeg + p.kchunk);
    const uint16_t* A = p.A + (int64_t)z * p.bsA;
    const uint16_t* B = p.B + (int64_t)z * p.bsB;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int li = lane & 15, lk = lane >> 4;

    v4f acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    v4u ra[4], rb[4];
    nt_load(A, p.lda, m0, p.M, kbeg, ra);
    nt_load(B, p.ldb, n0, p.N, kbeg, rb);
    nt_store(sAh[0], sAl[0], ra);
    nt_store(sBh[0], sBl[0], rb);
    __syncthreads();

    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        const bool more = k0 + BK < kend;
        if (more) {
            nt_load(A, p.lda, m0, p.M, k0 + BK, ra);
            nt_load(B, p.ldb, n0, p.N, k0 + BK, rb);
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
            nt_store(sAh[buf ^ 1], sAl[buf ^ 1], ra);
            nt_store(sBh[buf ^ 1], sBl[buf ^ 1], rb);
        }
        __syncthreads();
        buf ^= 1;
    }

    if (p.splits > 1) {  // FP32 partial sums of this K range, reduced by splitk_reduce_kernel
        float* P = p.part + (int64_t)zz * p.M * p.N;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = n0 + wn * 64 + j * 16 + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * 64 + i * 16 + lk * 4 + r;
                    if (m < p.M && n < p.N) P[(int64_t)m * p.N + n] = acc[i][j][r];
                }
            }
        return;
    }
    TO* C = (TO*)p.C + (int64_t)z * p.bsC;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + wn * 64 + j * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) store_out(p, C, m0 + wm * 64 + i * 16 + lk * 4 + r, n, acc[i][j][r]);
        }
}

// 6xBF16 on 128x128x32 blocks (problems too small to give every CU a 256x128
// tile): 4 waves (2 x 2) of 64 x 64, operands staged through registers into a
// double-buffered LDS (h, m and l planes of both tiles, 120 KiB), h*H and the
// five smaller products in separate accumulators as in the 256x128 kernel.
// Thread t copies the 16 h, 16 m and 16 l values of k-half t / 128 of row
// t % 128 (six 16-byte loads); rows past the edge re-read the last row.
__device__ __forceinline__ void nt6_load(const uint16_t* P, int64_t ld, const uint16_t* Pl, int64_t ldl, int R0,
                                         int RN, int k0, v4u* r) {
    const int row = min(R0 + (int)(threadIdx.x & 127), RN - 1), kh = threadIdx.x >> 7;
    const v4u* q = (const v4u*)(P + (int64_t)row * ld + 2 * k0 + 16 * kh);
    const v4u* ql = (const v4u*)(Pl + (int64_t)row * ldl + k0 + 16 * kh);
    r[0] = q[0];
    r[1] = q[1];
    r[2] = q[4];  // lo values sit 32 elements (64 B) after the hi ones
    r[3] = q[5];
    r[4] = ql[0];
    r[5] = ql[1];
}

__device__ __forceinline__ void nt6_store(uint16_t (*Sh)[LDK], uint16_t (*Sm)[LDK], uint16_t (*Sl)[LDK],
                                          const v4u* r) {
    const int row = threadIdx.x & 127, kh = threadIdx.x >> 7;
    v4u* h = (v4u*)&Sh[row][16 * kh];
    v4u* m = (v4u*)&Sm[row][16 * kh];
    v4u* l = (v4u*)&Sl[row][16 * kh];
    h[0] = r[0];
    h[1] = r[1];
    m[0] = r[2];
    m[1] = r[3];
    l[0] = r[4];
    l[1] = r[5];
}

template <typename TO>
__global__ __launch_bounds__(256) void gemm_bf16x6_nt_kernel(X6Args q) {
    const NtArgs& p = q.nt;
    if (p.active && *p.active == 0) return;
    __shared__ __attribute__((aligned(16))) uint16_t sAh[2][SM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sAm[2][SM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sAl[2][SM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBh[2][SM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBm[2][SM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t sBl[2][SM][LDK];

    int tm, tn;
    gemm_tile_coords(blockIdx.x, gridDim.x, p.tilesM, p.tilesN, tm, tn);
    const int m0 = tm * SM, n0 = tn * SM;
    const int zz = blockIdx.z;  // batch * splits + split
    const int z = zz / p.splits;
    const int kbeg = (zz - z * p.splits) * p.kchunk;
    const int kend = min(p.K32, kbeg + p.kchunk);
    const uint16_t* A = p.A + (int64_t)z * p.bsA;
    const uint16_t* B = p.B + (int64_t)z * p.bsB;
    const uint16_t* Al = q.Al + (int64_t)z * q.bsAl;
    const uint16_t* Bl = q.Bl + (int64_t)z * q.bsBl;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int li = lane & 15, lk = lane >> 4;

    v4f acc[4][4], acs[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
            acs[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
        }

    v4u ra[6], rb[6];
    if (kbeg < kend) {
        nt6_load(A, p.lda, Al, q.ldal, m0, p.M, kbeg, ra);
        nt6_load(B, p.ldb, Bl, q.ldbl, n0, p.N, kbeg, rb);
        nt6_store(sAh[0], sAm[0], sAl[0], ra);
        nt6_store(sBh[0], sBm[0], sBl[0], rb);
    }
    __syncthreads();

    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        const bool more = k0 + BK < kend;
        if (more) {
            nt6_load(A, p.lda, Al, q.ldal, m0, p.M, k0 + BK, ra);
            nt6_load(B, p.ldb, Bl, q.ldbl, n0, p.N, k0 + BK, rb);
        }
        v8bf ah[4], am[4], al[4], bh[4], bm[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = wm * 64 + i * 16 + li;
            ah[i] = __builtin_bit_cast(v8bf, *(const v4u*)&sAh[buf][row][8 * lk]);
            am[i] = __builtin_bit_cast(v8bf, *(const v4u*)&sAm[buf][row][8 * lk]);
            al[i] = __builtin_bit_cast(v8bf, *(const v4u*)&sAl[buf][row][8 * lk]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = wn * 64 + j * 16 + li;
            bh[j] = __builtin_bit_cast(v8bf, *(const v4u*)&sBh[buf][col][8 * lk]);
            bm[j] = __builtin_bit_cast(v8bf, *(const v4u*)&sBm[buf][col][8 * lk]);
            bl[j] = __builtin_bit_cast(v8bf, *(const v4u*)&sBl[buf][col][8 * lk]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acs[i][j], 0, 0, 0);
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acs[i][j], 0, 0, 0);
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[i], bm[j], acs[i][j], 0, 0, 0);
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[i], bh[j], acs[i][j], 0, 0, 0);
                acs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bm[j], acs[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        if (more) {
            nt6_store(sAh[buf ^ 1], sAm[buf ^ 1], sAl[buf ^ 1], ra);
            nt6_store(sBh[buf ^ 1], sBm[buf ^ 1], sBl[buf ^ 1], rb);
        }
        __syncthreads();
        buf ^= 1;
    }

#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += acs[i][j];
    if (p.splits > 1) {  // FP32 partial sums of this K range, reduced by splitk_reduce
        float* P = p.part + (int64_t)zz * p.M * p.N;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = n0 + wn * 64 + j * 16 + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * 64 + i * 16 + lk * 4 + r;
                    if (m < p.M && n < p.N) P[(int64_t)m * p.N + n] = acc[i][j][r];
                }
            }
        return;
    }
    TO* C = (TO*)p.C + (int64_t)z * p.bsC;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + wn * 64 + j * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) store_out(p, C, m0 + wm * 64 + i * 16 + lk * 4 + r, n, acc[i][j][r]);
        }
}

}  // namespace

int split_bf16(Handle* h, int tX, const void* X, int rows, int K, int64_t s_row, int64_t s_k, int batch,
               int64_t bsX, uint16_t* out, int64_t ld_out, int64_t bs_out, const int* active, uint16_t* outl,
               int64_t ldl, int64_t bsl) {
    if (rows <= 0 || K <= 0 || batch <= 0) return 0;
    if (outl && ldl < split_k32(K)) {
        set_last_error("split_bf16: the l plane needs ldl >= roundup(K, 32)");
        return CT_STATUS_SHAPE;
    }
    const bool l_vec = !outl || (ldl % 8 == 0 && bsl % 8 == 0 && ((uintptr_t)outl % 16) == 0);
    if (s_row == 1 && ld_out % 8 == 0 && bs_out % 8 == 0 && ((uintptr_t)out % 16) == 0 && l_vec) {
        dim3 grid((unsigned)((rows + 63) / 64), (unsigned)(split_k32(K) / 32), (unsigned)batch), block(256);
        if (tX == CT_F64)
            hipLaunchKernelGGL((split_t_kernel<double, false>), grid, block, 0, h->stream, (const double*)X, rows, K,
                               K, s_k, (int64_t)0, bsX, out, ld_out, bs_out, active, outl, ldl, bsl);
        else
            hipLaunchKernelGGL((split_t_kernel<float, false>), grid, block, 0, h->stream, (const float*)X, rows, K, K,
                               s_k, (int64_t)0, bsX, out, ld_out, bs_out, active, outl, ldl, bsl);
        CT_HIP(hipGetLastError());
        return 0;
    }
    dim3 grid((unsigned)(split_k32(K) / 32), (unsigned)((rows + 31) / 32), (unsigned)batch), block(256);
    if (tX == CT_F64)
        hipLaunchKernelGGL(split_kernel<double>, grid, block, 0, h->stream, (const double*)X, rows, K, s_row, s_k,
                           bsX, out, ld_out, bs_out, active, outl, ldl, bsl);
    else
        hipLaunchKernelGGL(split_kernel<float>, grid, block, 0, h->stream, (const float*)X, rows, K, s_row, s_k,
                           bsX, out, ld_out, bs_out, active, outl, ldl, bsl);
    CT_HIP(hipGetLastError());
    return 0;
}

int split_bf16_kk(Handle* h, int tX, const void* X, int rows, int K, int kin, int64_t s_k, int64_t s_kc, int batch,
                  int64_t bsX, uint16_t* out, int64_t ld_out, int64_t bs_out, const int* active, uint16_t* outl,
                  int64_t ldl, int64_t bsl) {
    if (rows <= 0 || K <= 0 || batch <= 0) return 0;
    if (kin < 1 || ld_out % 8 || bs_out % 8 || ((uintptr_t)out % 16) || ld_out < 2 * split_k32(K) ||
        (outl && (ldl % 8 || bsl % 8 || ((uintptr_t)outl % 16) || ldl < split_k32(K)))) {
        set_last_error("split_bf16_kk: needs kin >= 1, ld >= 2*roundup(K,32) (l plane: ldl >= roundup(K,32)), "
                       "strides multiples of 8, 16-byte aligned outputs");
        return CT_STATUS_SHAPE;
    }
    dim3 grid((unsigned)((rows + 63) / 64), (unsigned)(split_k32(K) / 32), (unsigned)batch), block(256);
    if (tX == CT_F64)
        hipLaunchKernelGGL((split_t_kernel<double, true>), grid, block, 0, h->stream, (const double*)X, rows, K, kin,
                           s_k, s_kc, bsX, out, ld_out, bs_out, active, outl, ldl, bsl);
    else
        hipLaunchKernelGGL((split_t_kernel<float, true>), grid, block, 0, h->stream, (const float*)X, rows, K, kin,
                           s_k, s_kc, bsX, out, ld_out, bs_out, active, outl, ldl, bsl);
    CT_HIP(hipGetLastError());
    return 0;
}

int gemm_bf16x3_nt(Handle* h, const SplitOp& A, const SplitOp& B, int tC, void* C, int M, int N, int K,
                   int64_t ldc, int64_t bsC, int batch, double alpha, double beta, const int* active) {
    if (M <= 0 || N <= 0 || batch <= 0) return 0;
    const int K32 = (int)split_k32(K);
    auto al16 = [](const void* q) { return ((uintptr_t)q % 16) == 0; };
    if (A.ld < 2 * K32 || B.ld < 2 * K32 || A.ld % 8 || B.ld % 8 || A.bs % 8 || B.bs % 8 || !al16(A.p) ||
        !al16(B.p)) {
        set_last_error("3xBF16 GEMM: split rows need ld >= 2*roundup(K,32), ld and batch stride multiples of 8, "
                       "16-byte aligned bases");
        return CT_STATUS_SHAPE;
    }
    NtArgs p;
    p.A = A.p; p.B = B.p;
    p.C = C;
    p.M = M; p.N = N; p.K32 = K32;
    p.lda = A.ld; p.ldb = B.ld; p.bsA = A.bs; p.bsB = B.bs;
    p.ldc = ldc; p.bsC = bsC;
    p.alpha = alpha; p.beta = beta; p.active = active;
    p.GM = 8;
    p.splits = 1;
    p.batch_xcd = 0;
    p.kchunk = K32;
    p.part = nullptr;
    if (K == 0) {  // empty contraction: C = beta * C
        p.K32 = 0;
    }

    // 256x256 tiling when it still gives every CU a block (after splitting
    // long K ranges), else the 128x128 tiling.
    const int64_t t256 = (int64_t)((M + BB - 1) / BB) * ((N + BB - 1) / BB) * batch;
    int splits = 1;
    while (t256 * splits < 256 && K32 / (splits * 2) >= 1024) splits *= 2;
    if (t256 * splits >= 256) {
        p.tilesM = (M + BB - 1) / BB;
        p.tilesN = (N + BB - 1) / BB;
        // Per-XCD share of one z-plane; group tile rows so that share spans
        // whole rows of tiles (each A panel then leaves HBM once).
        const int per_xcd = (p.tilesM * p.tilesN) / 8;
        p.GM = per_xcd >= p.tilesN ? per_xcd / p.tilesN : 1;
        if (p.GM < 1) p.GM = 1;
        p.splits = splits;
        p.kchunk = (K32 / splits + BK - 1) / BK * BK;
        const int64_t total = (int64_t)p.tilesM * p.tilesN * batch * splits;
        if (splits > 1 && total % 8 == 0) {
            // one K range per XCD (gemm coords mode 2): the XCD's share of one
            // z-plane, grouped into whole rows of tiles
            const int64_t share = std::min<int64_t>((int64_t)p.tilesM * p.tilesN, total / 8);
            p.batch_xcd = 2;
            p.GM = share >= p.tilesN ? (int)(share / p.tilesN) : 1;
        }
        // few tiles per batch: keep each batch on one XCD so its panels are fetched once
        if (splits == 1 && batch % 8 == 0 && p.tilesM * p.tilesN <= 32) {
            p.batch_xcd = 1;
            p.GM = p.tilesM;
        }
        dim3 grid(p.tilesM * p.tilesN, 1, batch * splits), block(256);
        if (splits > 1) {
            void* w;
            CT_CHECK(ws_get(h, WS_SPLITK, (size_t)batch * splits * M * N * 4, &w));
            p.part = (float*)w;
            hipLaunchKernelGGL((gemm_bf16x3_dma_kernel<float, true>), grid, block, 0, h->stream, p);
            if (tC == CT_F64)
                splitk_reduce<double>(h->stream, p.part, (double*)C, M, N, splits, batch, ldc, bsC, alpha, beta,
                                      active);
            else
                splitk_reduce<float>(h->stream, p.part, (float*)C, M, N, splits, batch, ldc, bsC, alpha, beta,
                                     active);
        } else if (tC == CT_F64) {
            hipLaunchKernelGGL((gemm_bf16x3_dma_kernel<double, false>), grid, block, 0, h->stream, p);
        } else {
            hipLaunchKernelGGL((gemm_bf16x3_dma_kernel<float, false>), grid, block, 0, h->stream, p);
        }
        CT_HIP(hipGetLastError());
        return 0;
    }
    p.tilesM = (M + SM - 1) / SM;
    p.tilesN = (N + SM - 1) / SM;
    // split K when the 128x128 tiles leave most CUs idle (at least 8 K blocks per split)
    const int64_t t128 = (int64_t)p.tilesM * p.tilesN * batch;
    int s128 = 1;
    if (t128 < 128 && K32 >= 16 * BK) s128 = (int)std::min<int64_t>((256 + t128 - 1) / t128, K32 / (8 * BK));
    if (s128 > 1) {
        p.kchunk = ((K32 + s128 - 1) / s128 + BK - 1) / BK * BK;
        s128 = (K32 + p.kchunk - 1) / p.kchunk;
    }
    p.splits = s128;
    dim3 grid(p.tilesM * p.tilesN, 1, batch * s128), block(256);
    if (s128 > 1) {
        void* w;
        CT_CHECK(ws_get(h, WS_SPLITK, (size_t)batch * s128 * M * N * 4, &w));
        p.part = (float*)w;
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<float>, grid, block, 0, h->stream, p);
        if (tC == CT_F64)
            splitk_reduce<double>(h->stream, p.part, (double*)C, M, N, s128, batch, ldc, bsC, alpha, beta, active);
        else
            splitk_reduce<float>(h->stream, p.part, (float*)C, M, N, s128, batch, ldc, bsC, alpha, beta, active);
    } else if (tC == CT_F64) {
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<double>, grid, block, 0, h->stream, p);
    } else {
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<float>, grid, block, 0, h->stream, p);
    }
    CT_HIP(hipGetLastError());
    return 0;
}

int gemm_bf16x6_nt(Handle* h, const SplitOp& A, const SplitOp& B, int tC, void* C, int M, int N, int K,
                   int64_t ldc, int64_t bsC, int batch, double alpha, double beta, const int* active) {
    if (M <= 0 || N <= 0 || batch <= 0) return 0;
    const int K32 = (int)split_k32(K);
    auto al16 = [](const void* q) { return ((uintptr_t)q % 16) == 0; };
    if (!A.l || !B.l || A.ld < 2 * K32 || B.ld < 2 * K32 || A.ld % 8 || B.ld % 8 || A.bs % 8 || B.bs % 8 ||
        !al16(A.p) || !al16(B.p) || A.ldl < K32 || B.ldl < K32 || A.ldl % 8 || B.ldl % 8 || A.bsl % 8 ||
        B.bsl % 8 || !al16(A.l) || !al16(B.l)) {
        set_last_error("6xBF16 GEMM: split rows need ld >= 2*roundup(K,32), l planes ldl >= roundup(K,32), "
                       "strides multiples of 8 and 16-byte aligned bases");
        return CT_STATUS_SHAPE;
    }
    X6Args q;
    NtArgs& p = q.nt;
    p.A = A.p; p.B = B.p;
    p.C = C;
    p.M = M; p.N = N; p.K32 = K32;
    p.lda = A.ld; p.ldb = B.ld; p.bsA = A.bs; p.bsB = B.bs;
    p.ldc = ldc; p.bsC = bsC;
    p.alpha = alpha; p.beta = beta; p.active = active;
    p.batch_xcd = 0;
    p.part = nullptr;
    q.Al = A.l; q.Bl = B.l;
    q.ldal = A.ldl; q.ldbl = B.ldl; q.bsAl = A.bsl; q.bsBl = B.bsl;
    p.tilesM = (M + X6M - 1) / X6M;
    p.tilesN = (N + X6N - 1) / X6N;
    // split K while the tiles leave CUs idle and every split keeps >= 16 K steps
    const int64_t tiles = (int64_t)p.tilesM * p.tilesN * batch;
    int splits = 1;
    while (tiles * splits < 256 && K32 / (splits * 2) >= 512) splits *= 2;
    if (tiles * splits < 128) {
        // fewer 256x128 tiles than half the CUs even after splitting K: the
        // 128x128 tiling, split over K down to 4 K blocks per split (faster
        // below this point, slower above it: 1.6x at 4096 x 1024 x 8192)
        p.tilesM = (M + SM - 1) / SM;
        p.tilesN = (N + SM - 1) / SM;
        const int64_t t128 = (int64_t)p.tilesM * p.tilesN * batch;
        int s128 = 1;
        if (t128 < 256 && K32 >= 8 * BK) s128 = (int)std::min<int64_t>((256 + t128 - 1) / t128, K32 / (4 * BK));
        p.kchunk = K32;
        if (s128 > 1) {
            p.kchunk = ((K32 + s128 - 1) / s128 + BK - 1) / BK * BK;
            s128 = (K32 + p.kchunk - 1) / p.kchunk;
        }
        p.splits = s128;
        dim3 grid(p.tilesM * p.tilesN, 1, batch * s128), block(256);
        if (s128 > 1) {
            void* w;
            CT_CHECK(ws_get(h, WS_SPLITK, (size_t)batch * s128 * M * N * 4, &w));
            p.part = (float*)w;
            hipLaunchKernelGGL(gemm_bf16x6_nt_kernel<float>, grid, block, 0, h->stream, q);
            if (tC == CT_F64)
                splitk_reduce<double>(h->stream, p.part, (double*)C, M, N, s128, batch, ldc, bsC, alpha, beta, active);
            else
                splitk_reduce<float>(h->stream, p.part, (float*)C, M, N, s128, batch, ldc, bsC, alpha, beta, active);
        } else if (tC == CT_F64) {
            hipLaunchKernelGGL(gemm_bf16x6_nt_kernel<double>, grid, block, 0, h->stream, q);
        } else {
            hipLaunchKernelGGL(gemm_bf16x6_nt_kernel<float>, grid, block, 0, h->stream, q);
        }
        CT_HIP(hipGetLastError());
        return 0;
    }
    const int per_xcd = (p.tilesM * p.tilesN) / 8;
    p.GM = per_xcd >= p.tilesN ? per_xcd / p.tilesN : 1;
    if (p.GM < 1) p.GM = 1;
    p.splits = splits;
    p.kchunk = splits > 1 ? (K32 / splits + BK - 1) / BK * BK : K32;
    const int64_t total = tiles * splits;
    if (splits > 1 && total % 8 == 0) {
        const int64_t share = std::min<int64_t>((int64_t)p.tilesM * p.tilesN, total / 8);
        p.batch_xcd = 2;
        p.GM = share >= p.tilesN ? (int)(share / p.tilesN) : 1;
    }
    if (splits == 1 && batch % 8 == 0 && p.tilesM * p.tilesN <= 32) {
        p.batch_xcd = 1;
        p.GM = p.tilesM;
    }
    dim3 grid(p.tilesM * p.tilesN, 1, batch * splits), block(256);
    if (splits > 1) {
        void* w;
        CT_CHECK(ws_get(h, WS_SPLITK, (size_t)batch * splits * M * N * 4, &w));
        p.part = (float*)w;
        hipLaunchKernelGGL((gemm_bf16x6_dma_kernel<float, true>), grid, block, 0, h->stream, q);
        if (tC == CT_F64)
            splitk_reduce<double>(h->stream, p.part, (double*)C, M, N, splits, batch, ldc, bsC, alpha, beta, active);
        else
            splitk_reduce<float>(h->stream, p.part, (float*)C, M, N, splits, batch, ldc, bsC, alpha, beta, active);
    } else if (tC == CT_F64) {
        hipLaunchKernelGGL((gemm_bf16x6_dma_kernel<double, false>), grid, block, 0, h->stream, q);
    } else {
        hipLaunchKernelGGL((gemm_bf16x6_dma_kernel<float, false>), grid, block, 0, h->stream, q);
    }
    CT_HIP(hipGetLastError());
    return 0;
}

// Generic entry used by ct_tensordot / the contractions when no pre-split
// operand is available: split both operands into workspace, then run the NT
// kernels.  A(m,k) and B(k,n) follow GemmDesc.
int gemm_bf16x3(Handle* h, int tA, const void* A, int tB, const void* B, int tC, void* C, const GemmDesc& g) {
    const int64_t ld = 2 * split_k32(g.K);
    const int64_t bsA = (int64_t)g.M * ld, bsB = (int64_t)g.N * ld;
    void* wa;
    void* wb;
    CT_CHECK(ws_get(h, WS_CONV_A, (size_t)bsA * g.batch * 2, &wa));
    CT_CHECK(ws_get(h, WS_CONV_B, (size_t)bsB * g.batch * 2, &wb));
    // A(m, k): row = m, stride sAm; k stride sAk
    CT_CHECK(split_bf16(h, tA, A, g.M, g.K, g.sAm, g.sAk, g.batch, g.bsA, (uint16_t*)wa, ld, bsA, g.active));
    // B(k, n) -> row = n: stride sBn; k stride sBk
    CT_CHECK(split_bf16(h, tB, B, g.N, g.K, g.sBn, g.sBk, g.batch, g.bsB, (uint16_t*)wb, ld, bsB, g.active));
    SplitOp sa, sb;
    sa.p = (const uint16_t*)wa; sa.ld = ld; sa.bs = bsA;
    sb.p = (const uint16_t*)wb; sb.ld = ld; sb.bs = bsB;
    return gemm_bf16x3_nt(h, sa, sb, tC, C, g.M, g.N, g.K, g.ldc, g.bsC, g.batch, g.alpha, g.beta, g.active);
}

// The 6xBF16 form of gemm_bf16x3: both operands split with their l planes.
int gemm_bf16x6(Handle* h, int tA, const void* A, int tB, const void* B, int tC, void* C, const GemmDesc& g) {
    const int64_t K32 = split_k32(g.K), ld = 2 * K32;
    const int64_t bsA = (int64_t)g.M * ld, bsB = (int64_t)g.N * ld;
    const int64_t bsAl = (int64_t)g.M * K32, bsBl = (int64_t)g.N * K32;
    void *wa, *wb, *wal, *wbl;
    CT_CHECK(ws_get(h, WS_CONV_A, (size_t)bsA * g.batch * 2, &wa));
    CT_CHECK(ws_get(h, WS_CONV_B, (size_t)bsB * g.batch * 2, &wb));
    CT_CHECK(ws_get(h, WS_CONV_AL, (size_t)bsAl * g.batch * 2, &wal));
    CT_CHECK(ws_get(h, WS_CONV_BL, (size_t)bsBl * g.batch * 2, &wbl));
    CT_CHECK(split_bf16(h, tA, A, g.M, g.K, g.sAm, g.sAk, g.batch, g.bsA, (uint16_t*)wa, ld, bsA, g.active,
                        (uint16_t*)wal, K32, bsAl));
    CT_CHECK(split_bf16(h, tB, B, g.N, g.K, g.sBn, g.sBk, g.batch, g.bsB, (uint16_t*)wb, ld, bsB, g.active,
                        (uint16_t*)wbl, K32, bsBl));
    SplitOp sa, sb;
    sa.p = (const uint16_t*)wa; sa.ld = ld; sa.bs = bsA;
    sa.l = (const uint16_t*)wal; sa.ldl = K32; sa.bsl = bsAl;
    sb.p = (const uint16_t*)wb; sb.ld = ld; sb.bs = bsB;
    sb.l = (const uint16_t*)wbl; sb.ldl = K32; sb.bsl = bsBl;
    return gemm_bf16x6_nt(h, sa, sb, tC, C, g.M, g.N, g.K, g.ldc, g.bsC, g.batch, g.alpha, g.beta, g.active);
}

}  // namespace ct
