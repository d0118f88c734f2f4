// MFMA GEMM kernels for gfx950 (CDNA4, wave64).
//
// One templated kernel serves the FP64 (v_mfma_f64_16x16x4_f64) and FP32
// (v_mfma_f32_16x16x4_f32) paths: both MFMAs take one A and one B element per
// lane (row/col = lane%16, k = lane/16) and return four accumulators per lane
// in column lane%16 (rows: see Mfma<T>::row).  The 3xBF16 path (ct_gemm_bf16.hip)
// uses v_mfma_f32_16x16x32_bf16 with a k-contiguous LDS layout.
//
// Block tile 128x128 (64x64 for small problems), K-tile 16 (FP64, guarded
// FP32 tiles) or 32 (unguarded FP32 tiles), 256 threads = 4 waves in a 2x2
// grid, each wave owning a 64x64 (32x32) sub-tile = 4x4 (2x2) MFMA tiles.  Operands may be m/n-contiguous or
// k-contiguous (the transposes the reference performs with separate permute
// kernels, dense_tensor.mojo:1249-1257, happen here inside the tile load).
// An m/n-contiguous operand is staged in LDS k-major (sX[k][x]), a
// k-contiguous one x-major (sX[x][k], rows padded to 16-byte multiples):
// either way the tile store writes each thread's loaded vector contiguously
// instead of transposing it element by element (measured on the
// [4096 x 8192] . [8192 x 1024] product: FP32 0.577 -> 0.530 ms, FP64
// 1.052 -> 1.035 ms).  The next K-tile is prefetched into registers while
// the MFMAs of the current one run.
#include <algorithm>

#include "ct_internal.h"
#include "ct_device.h"

namespace ct {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// row(lk, r): row of the 16x16 output tile held in accumulator r of a lane
// with lane/16 == lk (the column is lane%16 for both).  The FP64 MFMA
// interleaves rows across the four lane groups; the FP32 one gives each lane
// group four consecutive rows.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

struct KArgs {
    const void* A;
    const void* B;
    void* C;
    int M, N, K;
    int tilesM, tilesN;
    int vecA, vecB;  // 1 when 16-byte vector loads are legal for the operand
    int64_t sAm, sAk, sBk, sBn, ldc, bsA, bsB, bsC;
    double alpha, beta;
    const int* active;
    int splits;  // split-K factor (blockIdx.z = batch * splits + split)
    int kchunk;  // K range per split, a multiple of the kernel's K step
    void* part;  // [batch*splits][M][N] partial sums (TC) when splits > 1
};

// Output tiles of BT x BT (BT = 128, or 64 for problems that 128-wide tiles
// cannot spread over the CUs), K steps of 16, or 32 for FP32 products whose
// tiles need no bounds checks (the FP32 MFMA takes half the cycles per K
// element, and with 32 the per-step tile store and barrier are spread over as
// many MFMA cycles as in FP64).  Guarded FP32 tiles keep 16: at the 2248 x
// 2004 x 512 MPO-mixing products of a w = 562 bond, 32 ran 22.0 ms against
// 18.1 ms.  BK is the unit of the split-K decision.
constexpr int BK = 16;
constexpr int kStepFp32Full = 32;
// padding of an x-major LDS row (sX[x][k]): keeps rows 16-byte aligned
template <typename TC> struct KPad { static constexpr int value = sizeof(TC) == 8 ? 2 : 4; };

// Loads 4 consecutive elements (16-byte aligned when vec) and converts.
template <typename TS, typename TD>
__device__ __forceinline__ void load4(const TS* p, TD* o);
template <>
__device__ __forceinline__ void load4<double, double>(const double* p, double* o) {
    v2d a = *(const v2d*)p, b = *(const v2d*)(p + 2);
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
template <>
__device__ __forceinline__ void load4<double, float>(const double* p, float* o) {
    v2d a = *(const v2d*)p, b = *(const v2d*)(p + 2);
    o[0] = (float)a.x; o[1] = (float)a.y; o[2] = (float)b.x; o[3] = (float)b.y;
}
template <>
__device__ __forceinline__ void load4<float, float>(const float* p, float* o) {
    v4f a = *(const v4f*)p;
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}

// Loads the 16 / sizeof(TS) consecutive elements of one 16-byte vector and converts.
template <typename TS, typename TD>
__device__ __forceinline__ void load16(const TS* p, TD* o) {
    if constexpr (sizeof(TS) == 8) {
        v2d a = *(const v2d*)p;
        o[0] = (TD)a.x; o[1] = (TD)a.y;
    } else {
        v4f a = *(const v4f*)p;
        o[0] = (TD)a.x; o[1] = (TD)a.y; o[2] = (TD)a.z; o[3] = (TD)a.w;
    }
}

// Tile loader.  KC == false: operand X[k][x] (x contiguous, row stride sk);
// KC == true: operand X[x][k] (k contiguous, row stride sx).  Fills
// BT*BK/256 register values per thread (8 for BT = 128, 4 for 64) for a
// BK x BT tile at (k0, x0).  FULL: the tile is entirely inside the operand
// and 16-byte loads are legal (no checks).  For KC, the BK elements of one
// row are read by BK / VE adjacent lanes (VE elements = 16 bytes each), so
// one load instruction covers whole rows' K slices.
template <typename TS, typename TC, bool KC, bool FULL, int BT, int BK>
__device__ __forceinline__ void load_tile(const TS* X, int64_t sx, int64_t sk, int X0, int XN, int k0, int K,
                                          bool vec, TC* r) {
    const int tid = threadIdx.x;
    if (!KC) {
        constexpr int CPR = BT / 4, RPP = 256 / CPR, PASSES = BK / RPP;  // 4-column chunks per row, rows per pass
        const int c = tid % CPR, row = tid / CPR;
#pragma unroll
        for (int h = 0; h < PASSES; ++h) {
            const int k = k0 + row + RPP * h;
            const int x = X0 + 4 * c;
            TC* o = r + 4 * h;
            if (FULL || (vec && k < K && x + 3 < XN)) {
                load4<TS, TC>(X + (int64_t)k * sk + x, o);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    o[e] = (k < K && x + e < XN) ? (TC)X[(int64_t)k * sk + (x + e)] : (TC)0;
            }
        }
    } else {
        constexpr int VE = 16 / (int)sizeof(TS), LPR = BK / VE, RPP = 256 / LPR, PASSES = BT / RPP;
        const int kq = tid % LPR, xr = tid / LPR;
        const int k = k0 + VE * kq;
#pragma unroll
        for (int h = 0; h < PASSES; ++h) {
            const int x = X0 + xr + RPP * h;
            TC* o = r + VE * h;
            if (FULL || (vec && x < XN && k + VE - 1 < K)) {
                load16<TS, TC>(X + (int64_t)x * sx + k, o);
            } else {
#pragma unroll
                for (int e = 0; e < VE; ++e) o[e] = (x < XN && k + e < K) ? (TC)X[(int64_t)x * sx + (k + e)] : (TC)0;
            }
        }
    }
}

// k-major store of an m/n-contiguous operand: S[k][x]
template <typename TC, int LD, int BT, int BK>
__device__ __forceinline__ void store_tile_k(TC (*S)[LD], const TC* r) {
    const int tid = threadIdx.x;
    constexpr int CPR = BT / 4, RPP = 256 / CPR, PASSES = BK / RPP;
    const int c = tid % CPR, row = tid / CPR;
#pragma unroll
    for (int h = 0; h < PASSES; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e) S[row + RPP * h][4 * c + e] = r[4 * h + e];
}

// x-major store of a k-contiguous operand: S[x][k], each thread's VE
// consecutive k written contiguously
template <typename TS, typename TC, int LD, int BT, int BK>
__device__ __forceinline__ void store_tile_x(TC (*S)[LD], const TC* r) {
    const int tid = threadIdx.x;
    constexpr int VE = 16 / (int)sizeof(TS), LPR = BK / VE, RPP = 256 / LPR, PASSES = BT / RPP;
    const int kq = tid % LPR, xr = tid / LPR;
#pragma unroll
    for (int h = 0; h < PASSES; ++h)
#pragma unroll
        for (int e = 0; e < VE; ++e) S[xr + RPP * h][VE * kq + e] = r[VE * h + e];
}

// LDS is double-buffered: the tile for k0+BK is loaded into registers before
// the MFMAs on k0 and stored to the other buffer after them, so each K step
// needs a single barrier.
template <typename TC, typename TA, typename TB, typename TO, bool AKC, bool BKC, bool FULL, int BT, int KSTEP>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(KArgs p) {
    if (p.active && *p.active == 0) return;
    constexpr int BK = KSTEP;
    constexpr int PAD = 4, KP = BK + KPad<TC>::value;
    constexpr int NI = BT / 32;   // 16x16 accumulator blocks per wave and dimension (2 x 2 waves)
    constexpr int NR = BT * BK / 256;
    // k-major (x-contiguous operand) or x-major (k-contiguous operand) LDS tiles
    __shared__ __attribute__((aligned(16))) TC sAk[AKC ? 1 : 2][AKC ? 1 : BK][BT + PAD];
    __shared__ __attribute__((aligned(16))) TC sAx[AKC ? 2 : 1][AKC ? BT : 1][KP];
    __shared__ __attribute__((aligned(16))) TC sBk[BKC ? 1 : 2][BKC ? 1 : BK][BT + PAD];
    __shared__ __attribute__((aligned(16))) TC sBx[BKC ? 2 : 1][BKC ? BT : 1][KP];
    typedef typename Mfma<TC>::acc_t acc_t;

    int tm, tn;
    gemm_tile_coords(blockIdx.x, gridDim.x, p.tilesM, p.tilesN, tm, tn);
    const int m0 = tm * BT, n0 = tn * BT;
    const int zz = blockIdx.z;
    const int z = zz / p.splits;
    const int kbeg = (zz - z * p.splits) * p.kchunk;
    const int kend = min(p.K, kbeg + p.kchunk);
    const TA* A = (const TA*)p.A + (int64_t)z * p.bsA;
    const TB* B = (const TB*)p.B + (int64_t)z * p.bsB;
    TO* C = (TO*)p.C + (int64_t)z * p.bsC;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int li = lane & 15, lk = lane >> 4;

    acc_t acc[NI][NI];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    TC ra[NR], rb[NR];
    const int64_t sAx_ = AKC ? p.sAm : 0, sAk_ = AKC ? 0 : p.sAk;
    const int64_t sBx_ = BKC ? p.sBn : 0, sBk_ = BKC ? 0 : p.sBk;
    auto store = [&](int buf) {
        if constexpr (AKC) store_tile_x<TA, TC, KP, BT, BK>(sAx[buf], ra);
        else store_tile_k<TC, BT + PAD, BT, BK>(sAk[buf], ra);
        if constexpr (BKC) store_tile_x<TB, TC, KP, BT, BK>(sBx[buf], rb);
        else store_tile_k<TC, BT + PAD, BT, BK>(sBk[buf], rb);
    };
    load_tile<TA, TC, AKC, FULL, BT, BK>(A, sAx_, sAk_, m0, p.M, kbeg, kend, p.vecA, ra);
    load_tile<TB, TC, BKC, FULL, BT, BK>(B, sBx_, sBk_, n0, p.N, kbeg, kend, p.vecB, rb);
    store(0);
    __syncthreads();

    int cur = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        const bool more = k0 + BK < kend;
        if (more) {
            load_tile<TA, TC, AKC, FULL, BT, BK>(A, sAx_, sAk_, m0, p.M, k0 + BK, kend, p.vecA, ra);
            load_tile<TB, TC, BKC, FULL, BT, BK>(B, sBx_, sBk_, n0, p.N, k0 + BK, kend, p.vecB, rb);
        }
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            TC a[NI], b[NI];
            const int k = kk * 4 + lk;
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                const int x = wm * (BT / 2) + i * 16 + li;
                a[i] = AKC ? sAx[cur][x][k] : sAk[cur][k][x];
            }
#pragma unroll
            for (int j = 0; j < NI; ++j) {
                const int x = wn * (BT / 2) + j * 16 + li;
                b[j] = BKC ? sBx[cur][x][k] : sBk[cur][k][x];
            }
#pragma unroll
            for (int i = 0; i < NI; ++i)
#pragma unroll
                for (int j = 0; j < NI; ++j) acc[i][j] = Mfma<TC>::op(a[i], b[j], acc[i][j]);
        }
        if (more) store(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }

    if (p.splits > 1) {  // raw partial sums of this K range; the reduction applies alpha / beta
        TC* P = (TC*)p.part + (int64_t)zz * p.M * p.N;
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < NI; ++j) {
                const int n = n0 + wn * (BT / 2) + j * 16 + li;
                if (!FULL && n >= p.N) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * (BT / 2) + i * 16 + Mfma<TC>::row(lk, r);
                    if (FULL || m < p.M) P[(int64_t)m * p.N + n] = acc[i][j][r];
                }
            }
        return;
    }
    const double alpha = p.alpha, beta = p.beta;
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int n = n0 + wn * (BT / 2) + j * 16 + li;
            if (!FULL && n >= p.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm * (BT / 2) + i * 16 + Mfma<TC>::row(lk, r);
                if (FULL || m < p.M) {
                    TO* c = C + (int64_t)m * p.ldc + n;
                    double v = alpha * (double)acc[i][j][r];
                    if (beta != 0.0) v += beta * (double)(*c);
                    *c = (TO)v;
                }
            }
        }
}

// C[b][m][n] = alpha * sum_s part[b][s][m][n] (summed in FP64) + beta * C
template <typename TP, typename TO>
__global__ __launch_bounds__(256) void gemm_splitk_reduce(const TP* __restrict__ part, TO* __restrict__ C, int M, int N,
                                                          int64_t ldc, int64_t bsC, int splits, int batch, double alpha,
                                                          double beta, const int* active) {
    if (active && *active == 0) return;
    const int64_t MN = (int64_t)M * N, total = MN * batch;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / MN, e = t - b * MN;
        const int64_t m = e / N, n = e - m * N;
        double s = 0.0;
        for (int k = 0; k < splits; ++k) s += (double)part[(b * splits + k) * MN + e];
        TO* c = C + b * bsC + m * ldc + n;
        double v = alpha * s;
        if (beta != 0.0) v += beta * (double)(*c);
        *c = (TO)v;
    }
}

template <typename T>
bool aligned16(const void* ptr, int64_t s1, int64_t s2) {
    constexpr int64_t q = 16 / sizeof(T);
    return ((uintptr_t)ptr % 16 == 0) && (s1 % q == 0) && (s2 % q == 0);
}

// Split K when the tiles alone would leave most CUs idle (small bond
// dimensions): at least 4 K-tiles per split, aiming at >= 256 workgroups.
// A short K loop is bound by the load latency of each step (a lone
// workgroup per CU hides none of it): on the [328 x 410] . [410 x 82]
// products of a D = 82, w = 5 matvec, 3 splits of 9 K steps took 15.7 us
// and the L = 64 XXZ sweep's Lanczos 0.126 s; 6 splits of 5 steps 0.108 s.
// With long K, also when they give fewer than two workgroups per CU: a
// lone workgroup has nothing to run while it waits at its per-step
// barrier (the [4096 x 8192] . [8192 x 1024] product of the D = 1024
// matvec is 256 tiles of 128 x 128).
void plan_splits(const GemmDesc& g, int BT, int BKT, int* splits_out, int* kchunk_out) {
    const int64_t tiles = (int64_t)((g.M + BT - 1) / BT) * ((g.N + BT - 1) / BT) * g.batch;
    int splits = 1;
    if (tiles < 128 && g.K >= 8 * BK) splits = (int)std::min<int64_t>((256 + tiles - 1) / tiles, g.K / (4 * BK));
    else if (tiles < 512 && g.K >= 128 * BK) splits = (int)std::min<int64_t>((512 + tiles - 1) / tiles, g.K / (64 * BK));
    int kchunk = g.K;
    if (splits > 1) {
        kchunk = (int)(((g.K + splits - 1) / splits + BKT - 1) / BKT * BKT);  // whole K steps of the kernel
        splits = (g.K + kchunk - 1) / kchunk;
    }
    *splits_out = splits;
    *kchunk_out = kchunk;
}

// K step of the kernel: kStepFp32Full for FP32 (elem_size 4) when the tiles
// need no bounds checks and K is a multiple of it, else BK.
int kstep_for(size_t elem_size, bool tiles_full, int64_t K) {
    return (elem_size == 4 && tiles_full && K % kStepFp32Full == 0) ? kStepFp32Full : BK;
}

// Tile size of the FP64/FP32 GEMM: 64x64 when 128x128 tiles would occupy
// fewer than half the CUs and split-K cannot make up for it (K < 2048, or
// fewer than 16 tiles): four times the workgroups for the same problem.  With
// long K the 128x128 tiles split over K fill the GPU at a better
// MFMA-to-load ratio (measured at D = 512).
int plan_tile(const GemmDesc& g) {
    const int64_t tiles128 = (int64_t)((g.M + 127) / 128) * ((g.N + 127) / 128) * g.batch;
    return (tiles128 < 128 && (g.K < 2048 || tiles128 < 16)) ? 64 : 128;
}

template <typename TC, typename TA, typename TB, typename TO, int BT>
int launch_bt(Handle* h, const void* A, const void* B, void* C, const GemmDesc& g) {
    KArgs p;
    p.A = A; p.B = B; p.C = C;
    p.M = g.M; p.N = g.N; p.K = g.K;
    p.tilesM = (g.M + BT - 1) / BT;
    p.tilesN = (g.N + BT - 1) / BT;
    p.sAm = g.sAm; p.sAk = g.sAk; p.sBk = g.sBk; p.sBn = g.sBn;
    p.ldc = g.ldc; p.bsA = g.bsA; p.bsB = g.bsB; p.bsC = g.bsC;
    p.alpha = g.alpha; p.beta = g.beta; p.active = g.active;
    const bool akc = g.a_kc != 0;
    const bool bkc = g.b_kc != 0;
    p.vecA = akc ? aligned16<TA>(A, g.sAm, g.bsA) : aligned16<TA>(A, g.sAk, g.bsA);
    p.vecB = bkc ? aligned16<TB>(B, g.sBn, g.bsB) : aligned16<TB>(B, g.sBk, g.bsB);
    const bool tiles_full = p.vecA && p.vecB && (g.M % BT == 0) && (g.N % BT == 0);
    const int BKT = kstep_for(sizeof(TC), tiles_full, g.K);
    int splits = 1;
    plan_splits(g, BT, BKT, &splits, &p.kchunk);
    p.splits = splits;
    p.part = nullptr;
    if (splits > 1) CT_CHECK(ws_get(h, WS_SPLITK, (size_t)g.batch * splits * g.M * g.N * sizeof(TC), &p.part));
    dim3 grid(p.tilesM * p.tilesN, 1, g.batch * splits), block(256);
    const bool full = tiles_full && (g.K % BKT == 0);
#define CT_LAUNCH(AK, BKc, F, KS) \
    hipLaunchKernelGGL((gemm_mfma_kernel<TC, TA, TB, TO, AK, BKc, F, BT, KS>), grid, block, 0, h->stream, p)
    if (BKT == kStepFp32Full) {  // FP32, unguarded tiles (kstep_for returns this only then)
        if constexpr (sizeof(TC) == 4) {
            if (!akc && !bkc) CT_LAUNCH(false, false, true, kStepFp32Full);
            else if (akc && !bkc) CT_LAUNCH(true, false, true, kStepFp32Full);
            else if (!akc && bkc) CT_LAUNCH(false, true, true, kStepFp32Full);
            else CT_LAUNCH(true, true, true, kStepFp32Full);
        }
    } else if (full) {
        if (!akc && !bkc) CT_LAUNCH(false, false, true, BK);
        else if (akc && !bkc) CT_LAUNCH(true, false, true, BK);
        else if (!akc && bkc) CT_LAUNCH(false, true, true, BK);
        else CT_LAUNCH(true, true, true, BK);
    } else {
        if (!akc && !bkc) CT_LAUNCH(false, false, false, BK);
        else if (akc && !bkc) CT_LAUNCH(true, false, false, BK);
        else if (!akc && bkc) CT_LAUNCH(false, true, false, BK);
        else CT_LAUNCH(true, true, false, BK);
    }
#undef CT_LAUNCH
    CT_HIP(hipGetLastError());
    if (splits > 1) {
        const int64_t total = (int64_t)g.batch * g.M * g.N;
        const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
        hipLaunchKernelGGL((gemm_splitk_reduce<TC, TO>), dim3(blocks), dim3(256), 0, h->stream, (const TC*)p.part,
                           (TO*)C, g.M, g.N, g.ldc, g.bsC, splits, g.batch, g.alpha, g.beta, g.active);
        CT_HIP(hipGetLastError());
    }
    return 0;
}

template <typename TC, typename TA, typename TB, typename TO>
int launch_typed(Handle* h, const void* A, const void* B, void* C, const GemmDesc& g) {
    if (plan_tile(g) == 64) return launch_bt<TC, TA, TB, TO, 64>(h, A, B, C, g);
    return launch_bt<TC, TA, TB, TO, 128>(h, A, B, C, g);
}

template <typename TC>
int dispatch_io(Handle* h, int tA, const void* A, int tB, const void* B, int tC, void* C, const GemmDesc& g) {
    // operand element types: FP64 or FP32 each
#define CT_CASE(a, b, c, TA, TB, TO) \
    if (tA == a && tB == b && tC == c) return launch_typed<TC, TA, TB, TO>(h, A, B, C, g);
    CT_CASE(CT_F64, CT_F64, CT_F64, double, double, double)
    CT_CASE(CT_F32, CT_F32, CT_F32, float, float, float)
    CT_CASE(CT_F32, CT_F64, CT_F32, float, double, float)
    CT_CASE(CT_F32, CT_F64, CT_F64, float, double, double)
    CT_CASE(CT_F32, CT_F32, CT_F64, float, float, double)
    CT_CASE(CT_F64, CT_F32, CT_F32, double, float, float)
    CT_CASE(CT_F64, CT_F64, CT_F32, double, double, float)
    CT_CASE(CT_F64, CT_F32, CT_F64, double, float, double)
#undef CT_CASE
    set_last_error("gemm: unsupported operand type combination");
    return CT_STATUS_NOT_SUPPORTED;
}

}  // namespace

int gemm(Handle* h, int prec, int tA, const void* A, int tB, const void* B, int tC, void* C, const GemmDesc& g) {
    if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return 0;
    if (g.K <= 0) {
        // C = beta * C: only beta == 0 occurs in this library
        if (g.beta != 0.0) {
            set_last_error("gemm: K == 0 with beta != 0");
            return CT_STATUS_NOT_SUPPORTED;
        }
        for (int b = 0; b < g.batch; ++b) {
            size_t es = tC == CT_F64 ? 8 : 4;
            char* c = (char*)C + (size_t)b * g.bsC * es;
            CT_HIP(hipMemset2DAsync(c, (size_t)g.ldc * es, 0, (size_t)g.N * es, (size_t)g.M, h->stream));
        }
        return 0;
    }
    if (prec == CT_PREC_FP64) {
        if (tA != CT_F64 || tB != CT_F64 || tC != CT_F64) {
            set_last_error("FP64 contraction requires FP64 buffers");
            return CT_STATUS_NOT_SUPPORTED;
        }
        return launch_typed<double, double, double, double>(h, A, B, C, g);
    }
    if (prec == CT_PREC_FP32) return dispatch_io<float>(h, tA, A, tB, B, tC, C, g);
    if (prec == CT_PREC_BF16X3) return gemm_bf16x3(h, tA, A, tB, B, tC, C, g);
    if (prec == CT_PREC_BF16X6) return gemm_bf16x6(h, tA, A, tB, B, tC, C, g);
    set_last_error("unknown precision");
    return CT_STATUS_NOT_SUPPORTED;
}

size_t gemm_splitk_bytes(int prec, const GemmDesc& g) {
    if (g.M <= 0 || g.N <= 0 || g.batch <= 0 || g.K <= 0) return 0;
    // the K step also depends on the operands' alignment, known only at launch:
    // take the larger partial-sum buffer of the two possible steps
    const size_t es = prec == CT_PREC_FP64 ? 8 : 4;
    size_t bytes = 0;
    for (int kstep : {BK, kStepFp32Full}) {
        if (es == 8 && kstep != BK) continue;
        int splits = 1, kchunk = 0;
        plan_splits(g, plan_tile(g), kstep, &splits, &kchunk);
        if (splits > 1) bytes = std::max(bytes, (size_t)g.batch * splits * g.M * g.N * es);
    }
    return bytes;
}

}  // namespace ct
