// Device-side helpers shared by the kernel translation units (each TU gets
// its own inline copy; the library is built without relocatable device code).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ct {

// Sum over the 64 lanes of a wavefront; the result is valid in every lane.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the 64 lanes of a wavefront through DPP lane moves (no LDS
// crossbar): butterfly inside each row of 16 lanes, then the four row sums in
// a fixed order.  Every lane gets the same value, and every wavefront that
// sums the same 64 inputs gets the same bits.
template <int CTRL>
__device__ __forceinline__ double dpp_mov64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ double readlane64(double v, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_mov64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov64<0x141>(v);  // row_half_mirror
    v += dpp_mov64<0x140>(v);  // row_mirror
    return (readlane64(v, 0) + readlane64(v, 16)) + (readlane64(v, 32) + readlane64(v, 48));
}

// Sum over a 256-thread block (4 wavefronts); valid in every thread.
// Contains barriers: call from all threads of the block.
__device__ __forceinline__ double block_sum256(double v) {
    __shared__ double red[4];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// Tile coordinates of a GEMM workgroup.  Consecutive workgroup ids are
// dispatched round-robin over the 8 XCDs of an MI355X, each with its own L2.
// The id is first regrouped so that every XCD owns a contiguous range of
// tiles, then walked in groups of GM tile rows (column-major inside a group),
// so the A row-panels and B column-panels one XCD touches overlap and stay in
// its L2.
__device__ __forceinline__ void gemm_group_coords(int bid, int tilesM, int tilesN, int GM, int& tm, int& tn) {
    const int group = bid / (GM * tilesN);
    const int first = group * GM;
    const int gm = min(tilesM - first, GM);
    const int in = bid - group * GM * tilesN;
    tm = first + in % gm;
    tn = in / gm;
}

__device__ __forceinline__ void gemm_tile_coords(int bid, int nb, int tilesM, int tilesN, int GM, int& tm, int& tn) {
    if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
    gemm_group_coords(bid, tilesM, tilesN, GM, tm, tn);
}

__device__ __forceinline__ void gemm_tile_coords(int bid, int nb, int tilesM, int tilesN, int& tm, int& tn) {
    gemm_tile_coords(bid, nb, tilesM, tilesN, 8, tm, tn);
}

// Batched launches with few tiles per batch (grid = tiles x 1 x batch, batch a
// multiple of 8): workgroup f = z*nb + x runs on XCD f mod 8, so give XCD c
// whole batches c, c+8, ... and walk their tiles in order.  Every batch's
// operand panels then leave HBM once instead of once per XCD.  A bijection of
// (x, z) onto (batch, tile).
__device__ __forceinline__ void gemm_batch_per_xcd(int nb, int& b, int& tile) {
    const int f = (int)blockIdx.z * nb + (int)blockIdx.x;
    const int s = f >> 3;
    b = (f & 7) + 8 * (s / nb);
    tile = s % nb;
}

}  // namespace ct
