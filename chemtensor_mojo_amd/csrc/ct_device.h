// Device-side helpers shared by the kernel translation units (each TU gets
// its own inline copy; the library is built without relocatable device code).
#pragma once
#include <hip/hip_runtime.h>

namespace ct {

// Sum over the 64 lanes of a wavefront; the result is valid in every lane.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
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
__device__ __forceinline__ void gemm_tile_coords(int bid, int nb, int tilesM, int tilesN, int GM, int& tm, int& tn) {
    if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
    const int group = bid / (GM * tilesN);
    const int first = group * GM;
    const int gm = min(tilesM - first, GM);
    const int in = bid - group * GM * tilesN;
    tm = first + in % gm;
    tn = in / gm;
}

__device__ __forceinline__ void gemm_tile_coords(int bid, int nb, int tilesM, int tilesN, int& tm, int& tn) {
    gemm_tile_coords(bid, nb, tilesM, tilesN, 8, tm, tn);
}

}  // namespace ct
