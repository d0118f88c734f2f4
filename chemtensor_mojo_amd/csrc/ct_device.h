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

}  // namespace ct
