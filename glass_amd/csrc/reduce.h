// reduce.h -- deterministic FP64 block reductions (wave64 shuffles + LDS).
#pragma once
#include <hip/hip_runtime.h>

namespace glass {

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;  // valid in lane 0
}

// Sum over the block; the result is valid in thread 0.  Fixed order, so
// results are reproducible run to run.
template <int NTHR>
__device__ inline double block_sum(double v)
{
    static_assert(NTHR % 64 == 0, "wave64");
    __shared__ double part[NTHR / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NTHR / 64; i++) r += part[i];
    __syncthreads();
    return r;
}

}  // namespace glass
