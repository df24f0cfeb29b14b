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

// K sums over the block at once: each wave's tree sum of each value, one
// barrier, then thread k adds the waves' partials of value k in order into
// out[k] (LDS or global); ends with a barrier.  The per-value order is
// block_sum's, so the results are the same as K separate block_sum calls.
template <int NTHR, int K>
__device__ inline void block_sum_many(const double (&v)[K], double* out)
{
    static_assert(NTHR % 64 == 0 && K <= NTHR, "wave64 workgroups, K <= threads");
    __shared__ double part[NTHR / 64][K];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++)
    {
        const double x = wave_sum(v[k]);
        if (lane == 0) part[w][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < K)
    {
        double r = 0.0;
        for (int i = 0; i < NTHR / 64; i++) r += part[i][threadIdx.x];
        out[threadIdx.x] = r;
    }
    __syncthreads();
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
