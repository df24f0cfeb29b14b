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

// The least power of two >= k (k >= 1).  Written as a loop, and used only
// to initialise constexpr variables, so it is always evaluated at compile
// time: a recursive helper called at run time would make the calling kernel's
// stack size unbounded.
constexpr int pow2_at_least(int k)
{
    int p = 1;
    while (p < k) p *= 2;
    return p;
}

// One exchange step at distance D of wave_sum_transpose with N values left.
template <int D, int N, int P>
__device__ __forceinline__ void wave_sum_transpose_step(double (&a)[P], int lane)
{
    if constexpr (N > 1)
    {
        constexpr int H = N / 2;
        const bool up = (lane & D) != 0;
#pragma unroll
        for (int i = 0; i < H; i++)
        {
            const double send = up ? a[i] : a[i + H];
            const double keep = up ? a[i + H] : a[i];
            a[i] = keep + __shfl_xor(send, D, 64);
        }
    }
    else
        a[0] += __shfl_xor(a[0], D, 64);
    if constexpr (D > 1) wave_sum_transpose_step<D / 2, (N > 1 ? N / 2 : 1), P>(a, lane);
}

// K per-lane values summed over the wave by transposition: at each exchange
// step (partner lane ^ d, d = 32, 16, ...) a lane keeps half of its values,
// the lower half where bit d of the lane is clear, and adds the partner's
// copies of them; with one value left, the remaining steps are plain
// butterflies.  K values padded to P = 2^p cost P - 1 double shuffles (63
// for 45 values against 270 for 45 separate tree sums).  Lane L returns the
// sum of value index L / (64 / P) (every lane of that group holds it).
// Fixed order: reproducible run to run.
template <int K>
__device__ inline double wave_sum_transpose(const double (&v)[K])
{
    static_assert(K >= 1 && K <= 64, "at most 64 values");
    constexpr int P = pow2_at_least(K);
    double a[P];
#pragma unroll
    for (int i = 0; i < P; i++) a[i] = i < K ? v[i] : 0.0;
    wave_sum_transpose_step<32, P, P>(a, threadIdx.x & 63);
    return a[0];
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
