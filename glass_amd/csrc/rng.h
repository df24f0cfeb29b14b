// rng.h -- the reference's random numbers on the device: rand_r of the GNU C
// library and the two draws built on it (utils/src/glass_utils.c:19-35).
// Every chain slot carries its own 32-bit rand_r state, as chain->r[ic].
#pragma once
#include <hip/hip_runtime.h>

namespace glass {

// rand_r of the GNU C library: three steps of its linear congruential
// generator contribute 11 + 10 + 10 bits; RAND_MAX = 2^31 - 1
__device__ inline int libc_rand_r(unsigned* seed)
{
    unsigned next = *seed;
    int result;
    next *= 1103515245u;
    next += 12345u;
    result = (int)((next / 65536u) % 2048u);
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (int)((next / 65536u) % 1024u);
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (int)((next / 65536u) % 1024u);
    *seed = next;
    return result;
}

// rand_r_U_0_1
__device__ inline double rand_U(unsigned* s) { return (double)libc_rand_r(s) / 2147483647.0; }

// rand_r_N_0_1 (Marsaglia's polar method)
__device__ inline double rand_N(unsigned* s)
{
    // rounded products, as the C library build computes them: a fused
    // u*u + v*v moves q by an ulp, which log(q) magnifies by 1/(1-q)
#pragma clang fp contract(off)
    double u, v, q;
    do
    {
        u = 2 * rand_U(s) - 1;
        v = 2 * rand_U(s) - 1;
        q = u * u + v * v;
    } while (q >= 1 || q == 0);
    return u * sqrt(-2 * log(q) / q);
}

}  // namespace glass
