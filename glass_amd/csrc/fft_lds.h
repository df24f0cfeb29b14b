// fft_lds.h -- batched in-LDS complex FFTs for one workgroup (gfx950).
//
// NL independent length-N forward transforms stored link-major in LDS
// (buf[l*N + n]), unnormalised, exp(-2 pi i n k / N) convention, as
// kiss_fft forward (external/kissfft/kiss_fft.c:337-369).  Stockham
// autosort formulation: one radix-3 stage when 3 divides N (N = 3*2^a, the
// WDM sizes), one radix-2 stage when the power-of-two part has an odd
// exponent, radix 4 for the rest.
// Each stage reads every butterfly's inputs into registers, synchronises,
// then writes the outputs into the same buffer, so a single LDS buffer
// suffices (96*N bytes for the six link series of a template).
//
// Twiddles come from a table tw[k] = exp(-2 pi i k / NTW) of NTW entries
// (NTW a multiple of N: 4096 or 16384 for N = 2^a, 12288 for N = 3*2^a) built once
// per device with the phase formed as -2*pi*k/NTW, the kiss_fft twiddle
// formula.
#pragma once
#include <hip/hip_runtime.h>

namespace glass {

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Tables that do not change between templates (FFT twiddles, spacecraft
// positions) are read through opaque(p) wherever a loop or a second pass
// would read them again.  Left alone, the compiler keeps the values of the
// first read live in registers until the second (every twiddle of every
// FFT stage: ~30-100 VGPRs, halving occupancy).  Passing the pointer
// through an empty asm makes each read a fresh load.  The pointer goes
// through the asm as a global-memory pointer: a generic pointer of unknown
// origin would turn every table read into a flat load, which the hardware
// counts against the LDS wait counter as well, so each FFT stage's LDS
// wait would also wait for its twiddle loads.
template <typename T>
__device__ inline const T* opaque(const T* p)
{
    auto g = (const __attribute__((address_space(1))) T*)p;
    asm volatile("" : "+s"(g));
    return (const T*)g;
}

// LDS slot of complex element i of a buffer transformed with the swizzled
// layout (SWZ = true below): bits 0-3 of i XOR-ed with bits 2-5 and bit 0
// also with bit 3.  The map is linear over GF(2) (slot(a ^ b) = slot(a) ^
// slot(b)) and a bijection of every aligned block of 64 elements.  With it
// the reads of consecutive elements stay conflict-free and so do the
// strided writes of the first Stockham stages (64 bytes apart per lane at
// Ns = 1, runs of four lanes 256 bytes apart at Ns = 4), which in the plain
// layout put 2-4 lanes of every 8 or 16 on the same LDS banks.  Linearity
// lets a butterfly swizzle its base index once: its R elements sit at
// disjoint-bit offsets m*span, whose slots are compile-time constants.
__device__ __forceinline__ int lds_swz(int i) { return i ^ ((i >> 2) & 15) ^ ((i >> 3) & 1); }
template <bool SWZ>
__device__ __forceinline__ int lds_slot(int i)
{
    return SWZ ? lds_swz(i) : i;
}
// Slot of a + c for a and c with no set bit in common, from the slot of a.
template <bool SWZ>
__device__ __forceinline__ int lds_slot_add(int slot_a, int c)
{
    return SWZ ? (slot_a ^ lds_swz(c)) : slot_a + c;
}

// Radix-R butterfly (R = 2, 3, 4) of twiddled inputs v into y:
// y[m] = sum_q v[q] exp(-2 pi i q m / R).
template <int R>
__device__ inline void radix_butterfly(const double2 (&v)[R], double2 (&y)[R])
{
    if constexpr (R == 3)
    {
        // w = exp(-2 pi i/3): y1 = a0 - s/2 - i(sqrt3/2) d, y2 = conj rotation
        const double h3 = 0.86602540378443864676;
        double2 a0 = v[0], a1 = v[1], a2 = v[2];
        double2 s12 = make_double2(a1.x + a2.x, a1.y + a2.y);
        double2 d12 = make_double2(a1.x - a2.x, a1.y - a2.y);
        double2 t = make_double2(a0.x - 0.5 * s12.x, a0.y - 0.5 * s12.y);
        y[0] = make_double2(a0.x + s12.x, a0.y + s12.y);
        y[1] = make_double2(t.x + h3 * d12.y, t.y - h3 * d12.x);
        y[2] = make_double2(t.x - h3 * d12.y, t.y + h3 * d12.x);
    }
    else if constexpr (R == 4)
    {
        double2 a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        double2 s02 = make_double2(a0.x + a2.x, a0.y + a2.y);
        double2 d02 = make_double2(a0.x - a2.x, a0.y - a2.y);
        double2 s13 = make_double2(a1.x + a3.x, a1.y + a3.y);
        double2 d13 = make_double2(a1.x - a3.x, a1.y - a3.y);
        y[0] = make_double2(s02.x + s13.x, s02.y + s13.y);
        y[2] = make_double2(s02.x - s13.x, s02.y - s13.y);
        // -i * d13 = (d13.y, -d13.x)
        y[1] = make_double2(d02.x + d13.y, d02.y - d13.x);
        y[3] = make_double2(d02.x - d13.y, d02.y + d13.x);
    }
    else
    {
        y[0] = make_double2(v[0].x + v[1].x, v[0].y + v[1].y);
        y[1] = make_double2(v[0].x - v[1].x, v[0].y - v[1].y);
    }
}

// One Stockham stage of radix R over NL transforms; butterflies are spread
// over the workgroup's NTHR threads.  Ns is the span of the previous stage.
// The twiddles are loaded first, so their memory latency overlaps the LDS
// reads and the barrier; when the workgroup spans whole series (NTHR a
// multiple of N/R) every butterfly of a thread has the same index j and
// shares one set.  Twiddle indices are in range for any b (j < N/R), so
// the loads need no guard.
// Orders a wave's LDS accesses before and after it, for the other lanes of
// the same wave: the compiler keeps loads and stores on their side, and the
// LDS serves one wave's instructions in issue order, so no instruction is
// emitted (wavefront-scope fences).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// WAVE: every series of the stage is read and written by the lanes of one
// wave, the same wave as in the neighbouring stages (lds_fft decides), so
// the stage's two synchronisations are wave-level; LAST: its final one is a
// workgroup barrier all the same (the caller reads across waves after it).
template <int N, int NL, int R, int NTHR, bool SWZ = false, bool WAVE = false, bool LAST = true>
__device__ __forceinline__ void stockham_stage(double2* buf, int Ns, const double2* __restrict__ tw, int twstride,
                                               int tid = threadIdx.x)
{
    static_assert(!SWZ || (N % 64 == 0 && R != 3), "swizzled layout: power-of-two N >= 64");
    constexpr int NB = NL * (N / R);                 // butterflies in the stage
    constexpr int PER = (NB + NTHR - 1) / NTHR;      // per thread
    constexpr bool SAMEJ = NTHR % (N / R) == 0;
    constexpr int NW = SAMEJ ? 1 : PER;
    // twiddle exp(-2 pi i m k /(R Ns)) = tw[m*k*(NTW/(R*Ns))]; the first
    // stage (Ns = 1, a constant once inlined) has k = 0, all twiddles 1,
    // and skips the loads and products
    double2 w[NW][R - 1];
    if (Ns > 1)
#pragma unroll
        for (int q = 0; q < NW; q++)
        {
            const int k = ((tid + q * NTHR) % (N / R)) % Ns;
            const int step = k * twstride * (N / (R * Ns));
#pragma unroll
            for (int m = 1; m < R; m++) w[q][m - 1] = tw[m * step];
        }
    double2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; p++)
    {
        int b = tid + p * NTHR;
        if (NB % NTHR == 0 || b < NB)
        {
            int l = b / (N / R), j = b % (N / R);
            const int sj = lds_slot<SWZ>(j);  // j < N/R: disjoint from m*(N/R)
#pragma unroll
            for (int m = 0; m < R; m++) v[p][m] = buf[l * N + lds_slot_add<SWZ>(sj, m * (N / R))];
        }
    }
    if constexpr (WAVE)
        wave_lds_sync();
    else
        __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; p++)
    {
        int b = tid + p * NTHR;
        if (NB % NTHR == 0 || b < NB)
        {
            int l = b / (N / R), j = b % (N / R);
            int k = j % Ns;
            if (Ns > 1)
#pragma unroll
                for (int m = 1; m < R; m++) v[p][m] = cmul(v[p][m], w[SAMEJ ? 0 : p][m - 1]);
            double2 y[R];
            radix_butterfly<R>(v[p], y);
            // base has no bits in [log2 Ns, log2 (R Ns)), where m*Ns lives
            const int sb = lds_slot<SWZ>((j / Ns) * Ns * R + k);
#pragma unroll
            for (int m = 0; m < R; m++) buf[l * N + lds_slot_add<SWZ>(sb, m * Ns)] = y[m];
        }
    }
    if constexpr (WAVE && !LAST)
        wave_lds_sync();
    else
        __syncthreads();
}

// Full forward FFT of NL length-N series in LDS. twstride = NTW / N.
// SWZ: element n of series l lives at buf[l*N + lds_swz(n)] on entry and
// on return.
template <int N, int NL, int NTHR, bool SWZ = false>
__device__ __forceinline__ void lds_fft(double2* buf, const double2* __restrict__ tw, int twstride)
{
    constexpr bool R3 = N % 3 == 0;
    constexpr int P2 = R3 ? N / 3 : N;
    static_assert((P2 & (P2 - 1)) == 0 && N >= 4, "N = 2^a or 3*2^a");
    constexpr int LOG2 = __builtin_ctz(P2);
    int tid = threadIdx.x;
    if constexpr (SWZ)
    {
        // a fresh thread index per transform: its swizzled addresses are
        // not shared with another transform's, so none stays live between
        asm volatile("" : "+v"(tid));
        tid &= 1023;  // the range threadIdx.x has
    }
    int Ns = 1;
    if (R3)
    {
        stockham_stage<N, NL, 3, NTHR>(buf, Ns, tw, twstride);
        Ns *= 3;
    }
    if (LOG2 & 1)
    {
        stockham_stage<N, NL, 2, NTHR, SWZ>(buf, Ns, tw, twstride, tid);
        Ns *= 2;
    }
    // radix-4 stages only, N/4 <= 64 butterflies per series: butterfly b =
    // tid + p*NTHR of series b / (N/4) lies in one wave (N/4 divides 64 and
    // NTHR), the same wave in every stage, so only the last stage ends with
    // a workgroup barrier (the caller has synchronised the input)
    constexpr bool WAVE = !R3 && !(LOG2 & 1) && N / 4 <= 64 && NTHR % 64 == 0;
#pragma unroll
    for (int s = 0; s < LOG2 / 2; s++)
    {
        if (s == LOG2 / 2 - 1)
            stockham_stage<N, NL, 4, NTHR, SWZ, WAVE, true>(buf, Ns, tw, twstride, tid);
        else
            stockham_stage<N, NL, 4, NTHR, SWZ, WAVE, false>(buf, Ns, tw, twstride, tid);
        Ns *= 4;
    }
}

// One Stockham stage of radix R from src to dst (two buffers, any memory:
// transforms too long for LDS run through a workgroup's global scratch).
// Same butterflies, twiddles and order as stockham_stage, so the results
// are those of lds_fft.  Ends with a barrier (the workgroup's writes are
// then visible to all of its threads).
template <int N, int R, int NTHR>
__device__ inline void pingpong_stage(const double2* src, double2* dst, int Ns, const double2* __restrict__ tw,
                                      int twstride)
{
    for (int j = threadIdx.x; j < N / R; j += NTHR)
    {
        double2 v[R];
#pragma unroll
        for (int m = 0; m < R; m++) v[m] = src[j + m * (N / R)];
        const int k = j % Ns;
        const int step = k * twstride * (N / (R * Ns));
        if (Ns > 1)
#pragma unroll
            for (int m = 1; m < R; m++) v[m] = cmul(v[m], tw[m * step]);
        double2 y[R];
        radix_butterfly<R>(v, y);
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int m = 0; m < R; m++) dst[base + m * Ns] = y[m];
    }
    __syncthreads();
}

// Forward FFT of one length-N series between buffers a and b (the stage
// schedule of lds_fft); returns the buffer that holds the result.
template <int N, int NTHR>
__device__ inline double2* pingpong_fft(double2* a, double2* b, const double2* __restrict__ tw, int twstride)
{
    constexpr bool R3 = N % 3 == 0;
    constexpr int P2 = R3 ? N / 3 : N;
    static_assert((P2 & (P2 - 1)) == 0 && N >= 4, "N = 2^a or 3*2^a");
    constexpr int LOG2 = __builtin_ctz(P2);
    int Ns = 1;
    double2 *src = a, *dst = b;
    if (R3)
    {
        pingpong_stage<N, 3, NTHR>(src, dst, Ns, tw, twstride);
        Ns *= 3;
        double2* t = src; src = dst; dst = t;
    }
    if (LOG2 & 1)
    {
        pingpong_stage<N, 2, NTHR>(src, dst, Ns, tw, twstride);
        Ns *= 2;
        double2* t = src; src = dst; dst = t;
    }
#pragma unroll 1
    for (int s = 0; s < LOG2 / 2; s++)
    {
        pingpong_stage<N, 4, NTHR>(src, dst, Ns, tw, twstride);
        Ns *= 4;
        double2* t = src; src = dst; dst = t;
    }
    return src;
}

}  // namespace glass
