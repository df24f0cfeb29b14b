// fd_waveform.h -- device building blocks of the frequency-domain UCB
// template: slow series -> six LDS FFTs -> fftshift/normalise -> TDI.
// ucb/src/glass_ucb_waveform.c:738-1001; TDI utils/src/glass_lisa.c:458-710.
//
// Work split: one workgroup of NTHR threads per template of N bins.  Thread
// t generates slow samples n = t + k*NTHR and afterwards owns output bins
// m = t + k*NTHR (k < PER = N/NTHR).  The six link series go through LDS in
// passes of NLP links (6 when 96*N bytes fit in 96 KB of the 160 KB LDS,
// i.e. N <= 1024; 3 at N = 2048; 2 at N = 4096); after each pass's FFT every thread
// copies its bins' link values into registers, so the TDI combination reads
// registers only.
#pragma once
#include "fft_lds.h"
#include "ucb_common.h"

namespace glass {

template <int N, int NTHR>
struct FdShape
{
    static constexpr int PER = (N + NTHR - 1) / NTHR;  // samples (and bins) per thread
    // links per LDS pass: all six while they fit in 96 KB, else 3, 2, 1
    static constexpr int NLP = (6 * N * 16 <= 98304) ? 6 : (3 * N * 16 <= 98304) ? 3 : (2 * N * 16 <= 131072) ? 2 : 1;
    static constexpr int LDS = NLP * N;                          // double2 elements
};

// Build the six shifted, normalised link spectra of one template: on return
// d[k][l] holds link l at output bin m = threadIdx.x + k*NTHR, i.e.
// FFT_l[(m + N/2) mod N] / (2N)  (glass_ucb_waveform.c:940-962).
// Link order 12, 21, 31, 13, 23, 32.  Ends with a barrier, so lds may be
// reused immediately.
template <int N, int NTHR>
__device__ inline void fd_link_spectra(const UcbConst& c, double2* lds, const double* __restrict__ pos,
                                       int posstride, const double2* __restrict__ tw, int twstride,
                                       double2 (&d)[FdShape<N, NTHR>::PER][6])
{
    using S = FdShape<N, NTHR>;
    double2 sv[S::PER][6];
#pragma unroll
    for (int k = 0; k < S::PER; k++)
    {
        const int n = threadIdx.x + k * NTHR;
        if (n < N)
        {
            double t = ucb_sample_time(c.T, c.t0, n, N);
            double s[12];
            ucb_slow_sample(c, t, pos + (size_t)n * posstride * 9, s);
#pragma unroll
            for (int l = 0; l < 6; l++) sv[k][l] = make_double2(s[2 * l], s[2 * l + 1]);
        }
    }
    const double inv2N = 1. / (double)(2 * N);
#pragma unroll
    for (int pass = 0; pass < 6 / S::NLP; pass++)
    {
#pragma unroll
        for (int k = 0; k < S::PER; k++)
        {
            const int n = threadIdx.x + k * NTHR;
            if (n < N)
#pragma unroll
                for (int l = 0; l < S::NLP; l++) lds[l * N + n] = sv[k][pass * S::NLP + l];
        }
        __syncthreads();
        lds_fft<N, S::NLP, NTHR>(lds, tw, twstride);
#pragma unroll
        for (int k = 0; k < S::PER; k++)
        {
            const int m = threadIdx.x + k * NTHR;
            const int s = (m + N / 2) & (N - 1);
            if (m < N)
#pragma unroll
                for (int l = 0; l < S::NLP; l++)
                {
                    double2 v = lds[l * N + s];
                    d[k][pass * S::NLP + l] = make_double2(v.x * inv2N, v.y * inv2N);
                }
        }
        __syncthreads();
    }
}

// TDI of template bin m from its six link values d (link order 12, 21, 31,
// 13, 23, 32).  out: X, Y, Z, A, E.  Trig of f/f* is evaluated directly per
// bin; the reference steps it with a rotation recurrence
// (glass_lisa.c:403-415), which agrees to ~N ulp.
template <int N>
__device__ inline void fd_tdi(const UcbConst& c, int m, int fmt, const double2* d, double2* out)
{
    const double2 d12 = d[0], d21 = d[1], d31 = d[2], d13 = d[3], d23 = d[4], d32 = d[5];
    const double f = ((double)(c.q + m - N / 2)) / c.T;
    const double fonfs = f * c.invfstar;
    double s1, c1;
    sincos(fonfs, &s1, &c1);
    const double c2 = 2. * c1 * c1 - 1.0, s2 = 2. * s1 * c1;
    const double sqT = sqrt(c.T);
    // link combinations a = d_ij - d_ik, b = d_ji - d_ki per channel
    const double2 aX = make_double2(d12.x - d13.x, d12.y - d13.y), bX = make_double2(d21.x - d31.x, d21.y - d31.y);
    const double2 aY = make_double2(d23.x - d21.x, d23.y - d21.y), bY = make_double2(d32.x - d12.x, d32.y - d12.y);
    const double2 aZ = make_double2(d31.x - d32.x, d31.y - d32.y), bZ = make_double2(d13.x - d23.x, d13.y - d23.y);
    double2 X, Y, Z;
    if (fmt == TDI_SANGRIA)
    {
        // glass_lisa.c:645-710: pf [a e^{-2ix} + b e^{-ix}], pf = 4 sqrt(T) (f/f*) sin(f/f*)
        const double pf = f * (4.0 * c.invfstar * sqT) * s1;
        auto ch = [&](double2 a, double2 b) {
            return make_double2(pf * (a.x * c2 + a.y * s2 + b.x * c1 + b.y * s1),
                                pf * (a.y * c2 - a.x * s2 + b.y * c1 - b.x * s1));
        };
        X = ch(aX, bX);
        Y = ch(aY, bY);
        Z = ch(aZ, bZ);
    }
    else
    {
        // glass_lisa.c:458-643: pf [a e^{-3ix} + b e^{-2ix} - a e^{-ix} - b],
        // pf = sqrt(T) ("phase") or sqrt(T) 2 f/f* ("frequency")
        const double c3 = 2. * c1 * c2 - c1, s3 = 2. * s1 * c2 + s1;
        const double pf = (fmt == TDI_PHASE) ? sqT : sqT * (2. * fonfs);
        auto ch = [&](double2 a, double2 b) {
            return make_double2(pf * (a.x * c3 + a.y * s3 + b.x * c2 + b.y * s2 - a.x * c1 - a.y * s1 - b.x),
                                pf * (a.y * c3 - a.x * s3 + b.y * c2 - b.x * s2 - a.y * c1 + a.x * s1 - b.y));
        };
        X = ch(aX, bX);
        Y = ch(aY, bY);
        Z = ch(aZ, bZ);
    }
    double2 A = make_double2((Z.x - X.x) * kSQRT1_2, (Z.y - X.y) * kSQRT1_2);
    double2 E = make_double2((X.x - 2 * Y.x + Z.x) * kINVSQ6, (X.y - 2 * Y.y + Z.y) * kINVSQ6);
    if (fmt == TDI_PHASE)
    {
        A.y = -A.y;
        E.y = -E.y;
    }
    out[0] = X;
    out[1] = Y;
    out[2] = Z;
    out[3] = A;
    out[4] = E;
}

}  // namespace glass
