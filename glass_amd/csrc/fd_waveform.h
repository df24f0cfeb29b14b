// fd_waveform.h -- device building blocks of the frequency-domain UCB
// template: slow series -> six LDS FFTs -> fftshift/normalise -> TDI.
// ucb/src/glass_ucb_waveform.c:738-1001; TDI utils/src/glass_lisa.c:458-710.
#pragma once
#include "fft_lds.h"
#include "ucb_common.h"

namespace glass {

// Generate the six slow series of one template into LDS (link-major, N
// complex each) and FFT them.  pos: positions table of NPOS samples over
// [t0, t0+T) (9 doubles per sample), sampled at stride NPOS/N.
template <int N, int NTHR>
__device__ inline void fd_slow_fft(const UcbConst& c, double2* lds, const double* __restrict__ pos,
                                   int posstride, const double2* __restrict__ tw, int twstride)
{
    for (int n = threadIdx.x; n < N; n += NTHR)
    {
        double t = ucb_sample_time(c.T, c.t0, n, N);
        double s[12];
        ucb_slow_sample(c, t, pos + (size_t)n * posstride * 9, s);
#pragma unroll
        for (int l = 0; l < 6; l++) lds[l * N + n] = make_double2(s[2 * l], s[2 * l + 1]);
    }
    __syncthreads();
    lds_fft<N, 6, NTHR>(lds, tw, twstride);
}

// TDI of output bin m (template-local) from the FFT'd link series in LDS.
// Link order 12,21,31,13,23,32.  Returns X,Y,Z,A,E as complex values.
// fftshift + 1/(2N): d(m) = FFT[(m + N/2) mod N] / (2N)
// (glass_ucb_waveform.c:948-962).  Trig of f/f* is evaluated directly per
// bin; the reference steps it with a rotation recurrence
// (glass_lisa.c:403-415), which agrees to ~N ulp.
template <int N>
__device__ inline void fd_tdi_bin(const UcbConst& c, const double2* lds, int m, int fmt, double2* XYZAE)
{
    const int s = (m + N / 2) & (N - 1);
    const double inv2N = 1. / (double)(2 * N);
    double2 d12 = lds[0 * N + s], d21 = lds[1 * N + s], d31 = lds[2 * N + s];
    double2 d13 = lds[3 * N + s], d23 = lds[4 * N + s], d32 = lds[5 * N + s];
    d12.x *= inv2N; d12.y *= inv2N; d21.x *= inv2N; d21.y *= inv2N;
    d31.x *= inv2N; d31.y *= inv2N; d13.x *= inv2N; d13.y *= inv2N;
    d23.x *= inv2N; d23.y *= inv2N; d32.x *= inv2N; d32.y *= inv2N;

    const double f = ((double)(c.q + m - N / 2)) / c.T;
    const double fonfs = f * c.invfstar;
    double s1, c1;
    sincos(fonfs, &s1, &c1);
    const double c2 = 2. * c1 * c1 - 1.0, s2 = 2. * s1 * c1;
    const double sqT = sqrt(c.T);
    double2 X, Y, Z, A, E;
    if (fmt == TDI_SANGRIA)
    {
        // glass_lisa.c:645-710
        const double pf = f * (4.0 * c.invfstar * sqT) * s1;
        double ar = d12.x - d13.x, ai = d12.y - d13.y, br = d21.x - d31.x, bi = d21.y - d31.y;
        X = make_double2(pf * (ar * c2 + ai * s2 + br * c1 + bi * s1), pf * (ai * c2 - ar * s2 + bi * c1 - br * s1));
        ar = d23.x - d21.x; ai = d23.y - d21.y; br = d32.x - d12.x; bi = d32.y - d12.y;
        Y = make_double2(pf * (ar * c2 + ai * s2 + br * c1 + bi * s1), pf * (ai * c2 - ar * s2 + bi * c1 - br * s1));
        ar = d31.x - d32.x; ai = d31.y - d32.y; br = d13.x - d23.x; bi = d13.y - d23.y;
        Z = make_double2(pf * (ar * c2 + ai * s2 + br * c1 + bi * s1), pf * (ai * c2 - ar * s2 + bi * c1 - br * s1));
    }
    else
    {
        // glass_lisa.c:458-643 ("phase": prefactor sqrt(T); "frequency": sqrt(T)*2f/f*)
        const double c3 = 2. * c1 * c2 - c1, s3 = 2. * s1 * c2 + s1;
        const double pf = (fmt == TDI_PHASE) ? sqT : sqT * (2. * fonfs);
        auto chan = [&](double2 a, double2 b, double2 e, double2 g) {
            // a = dij - dik (c3), b = dji - dki (c2), -a (c1), -b (const)
            double ar = a.x, ai = a.y, br = b.x, bi = b.y;
            double re = ar * c3 + ai * s3 + br * c2 + bi * s2 - ar * c1 - ai * s1 - br;
            double im = ai * c3 - ar * s3 + bi * c2 - br * s2 - ai * c1 + ar * s1 - bi;
            (void)e; (void)g;
            return make_double2(pf * re, pf * im);
        };
        X = chan(make_double2(d12.x - d13.x, d12.y - d13.y), make_double2(d21.x - d31.x, d21.y - d31.y), d12, d12);
        Y = chan(make_double2(d23.x - d21.x, d23.y - d21.y), make_double2(d32.x - d12.x, d32.y - d12.y), d12, d12);
        Z = chan(make_double2(d31.x - d32.x, d31.y - d32.y), make_double2(d13.x - d23.x, d13.y - d23.y), d12, d12);
    }
    A = make_double2((Z.x - X.x) * kSQRT1_2, (Z.y - X.y) * kSQRT1_2);
    E = make_double2((X.x - 2 * Y.x + Z.x) * kINVSQ6, (X.y - 2 * Y.y + Z.y) * kINVSQ6);
    if (fmt == TDI_PHASE)
    {
        A.y = -A.y;
        E.y = -E.y;
    }
    XYZAE[0] = X;
    XYZAE[1] = Y;
    XYZAE[2] = Z;
    XYZAE[3] = A;
    XYZAE[4] = E;
}

}  // namespace glass
