// fd_waveform.h -- device building blocks of the frequency-domain UCB
// template: slow series -> six LDS FFTs -> fftshift/normalise -> TDI.
// ucb/src/glass_ucb_waveform.c:738-1001; TDI utils/src/glass_lisa.c:458-710.
//
// Work split: one workgroup of NTHR threads per template of N bins; thread
// t owns output bins m = t + k*NTHR (k < PER).
//  * N <= 1024: NTHR = max(N, 64) threads, one slow sample each, written
//    straight into LDS (6 links x N complex = 96 N bytes <= 96 KB), six
//    FFTs, then every thread reads its bin's six link values.
//  * N <= 1024 with NTHR = N/2 (fd_link_bins): thread t folds samples t
//    and t + N/2 by one radix-2 decimation-in-frequency step; the even-bin
//    series goes to LDS, the odd-bin series waits in registers, and the two
//    half-length transforms run in turn (48 N bytes of LDS).  Bins are
//    handed out per pass, one per thread (m = 2t, then 2t + 1), so only
//    one bin's six link values are live at a time.
//  * N = 2048 with NTHR = 512 (fd_link_bins): the split step on two sample
//    pairs per thread, the two 1024-point half-length transforms in turn.
//  * N = P*NS (N >= 2048, NS = 1024): decimation in time.  The samples
//    n = P*j + p of each phase p form an NS-point series; its six FFTs run
//    in LDS in turn and each thread folds X_{j+NS r} += w_N^{p j} w_P^{p r}
//    S_p[j] into registers for its sub-bins j, so LDS stays at 6*NS*16
//    bytes and no sample is held in registers across passes.
#pragma once
#include "fft_lds.h"
#include "ucb_common.h"

#ifndef GH_NTHR_1024
#define GH_NTHR_1024 512
#endif
#ifndef GH_DECIM_NS
#define GH_DECIM_NS 1024
#endif
#ifndef GH_NTHR_DECIM
#define GH_NTHR_DECIM 512
#endif

namespace glass {

// Workgroup size for a template of N bins: one sample per thread up to 512,
// two at N = 1024 (keeps the 256-VGPR budget of 8-wave workgroups), and
// 1024 threads for the decimated N >= 2048 path.
constexpr int fd_threads(int N) { return N < 64 ? 64 : (N <= 512 ? N : (N == 1024 ? GH_NTHR_1024 : GH_NTHR_DECIM)); }

// Default workgroup size of the batched template kernels (fd_link_bins):
// the split shape, N/2 threads with two bins each, from 128 to 1024 bins.
constexpr int fd_batch_threads(int N) { return N <= 64 ? 64 : (N <= 1024 ? N / 2 : GH_NTHR_DECIM); }

// SPLIT (fd_link_bins only, N <= 2048 with NTHR = N/2): one radix-2
// decimation-in-frequency step in registers, then the even- and odd-bin
// half-length transforms in LDS one after the other (half the LDS).
template <int N, int NTHR, bool SPLIT = false>
struct FdShape
{
    static_assert(!SPLIT || (N <= 2048 && N >= 8 && 2 * NTHR == N), "split shape: N <= 2048, N/2 threads");
    static constexpr int PER = (N + NTHR - 1) / NTHR;  // bins per thread
    // FFT length in LDS: the whole template up to 1024 bins (half of it
    // when split), else the sub-series length of the decimated path
    static constexpr int NS = SPLIT ? N / 2 : (N <= 1024 ? N : GH_DECIM_NS);
    static constexpr int LDS = 6 * NS;                  // double2 elements
};

// Split launch shapes: two bins per thread up to 2048 bins.
constexpr bool fd_split(int N, int NTHR) { return N <= 2048 && N >= 8 && 2 * NTHR == N; }

// Move a value that every lane holds identically into a scalar register.
__device__ inline double uniform(double x)
{
    int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

__device__ inline void make_uniform(UcbConst& c)
{
    double* f = &c.T;
    // every double member up to (not including) q
    constexpr int nd = (int)((offsetof(UcbConst, q)) / sizeof(double));
#pragma unroll
    for (int i = 0; i < nd; i++) f[i] = uniform(f[i]);
    c.q = ((long)__builtin_amdgcn_readfirstlane((int)(c.q >> 32)) << 32) |
          (unsigned)__builtin_amdgcn_readfirstlane((int)(c.q & 0xffffffff));
    c.np = __builtin_amdgcn_readfirstlane(c.np);
}

// Per-template constants computed once per workgroup (by thread 0) into
// LDS; every thread reads them from there.  Synchronises the workgroup on
// entry (readers of the previous template are done) and on exit.
__device__ inline const UcbConst& ucb_setup_shared(const double* p, int np, double T, double t0, double L,
                                                   double fstar)
{
    __shared__ UcbConst sc;
    __syncthreads();
    if (threadIdx.x == 0) ucb_setup(p, np, T, t0, L, fstar, sc);
    __syncthreads();
    return sc;
}

// Slow samples of one phase into LDS: sample n = P*j + p at LDS slot j.
template <int N, int NTHR, int P>
__device__ inline void fd_samples_to_lds(const UcbConst& c, double2* lds, const double* __restrict__ pos,
                                         int posstride, int p)
{
    constexpr int NS = N / P;
    for (int j = threadIdx.x; j < NS; j += NTHR)
    {
        const int n = P * j + p;
        double t = ucb_sample_time(c.T, c.t0, n, N);
        ucb_slow_sample(c, t, pos + (size_t)n * posstride * 9, lds, NS, j);
    }
    __syncthreads();
}

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
    const double inv2N = 1. / (double)(2 * N);
    if constexpr (N <= 1024)
    {
        fd_samples_to_lds<N, NTHR, 1>(c, lds, pos, posstride, 0);
        lds_fft<N, 6, NTHR>(lds, tw, twstride);
#pragma unroll
        for (int k = 0; k < S::PER; k++)
        {
            const int m = threadIdx.x + k * NTHR;
            const int s = (m + N / 2) & (N - 1);
            if (m < N)
#pragma unroll
                for (int l = 0; l < 6; l++)
                {
                    double2 v = lds[l * N + s];
                    d[k][l] = make_double2(v.x * inv2N, v.y * inv2N);
                }
        }
        __syncthreads();
    }
    else
    {
        // decimation in time: P phases of NS-point sub-series
        constexpr int NS = S::NS, P = N / NS, U = NS / NTHR;
        static_assert(NS % NTHR == 0 && N % NS == 0, "decimated shape");
        const int t = threadIdx.x;
        double2 X[U][P][6];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int r = 0; r < P; r++)
#pragma unroll
                for (int l = 0; l < 6; l++) X[u][r][l] = make_double2(0.0, 0.0);
#pragma unroll 1
        for (int p = 0; p < P; p++)
        {
            fd_samples_to_lds<N, NTHR, P>(c, lds, opaque(pos), posstride, p);
            tw = opaque(tw);
            lds_fft<NS, 6, NTHR>(lds, tw, twstride * P);
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const int j = t + u * NTHR;
                // w_N^{p j}; the radix-P rotations are w_P^{p r}
                const double2 wj = tw[(p * j) * twstride];
#pragma unroll
                for (int l = 0; l < 6; l++)
                {
                    const double2 v = cmul(lds[l * NS + j], wj);
#pragma unroll
                    for (int r = 0; r < P; r++)
                    {
                        const double2 w = tw[((p * r) % P) * NS * twstride];
                        const double2 q = cmul(v, w);
                        X[u][r][l].x += q.x;
                        X[u][r][l].y += q.y;
                    }
                }
            }
            __syncthreads();
        }
        // FFT index s = j + NS r, j = t + u NTHR, is output bin
        // m = (s + N/2) mod N = t + k NTHR with k as below
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int r = 0; r < P; r++)
            {
                const int k = (u + U * r + N / (2 * NTHR)) % (N / NTHR);
#pragma unroll
                for (int l = 0; l < 6; l++) d[k][l] = make_double2(X[u][r][l].x * inv2N, X[u][r][l].y * inv2N);
            }
    }
}

// f(m, d) for every output bin m of this thread, d[l] = link l at bin m
// as fd_link_spectra leaves it.  Ends with a barrier (TAIL below).  SOA (split shape
// only): pos is the N-sample structure-of-arrays block (9 rows of N
// samples) and posstride is unused; else pos is the record table read with
// stride posstride.
// pre(m) is called for each bin m before f(m, d) of that bin, as early as
// the shape allows: in the split shape at the start of the bin's pass,
// before its transform, so that global loads the caller issues there for
// bin m are in flight during the FFT stages.  TAIL = false (split shape):
// no barrier after the last LDS reads; the caller then passes a barrier of
// its own before anything writes lds again.
template <int N, int NTHR, bool SPLIT, bool SOA = false, bool TAIL = true, typename Pre, typename F>
__device__ inline void fd_link_bins(const UcbConst& c, double2* lds, const double* __restrict__ pos, int posstride,
                                    const double2* __restrict__ tw, int twstride, Pre&& pre, F&& f)
{
    static_assert(SPLIT || !SOA, "structure-of-arrays positions: split shape only");
    // 2048 bins on 512 threads: the split step on two sample pairs per
    // thread (n = t, t + 512 with partners n + 1024), the odd-bin series in
    // registers during the even-bin transform; no decimated accumulators
    // (3% faster than the decimated form, 173 VGPRs instead of ~235)
    constexpr bool SPLIT2 = !SPLIT && N == 2048 && NTHR == 512;
    if constexpr (SPLIT2)
    {
        constexpr int H = N / 2, Q = H / NTHR;
        const double inv2N = 1. / (double)(2 * N);
        const int t = threadIdx.x;
        double2 ya[Q][3], yb[Q][3];
#pragma unroll
        for (int q = 0; q < Q; q++)
        {
            const int n = t + q * NTHR;
            ucb_slow_sample(c, ucb_sample_time(c.T, c.t0, n, N), pos + (size_t)n * posstride * 9, lds, H, n);
            const double2 w = tw[n * twstride];
            ucb_slow_sample_to<false>(c, ucb_sample_time(c.T, c.t0, n + H, N), pos + (size_t)(n + H) * posstride * 9,
                                      [&](int l, double2 x1) {
                                          const double2 x0 = lds[l * H + n];
                                          lds[l * H + n] = make_double2(x0.x + x1.x, x0.y + x1.y);
                                          const double2 dv = cmul(make_double2(x0.x - x1.x, x0.y - x1.y), w);
                                          if (l < 3)
                                          {
                                              ya[q][0] = ya[q][1];
                                              ya[q][1] = ya[q][2];
                                              ya[q][2] = dv;
                                          }
                                          else
                                          {
                                              yb[q][0] = yb[q][1];
                                              yb[q][1] = yb[q][2];
                                              yb[q][2] = dv;
                                          }
                                      });
        }
        __syncthreads();
#pragma unroll
        for (int pass = 0; pass < 2; pass++)
        {
            lds_fft<H, 6, NTHR>(lds, opaque(tw), 2 * twstride);
            double2 d[Q][6];
#pragma unroll
            for (int q = 0; q < Q; q++)
            {
                // bin m = 2n + pass is FFT index (m + H) mod N, of parity pass
                const int m = 2 * (t + q * NTHR) + pass;
                const int sl = ((m + H) & (N - 1)) >> 1;
#pragma unroll
                for (int l = 0; l < 6; l++)
                {
                    const double2 v = lds[l * H + sl];
                    d[q][l] = make_double2(v.x * inv2N, v.y * inv2N);
                }
            }
            if (!pass || TAIL) __syncthreads();
            if (!pass)
#pragma unroll
                for (int q = 0; q < Q; q++)
#pragma unroll
                    for (int l = 0; l < 3; l++)
                    {
                        lds[l * H + t + q * NTHR] = ya[q][l];
                        lds[(l + 3) * H + t + q * NTHR] = yb[q][l];
                    }
#pragma unroll
            for (int q = 0; q < Q; q++)
            {
                const int m = 2 * (t + q * NTHR) + pass;
                pre(m);
                f(m, d[q]);
            }
            if (!pass) __syncthreads();
        }
    }
    else if constexpr (!SPLIT)
    {
        using S = FdShape<N, NTHR>;
        double2 dl[S::PER][6];
        fd_link_spectra<N, NTHR>(c, lds, pos, posstride, tw, twstride, dl);
#pragma unroll
        for (int k = 0; k < S::PER; k++)
        {
            const int m = threadIdx.x + k * NTHR;
            if (m < N)
            {
                pre(m);
                f(m, dl[k]);
            }
        }
    }
    else
    {
        // X[2k] = FFT_H(x[n] + x[n+H]), X[2k+1] = FFT_H((x[n] - x[n+H]) w_N^n),
        // H = N/2 = NTHR
        constexpr int H = N / 2;
        // swizzled LDS layout of the half-length series (conflict-free
        // Stockham writes) at 1024 bins: 4.5% faster there; at 256 and 512
        // bins the extra index arithmetic cancels the gain (measured)
        constexpr bool SW = H == 512;
        const double inv2N = 1. / (double)(2 * N);
        const int t = threadIdx.x;
        const int ts = lds_slot<SW>(t);
        constexpr int KS = SOA ? N : 1;
        const double* p0 = SOA ? pos + t : pos + (size_t)t * posstride * 9;
        const double* p1 = SOA ? pos + t + H : pos + (size_t)(t + H) * posstride * 9;
        ucb_slow_sample<KS>(c, ucb_sample_time(c.T, c.t0, t, N), p0, lds, H, ts);
        // links 0-2 and 3-5 of the odd-bin series, shifted in as the
        // spacecraft loop produces them (l < 3 is known per unrolled link)
        const double2 w = tw[t * twstride];
        double2 ya[3], yb[3];
        ucb_slow_sample_to<false, KS>(c, ucb_sample_time(c.T, c.t0, t + H, N), p1,
                                  [&](int l, double2 x1) {
                                      const double2 x0 = lds[l * H + ts];
                                      lds[l * H + ts] = make_double2(x0.x + x1.x, x0.y + x1.y);
                                      const double2 dv = cmul(make_double2(x0.x - x1.x, x0.y - x1.y), w);
                                      if (l < 3)
                                      {
                                          ya[0] = ya[1];
                                          ya[1] = ya[2];
                                          ya[2] = dv;
                                      }
                                      else
                                      {
                                          yb[0] = yb[1];
                                          yb[1] = yb[2];
                                          yb[2] = dv;
                                      }
                                  });
        __syncthreads();
#pragma unroll
        for (int pass = 0; pass < 2; pass++)
        {
            // bin m = 2t + pass is FFT index (m + H) mod N, of parity pass
            const int m = 2 * t + pass;
            pre(m);
            lds_fft<H, 6, NTHR, SW>(lds, opaque(tw), 2 * twstride);
            const int s = lds_slot<SW>(((m + H) & (N - 1)) >> 1);
            double2 d[6];
#pragma unroll
            for (int l = 0; l < 6; l++)
            {
                const double2 v = lds[l * H + s];
                d[l] = make_double2(v.x * inv2N, v.y * inv2N);
            }
            if (!pass || TAIL) __syncthreads();
            if (!pass)
            {
                // the odd-bin series takes the LDS before the even bins'
                // work, so it is not held in registers across that work
#pragma unroll
                for (int l = 0; l < 3; l++)
                {
                    lds[l * H + ts] = ya[l];
                    lds[(l + 3) * H + ts] = yb[l];
                }
            }
            f(m, d);
            if (!pass) __syncthreads();
        }
    }
}

template <int N, int NTHR, bool SPLIT, bool SOA = false, typename F>
__device__ inline void fd_link_bins(const UcbConst& c, double2* lds, const double* __restrict__ pos, int posstride,
                                    const double2* __restrict__ tw, int twstride, F&& f)
{
    fd_link_bins<N, NTHR, SPLIT, SOA>(c, lds, pos, posstride, tw, twstride, [](int) {}, f);
}

// TDI of template bin m from its six link values d (link order 12, 21, 31,
// 13, 23, 32).  out: X, Y, Z, A, E.  Trig of f/f* is evaluated directly per
// bin; the reference steps it with a rotation recurrence
// (glass_lisa.c:403-415), which agrees to ~N ulp.
template <int N>
__device__ inline void fd_tdi(const UcbConst& c, int m, int fmt, const double2* d, double2* out)
{
    const double2 d12 = d[0], d21 = d[1], d31 = d[2], d13 = d[3], d23 = d[4], d32 = d[5];
    // f = (q + m - N/2)/T (glass_lisa.c:664); amplitude path, so the
    // reciprocal is folded in
    const double f = ((double)(c.q + m - N / 2)) * c.invT;
    const double fonfs = f * c.invfstar;
    double s1, c1;
    sincos_red(fonfs, &s1, &c1);
    const double c2 = 2. * c1 * c1 - 1.0, s2 = 2. * s1 * c1;
    const double sqT = c.sqT;
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
