// fd_derivs.hip -- exact gradient and Fisher matrix of one galactic binary's
// likelihood contribution, for many sources at once (no reference
// counterpart: the reference's ucb_fisher, ucb/src/glass_ucb_waveform.c:
// 364-522, uses forward differences of 8 extra waveforms).
//
// For a source with template h(theta) over a residual R that excludes it,
// logL = const + L(theta),  L = (R|h) - (h|h)/2,  so
//   dL/dtheta_k = (R - h | dh/dtheta_k),   Gamma_kl = (dh_k | dh_l),
// with ( | ) the 3-channel noise-weighted product of gaussian_log_likelihood
// (utils/src/glass_model.c:758-767).  One workgroup per source: the six
// series sets of fd_tangent.h go through LDS in as few passes as 96 KB
// allows, each pass FFT'd and TDI'd into per-bin registers; then every
// thread accumulates L, the 8 gradient entries and the 36 distinct Fisher
// entries of its bins and the workgroup reduces the 45 sums.
#include <cmath>
#include <type_traits>

#include "../../include/glass_hip_batch.h"
#include "fd_derivs.h"
#include "fd_tangent.h"
#include "reduce.h"
#include "runtime.h"

namespace glass {

int parse_format(const char* format);

constexpr int kNSum = 45;  // L, grad[8], Fisher upper triangle[36]

// B(u,v) = 2 sum_ab Re(u_a conj v_b) w_ab with w = XX,YY,ZZ,XY,XZ,YZ
__device__ inline double bf3(const double2* u, const double2* v, const double* w)
{
    auto re = [](double2 a, double2 b) { return a.x * b.x + a.y * b.y; };
    return 2.0 * (w[0] * re(u[0], v[0]) + w[1] * re(u[1], v[1]) + w[2] * re(u[2], v[2]) +
                  w[3] * (re(u[0], v[1]) + re(u[1], v[0])) + w[4] * (re(u[0], v[2]) + re(u[2], v[0])) +
                  w[5] * (re(u[1], v[2]) + re(u[2], v[1])));
}

template <int NTHR>
__device__ inline void block_sum_vec(double (&v)[kNSum], double* out)
{
    constexpr int NW = (NTHR + 63) / 64;
    __shared__ double part[NW][kNSum];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // the 45 wave sums by transposition: lane k ends with value k's sum
    constexpr int GROUP = 64 / pow2_at_least(kNSum);  // lanes holding one sum
    const double x = wave_sum_transpose<kNSum>(v);
    const int k = lane / GROUP;
    if (k < kNSum && lane % GROUP == 0) part[w][k] = x;
    __syncthreads();
    if (threadIdx.x < kNSum)
    {
        double s = 0.0;
        for (int i = 0; i < NW; i++) s += part[i][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// LDS bytes for the series sets of one pass: sets per pass = this / 96 N
// (at least one).  48 KB up to 256 bins (two or more workgroups per CU at
// ~250 VGPRs), 96 KB above (one workgroup per CU either way, fewer passes).
constexpr int deriv_lds_budget(int N) { return N <= 256 ? 49152 : 98304; }

// One pass of series sets [S0, S0+NS): samples -> LDS -> FFT -> TDI into H.
template <int N, int NTHR, int S0, int NS>
__device__ inline void deriv_pass(const UcbConst& c, const UcbDiff& dd, double2* lds, const DerivArgs& A,
                                  double2 (&H)[FdShape<N, NTHR>::PER][6][3])
{
    using S = FdShape<N, NTHR>;
    for (int j = threadIdx.x; j < N; j += NTHR)
    {
        const double t = ucb_sample_time(c.T, c.t0, j, N);
        ucb_sample_sets<S0, NS>(c, dd, t, A.pos + (size_t)j * (4096 / N) * 9, lds, N, j);
    }
    __syncthreads();
    lds_fft<N, 6 * NS, NTHR>(lds, A.tw, kTwN / N);
    const double inv2N = 1. / (double)(2 * N);
#pragma unroll
    for (int k = 0; k < S::PER; k++)
    {
        const int m = threadIdx.x + k * NTHR;
        if (m >= N) continue;
        const int s = (m + N / 2) & (N - 1);
#pragma unroll
        for (int q = 0; q < NS; q++)
        {
            double2 d6[6], o5[5];
#pragma unroll
            for (int l = 0; l < 6; l++)
            {
                double2 v = lds[(q * 6 + l) * N + s];
                d6[l] = make_double2(v.x * inv2N, v.y * inv2N);
            }
            fd_tdi<N>(c, m, A.fmt, d6, o5);
#pragma unroll
            for (int ch = 0; ch < 3; ch++) H[k][S0 + q][ch] = o5[ch];
        }
    }
    __syncthreads();
}

// Passes over the sets below S1, the highest first: the tangent sets (2-5,
// the costly samples) run while few TDI values of earlier passes are held in
// registers, the basis sets (0, 1) last.
template <int N, int NTHR, int S1>
__device__ inline void deriv_passes(const UcbConst& c, const UcbDiff& dd, double2* lds, const DerivArgs& A,
                                    double2 (&H)[FdShape<N, NTHR>::PER][6][3])
{
    constexpr int CAP = (deriv_lds_budget(N) / (96 * N)) < 1 ? 1 : (deriv_lds_budget(N) / (96 * N));
    constexpr int NS = S1 < CAP ? S1 : CAP;
    deriv_pass<N, NTHR, S1 - NS, NS>(c, dd, lds, A, H);
    if constexpr (S1 - NS > 0) deriv_passes<N, NTHR, S1 - NS>(c, dd, lds, A, H);
}

// Waves per SIMD the register allocation must allow (the launch bound's
// second argument; 1 = the compiler's choice).  Only 256 bins can use more:
// their 50 KB of LDS fits three workgroups per CU, the larger and smaller
// sizes are held to two waves per SIMD by LDS.  Three waves cap the kernel
// at 168 VGPRs (from 195, with 144 B of spill) and run the 256-bin Fisher
// batch about 2% faster.
#ifndef GH_DERIV_W256
#define GH_DERIV_W256 3
#endif
constexpr int deriv_min_waves(int N) { return N == 256 ? GH_DERIV_W256 : 1; }

template <int N, int NTHR>
__global__ __launch_bounds__(NTHR, deriv_min_waves(N)) void deriv_kernel(DerivArgs A)
{
    using S = FdShape<N, NTHR>;
    constexpr int CAP = (deriv_lds_budget(N) / (96 * N)) < 1 ? 1 : (deriv_lds_budget(N) / (96 * N));
    constexpr int NSMAX = CAP < 6 ? CAP : 6;
    __shared__ double2 lds[6 * NSMAX * N];
    __shared__ UcbDiff sdd;
    __shared__ double red[kNSum];
    int e = blockIdx.x;
    if (A.list)
    {
        if (e >= *A.count) return;  // uniform over the workgroup
        e = A.list[e];
    }
    const double* p = A.params + (size_t)(A.pidx ? A.pidx[e] : e) * 8;
    // the template constants and their derivatives set up at once on two
    // waves (one workgroup per source: nothing of a previous source to wait for)
    __shared__ UcbConst sc;
    if (threadIdx.x == 0) ucb_setup(p, 8, A.T, A.t0, A.L, A.fstar, sc);
    if (threadIdx.x == (NTHR >= 128 ? 64 : 0)) ucb_diff_setup(p, A.T, sdd);
    __syncthreads();
    // the constants into scalar registers for the passes (2% faster than
    // reading them from LDS; the derivative constants stay in LDS, where
    // the scalar copies would spill)
    UcbConst cr = sc;
    make_uniform(cr);
    const UcbConst& c = cr;
    const UcbDiff& dd = sdd;
    double2 H[S::PER][6][3];
    deriv_passes<N, NTHR, 6>(c, dd, lds, A, H);

    const double2* R = A.iroff ? A.resid + A.iroff[e] : A.resid + (size_t)e * (A.rstride / 2);
    const double* iC = A.iioff ? A.invC + A.iioff[e] : A.invC + (size_t)e * A.istride;
    const int W = A.iW ? A.iW[e] : A.W, imin = A.imin[e];
    const double2* Ha = A.tadd ? A.tadd + A.taddoff[e] : nullptr;
    const int hcap = A.tadd ? A.taddcap[e] : 0;
    const double2 DP = make_double2(c.DPr, c.DPi), DC = make_double2(c.DCr, c.DCi);
    const double2 DP4 = make_double2(dd.DPr4, dd.DPi4), DC4 = make_double2(dd.DCr4, dd.DCi4);
    const double2 DP5 = make_double2(dd.DPr5, dd.DPi5), DC5 = make_double2(dd.DCr5, dd.DCi5);
    double acc[kNSum];
#pragma unroll
    for (int k = 0; k < kNSum; k++) acc[k] = 0.0;
#pragma unroll
    for (int k = 0; k < S::PER; k++)
    {
        const int m = threadIdx.x + k * NTHR;
        if (m >= N) continue;
        const int j = imin + m;
        const int wi = A.mode == 1 ? m : j;  // weight index
        if (wi < 0 || wi >= W) continue;
        double w[6];
#pragma unroll
        for (int q = 0; q < 6; q++) w[q] = iC[(size_t)q * W + wi];
        double2 h[3], dh[8][3], r[3];
#pragma unroll
        for (int ch = 0; ch < 3; ch++)
        {
            const double2 hp = H[k][0][ch], hc = H[k][1][ch];
            h[ch] = make_double2(cmulc(DP, hp).x + cmulc(DC, hc).x, cmulc(DP, hp).y + cmulc(DC, hc).y);
            dh[0][ch] = H[k][2][ch];
            dh[1][ch] = H[k][3][ch];
            dh[2][ch] = H[k][4][ch];
            dh[3][ch] = h[ch];                                  // d/d ln A
            dh[4][ch] = make_double2(cmulc(DP4, hp).x + cmulc(DC4, hc).x, cmulc(DP4, hp).y + cmulc(DC4, hc).y);
            dh[5][ch] = make_double2(cmulc(DP5, hp).x + cmulc(DC5, hc).x, cmulc(DP5, hp).y + cmulc(DC5, hc).y);
            dh[6][ch] = make_double2(h[ch].y, -h[ch].x);        // d/d phi0 = -i h
            dh[7][ch] = H[k][5][ch];
            const bool inwin = j >= 0 && j < W;
            r[ch] = (A.resid && inwin) ? R[(size_t)ch * W + j] : make_double2(0.0, 0.0);
            if (Ha && inwin)
            {
                // the residual of the whole model plus this source's stored
                // template: the residual without the source
                r[ch].x += Ha[(size_t)ch * hcap + m].x;
                r[ch].y += Ha[(size_t)ch * hcap + m].y;
            }
        }
        // likelihood terms only where the template lands in the window
        if (A.mode == 0 || (j >= 0 && j < W))
        {
            double wj[6];
            if (A.mode == 1)
            {
                for (int q = 0; q < 6; q++) wj[q] = (j >= 0 && j < W) ? iC[(size_t)q * W + j] : 0.0;
            }
            const double* ww = A.mode == 1 ? wj : w;
            acc[0] += bf3(r, h, ww) - 0.5 * bf3(h, h, ww);
            double2 rh[3];
#pragma unroll
            for (int ch = 0; ch < 3; ch++) rh[ch] = make_double2(r[ch].x - h[ch].x, r[ch].y - h[ch].y);
#pragma unroll
            for (int a = 0; a < 8; a++) acc[1 + a] += bf3(rh, dh[a], ww);
        }
        int idx = 9;
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int b = a; b < 8; b++) acc[idx++] += bf3(dh[a], dh[b], w);
    }
    block_sum_vec<NTHR>(acc, red);
    if (threadIdx.x < 64)
    {
        const int a = threadIdx.x / 8, b = threadIdx.x % 8;
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        const int idx = 9 + lo * 8 - lo * (lo - 1) / 2 + (hi - lo);
        A.fisher[(size_t)e * 64 + threadIdx.x] = red[idx];
        if (threadIdx.x < 8) A.grad[(size_t)e * 8 + threadIdx.x] = red[1 + threadIdx.x];
        if (threadIdx.x == 0) A.outL[e] = red[0];
    }
}

// The window and residual of item e.
struct DerivItem
{
    const double2* R;   // residual, 3 x W complex
    const double* iC;   // invC, 6 x W
    const double2* Ha;  // stored template added back (or NULL)
    int W, imin, hcap;
};

__device__ __forceinline__ DerivItem deriv_item(const DerivArgs& A, int e)
{
    DerivItem it;
    it.R = A.iroff ? A.resid + A.iroff[e] : A.resid + (size_t)e * (A.rstride / 2);
    it.iC = A.iioff ? A.invC + A.iioff[e] : A.invC + (size_t)e * A.istride;
    it.W = A.iW ? A.iW[e] : A.W;
    it.imin = A.imin[e];
    it.Ha = A.tadd ? A.tadd + A.taddoff[e] : nullptr;
    it.hcap = A.tadd ? A.taddcap[e] : 0;
    return it;
}

// Adds template bin m's terms to the 45 sums from its six series-set TDI
// values Hb[set][channel]: h and the eight dh/dtheta (sets 0,1 the plus /
// cross basis, 2..5 the tangent sets), L, the gradient and the Fisher
// upper triangle.
__device__ __forceinline__ void deriv_bin(const UcbConst& c, const UcbDiff& dd, const DerivArgs& A, const DerivItem& it, int m,
                                 const double2 (&Hb)[6][3], double (&acc)[kNSum])
{
    const int W = it.W;
    const int j = it.imin + m;
    const int wi = A.mode == 1 ? m : j;  // weight index
    if (wi < 0 || wi >= W) return;
    const double2 DP = make_double2(c.DPr, c.DPi), DC = make_double2(c.DCr, c.DCi);
    const double2 DP4 = make_double2(dd.DPr4, dd.DPi4), DC4 = make_double2(dd.DCr4, dd.DCi4);
    const double2 DP5 = make_double2(dd.DPr5, dd.DPi5), DC5 = make_double2(dd.DCr5, dd.DCi5);
    double w[6];
#pragma unroll
    for (int q = 0; q < 6; q++) w[q] = it.iC[(size_t)q * W + wi];
    double2 h[3], dh[8][3], r[3];
#pragma unroll
    for (int ch = 0; ch < 3; ch++)
    {
        const double2 hp = Hb[0][ch], hc = Hb[1][ch];
        h[ch] = make_double2(cmulc(DP, hp).x + cmulc(DC, hc).x, cmulc(DP, hp).y + cmulc(DC, hc).y);
        dh[0][ch] = Hb[2][ch];
        dh[1][ch] = Hb[3][ch];
        dh[2][ch] = Hb[4][ch];
        dh[3][ch] = h[ch];                                  // d/d ln A
        dh[4][ch] = make_double2(cmulc(DP4, hp).x + cmulc(DC4, hc).x, cmulc(DP4, hp).y + cmulc(DC4, hc).y);
        dh[5][ch] = make_double2(cmulc(DP5, hp).x + cmulc(DC5, hc).x, cmulc(DP5, hp).y + cmulc(DC5, hc).y);
        dh[6][ch] = make_double2(h[ch].y, -h[ch].x);        // d/d phi0 = -i h
        dh[7][ch] = Hb[5][ch];
        const bool inwin = j >= 0 && j < W;
        r[ch] = (A.resid && inwin) ? it.R[(size_t)ch * W + j] : make_double2(0.0, 0.0);
        if (it.Ha && inwin)
        {
            // the residual of the whole model plus this source's stored
            // template: the residual without the source
            r[ch].x += it.Ha[(size_t)ch * it.hcap + m].x;
            r[ch].y += it.Ha[(size_t)ch * it.hcap + m].y;
        }
    }
    // likelihood terms only where the template lands in the window
    if (A.mode == 0 || (j >= 0 && j < W))
    {
        double wj[6];
        if (A.mode == 1)
        {
            for (int q = 0; q < 6; q++) wj[q] = (j >= 0 && j < W) ? it.iC[(size_t)q * W + j] : 0.0;
        }
        const double* ww = A.mode == 1 ? wj : w;
        acc[0] += bf3(r, h, ww) - 0.5 * bf3(h, h, ww);
        double2 rh[3];
#pragma unroll
        for (int ch = 0; ch < 3; ch++) rh[ch] = make_double2(r[ch].x - h[ch].x, r[ch].y - h[ch].y);
#pragma unroll
        for (int a = 0; a < 8; a++) acc[1 + a] += bf3(rh, dh[a], ww);
    }
    int idx = 9;
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
        for (int b = a; b < 8; b++) acc[idx++] += bf3(dh[a], dh[b], w);
}

// Workgroup reduction of the 45 sums and the outputs of item e.
template <int NTHR>
__device__ __forceinline__ void deriv_write(const DerivArgs& A, int e, double (&acc)[kNSum], double* red)
{
    block_sum_vec<NTHR>(acc, red);
    if (threadIdx.x < 64)
    {
        const int a = threadIdx.x / 8, b = threadIdx.x % 8;
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        const int idx = 9 + lo * 8 - lo * (lo - 1) / 2 + (hi - lo);
        A.fisher[(size_t)e * 64 + threadIdx.x] = red[idx];
        if (threadIdx.x < 8) A.grad[(size_t)e * 8 + threadIdx.x] = red[1 + threadIdx.x];
        if (threadIdx.x == 0) A.outL[e] = red[0];
    }
    __syncthreads();
}

// N = 2048: the six link series of one set take 196 KB, more than a CU's
// LDS, so each set is built by decimation in time as in fd_link_spectra's
// N >= 2048 branch (two 1024-point sub-series, six 1024-point link FFTs in
// 96 KB of LDS, folded in registers), one set per pass; each set's TDI
// values go to the workgroup's slice of hs (6 sets x 3 channels x N
// complex) and are read back bin by bin for the sums.  Grid-striding over
// the class list (A.list / A.count required).
template <int NTHR>
__global__ __launch_bounds__(NTHR) void deriv_kernel_2048(DerivArgs A, double2* __restrict__ hs)
{
    constexpr int N = 2048, NS = 1024, P = N / NS, U = NS / NTHR, PER = N / NTHR;
    static_assert(NS % NTHR == 0, "decimated shape");
    __shared__ double2 lds[6 * NS];
    __shared__ UcbDiff sdd;
    __shared__ double red[kNSum];
    double2* H = hs + (size_t)blockIdx.x * 18 * N;
    const int cnt = A.list ? *A.count : A.nitems;
    const int t = threadIdx.x;
    const int twstride = kTwN / N;
    const double inv2N = 1. / (double)(2 * N);
    for (int w = blockIdx.x; w < cnt; w += gridDim.x)
    {
        const int e = A.list ? A.list[w] : w;
        const double* p = A.params + (size_t)(A.pidx ? A.pidx[e] : e) * 8;
        const UcbConst& c = ucb_setup_shared(p, 8, A.T, A.t0, A.L, A.fstar);
        if (t == 0) ucb_diff_setup(p, A.T, sdd);
        __syncthreads();
        const UcbDiff& dd = sdd;
        auto one_set = [&](auto set_tag) {
            constexpr int S = decltype(set_tag)::value;
            double2 X[U][P][6];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int r = 0; r < P; r++)
#pragma unroll
                    for (int l = 0; l < 6; l++) X[u][r][l] = make_double2(0.0, 0.0);
#pragma unroll 1
            for (int ph = 0; ph < P; ph++)
            {
                for (int j = t; j < NS; j += NTHR)
                {
                    const int n = P * j + ph;
                    const double ts = ucb_sample_time(c.T, c.t0, n, N);
                    ucb_sample_sets<S, 1>(c, dd, ts, A.pos + (size_t)n * (4096 / N) * 9, lds, NS, j);
                }
                __syncthreads();
                lds_fft<NS, 6, NTHR>(lds, A.tw, twstride * P);
#pragma unroll
                for (int u = 0; u < U; u++)
                {
                    const int j = t + u * NTHR;
                    const double2 wj = A.tw[(ph * j) * twstride];
#pragma unroll
                    for (int l = 0; l < 6; l++)
                    {
                        const double2 v = cmul(lds[l * NS + j], wj);
#pragma unroll
                        for (int r = 0; r < P; r++)
                        {
                            const double2 q = cmul(v, A.tw[((ph * r) % P) * NS * twstride]);
                            X[u][r][l].x += q.x;
                            X[u][r][l].y += q.y;
                        }
                    }
                }
                __syncthreads();
            }
            // FFT index s = j + NS r (j = t + u NTHR) is template bin
            // m = (s + N/2) mod N = t + k NTHR
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int r = 0; r < P; r++)
                {
                    const int k = (u + U * r + N / (2 * NTHR)) % PER;
                    const int m = t + k * NTHR;
                    double2 d6[6], o5[5];
#pragma unroll
                    for (int l = 0; l < 6; l++) d6[l] = make_double2(X[u][r][l].x * inv2N, X[u][r][l].y * inv2N);
                    fd_tdi<N>(c, m, A.fmt, d6, o5);
#pragma unroll
                    for (int ch = 0; ch < 3; ch++) H[(size_t)(S * 3 + ch) * N + m] = o5[ch];
                }
        };
        one_set(std::integral_constant<int, 0>{});
        one_set(std::integral_constant<int, 1>{});
        one_set(std::integral_constant<int, 2>{});
        one_set(std::integral_constant<int, 3>{});
        one_set(std::integral_constant<int, 4>{});
        one_set(std::integral_constant<int, 5>{});
        // each thread reads back only the bins it wrote: no barrier needed
        DerivItem it = deriv_item(A, e);
        double acc[kNSum];
#pragma unroll
        for (int q = 0; q < kNSum; q++) acc[q] = 0.0;
#pragma unroll 1
        for (int k = 0; k < PER; k++)
        {
            const int m = t + k * NTHR;
            double2 Hb[6][3];
#pragma unroll
            for (int S = 0; S < 6; S++)
#pragma unroll
                for (int ch = 0; ch < 3; ch++) Hb[S][ch] = H[(size_t)(S * 3 + ch) * N + m];
            deriv_bin(c, dd, A, it, m, Hb, acc);
        }
        deriv_write<NTHR>(A, e, acc, red);
    }
}

int launch_derivs_2048(const DerivArgs& A, int grid, double2* hs, hipStream_t s)
{
    if (grid <= 0) return GLASS_HIP_OK;
    if ((!A.list && A.nitems <= 0) || (A.list && !A.count) || !hs)
    {
        set_error("fd derivs: BW 2048 needs a class list or an item count, and scratch");
        return GLASS_HIP_EINVAL;
    }
    deriv_kernel_2048<kDeriv2048Threads><<<grid, kDeriv2048Threads, 0, s>>>(A, hs);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

// N = 4096: the 2048 kernel's decimation in time with P = 4 phases of
// 1024-point sub-series, but the P x 6 link accumulators of a thread's
// sub-bins (four times the 2048 kernel's, past any register budget) live in
// the workgroup's global scratch slice: acc[l][s] (6 x N complex), summed
// over the phases in order, then TDI'd into H[set][ch][m] (18 x N).  Each
// thread reads back only the accumulators and bins it wrote (FFT index
// s = j + NS r and bin m = (s + N/2) mod N keep j mod NTHR), so the slice
// needs no barrier.
template <int NTHR>
__global__ __launch_bounds__(NTHR) void deriv_kernel_4096(DerivArgs A, double2* __restrict__ hs)
{
    constexpr int N = 4096, NS = 1024, P = N / NS, U = NS / NTHR, PER = N / NTHR;
    static_assert(NS % NTHR == 0 && (N / 2) % NTHR == 0, "decimated shape");
    __shared__ double2 lds[6 * NS];
    __shared__ UcbDiff sdd;
    __shared__ double red[kNSum];
    double2* Acc = hs + (size_t)blockIdx.x * kDeriv4096Scratch;
    double2* H = Acc + 6 * N;
    const int cnt = A.list ? *A.count : A.nitems;
    const int t = threadIdx.x;
    const int twstride = kTwN / N;
    const double inv2N = 1. / (double)(2 * N);
    for (int w = blockIdx.x; w < cnt; w += gridDim.x)
    {
        const int e = A.list ? A.list[w] : w;
        const double* p = A.params + (size_t)(A.pidx ? A.pidx[e] : e) * 8;
        const UcbConst& c = ucb_setup_shared(p, 8, A.T, A.t0, A.L, A.fstar);
        if (t == 0) ucb_diff_setup(p, A.T, sdd);
        __syncthreads();
        const UcbDiff& dd = sdd;
        auto one_set = [&](auto set_tag) {
            constexpr int S = decltype(set_tag)::value;
#pragma unroll 1
            for (int ph = 0; ph < P; ph++)
            {
                for (int j = t; j < NS; j += NTHR)
                {
                    const int n = P * j + ph;
                    const double ts = ucb_sample_time(c.T, c.t0, n, N);
                    ucb_sample_sets<S, 1>(c, dd, ts, A.pos + (size_t)n * (4096 / N) * 9, lds, NS, j);
                }
                __syncthreads();
                lds_fft<NS, 6, NTHR>(lds, A.tw, twstride * P);
#pragma unroll
                for (int u = 0; u < U; u++)
                {
                    const int j = t + u * NTHR;
                    const double2 wj = A.tw[(ph * j) * twstride];
#pragma unroll
                    for (int l = 0; l < 6; l++)
                    {
                        const double2 v = cmul(lds[l * NS + j], wj);
#pragma unroll
                        for (int r = 0; r < P; r++)
                        {
                            const double2 q = cmul(v, A.tw[((ph * r) % P) * NS * twstride]);
                            double2* a = Acc + (size_t)l * N + j + NS * r;
                            if (ph == 0)
                                *a = q;
                            else
                                *a = make_double2(a->x + q.x, a->y + q.y);
                        }
                    }
                }
                __syncthreads();
            }
#pragma unroll 1
            for (int u = 0; u < U; u++)
#pragma unroll 1
                for (int r = 0; r < P; r++)
                {
                    const int sidx = t + u * NTHR + NS * r;
                    const int m = (sidx + N / 2) & (N - 1);
                    double2 d6[6], o5[5];
#pragma unroll
                    for (int l = 0; l < 6; l++)
                    {
                        const double2 a = Acc[(size_t)l * N + sidx];
                        d6[l] = make_double2(a.x * inv2N, a.y * inv2N);
                    }
                    fd_tdi<N>(c, m, A.fmt, d6, o5);
#pragma unroll
                    for (int ch = 0; ch < 3; ch++) H[(size_t)(S * 3 + ch) * N + m] = o5[ch];
                }
        };
        one_set(std::integral_constant<int, 0>{});
        one_set(std::integral_constant<int, 1>{});
        one_set(std::integral_constant<int, 2>{});
        one_set(std::integral_constant<int, 3>{});
        one_set(std::integral_constant<int, 4>{});
        one_set(std::integral_constant<int, 5>{});
        DerivItem it = deriv_item(A, e);
        double acc[kNSum];
#pragma unroll
        for (int q = 0; q < kNSum; q++) acc[q] = 0.0;
#pragma unroll 1
        for (int k = 0; k < PER; k++)
        {
            const int m = t + k * NTHR;
            double2 Hb[6][3];
#pragma unroll
            for (int S = 0; S < 6; S++)
#pragma unroll
                for (int ch = 0; ch < 3; ch++) Hb[S][ch] = H[(size_t)(S * 3 + ch) * N + m];
            deriv_bin(c, dd, A, it, m, Hb, acc);
        }
        deriv_write<NTHR>(A, e, acc, red);
    }
}

int launch_derivs_4096(const DerivArgs& A, int grid, double2* hs, hipStream_t s)
{
    if (grid <= 0) return GLASS_HIP_OK;
    if ((!A.list && A.nitems <= 0) || (A.list && !A.count) || !hs)
    {
        set_error("fd derivs: BW 4096 needs a class list or an item count, and scratch");
        return GLASS_HIP_EINVAL;
    }
    deriv_kernel_4096<kDeriv2048Threads><<<grid, kDeriv2048Threads, 0, s>>>(A, hs);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

int launch_derivs(int N, const DerivArgs& A, int grid, hipStream_t s)
{
    if (grid <= 0) return GLASS_HIP_OK;
    switch (N)
    {
#define GH_D(NN)                                                                   \
    case NN:                                                                       \
        deriv_kernel<NN, fd_threads(NN)><<<grid, fd_threads(NN), 0, s>>>(A);       \
        break;
        GH_D(32) GH_D(64) GH_D(128) GH_D(256) GH_D(512) GH_D(1024)
#undef GH_D
        default: set_error("fd derivs: BW=%d unsupported (power of two 32..1024)", N); return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

}  // namespace glass

using namespace glass;

extern "C" int glass_hip_fd_source_derivs(const glass_hip_orbit* orbit, const char* format, double T, double t0,
                                          int n, const double* d_params, int BW, const int* d_imin, int W,
                                          const double* d_resid, long long resid_stride, const double* d_invC,
                                          long long invC_stride, int mode, double* d_logL, double* d_grad,
                                          double* d_fisher, void* stream)
{
    int fmt = parse_format(format);
    if (fmt < 0)
    {
        set_error("Unsupported data format %s", format ? format : "(null)");
        return GLASS_HIP_EINVAL;
    }
    if (!orbit || n < 0 || !d_params || !d_imin || W < 1 || !d_invC || !d_logL || !d_grad || !d_fisher ||
        (mode != 0 && mode != 1) || resid_stride < 0 || invC_stride < 0)
    {
        set_error("fd_source_derivs: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    if (n == 0) return GLASS_HIP_OK;
    const double* pos;
    const double2* tw;
    int rc;
    if ((rc = positions(*orbit, T, t0, 4096, &pos)) || (rc = twiddles(&tw))) return rc;
    DerivArgs A{};
    A.params = d_params;
    A.imin = d_imin;
    A.W = W;
    A.resid = (const double2*)d_resid;
    A.rstride = resid_stride;
    A.invC = d_invC;
    A.istride = invC_stride;
    A.mode = mode;
    A.T = T;
    A.t0 = t0;
    A.L = orbit->L;
    A.fstar = orbit->fstar;
    A.pos = pos;
    A.tw = tw;
    A.fmt = fmt;
    A.outL = d_logL;
    A.grad = d_grad;
    A.fisher = d_fisher;
    if (BW <= 1024) return launch_derivs(BW, A, n, (hipStream_t)stream);
    if (BW != 2048 && BW != 4096)
    {
        set_error("fd derivs: BW=%d unsupported (power of two 32..4096)", BW);
        return GLASS_HIP_EINVAL;
    }
    // the decimated kernels stride over the items with a scratch slice per
    // workgroup (one workgroup per CU at most), held by the calling thread
    ThreadCtx* cx;
    if ((rc = ctx_get(&cx))) return rc;
    int ncu = 0;
    GH_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, cx->device));
    const int grid = n < ncu ? n : ncu;
    if ((rc = cx->dscr.ensure(sizeof(double2) * deriv_scratch(BW) * (size_t)grid))) return rc;
    A.nitems = n;
    double2* hs = (double2*)cx->dscr.p;
    return BW == 2048 ? launch_derivs_2048(A, grid, hs, (hipStream_t)stream)
                      : launch_derivs_4096(A, grid, hs, (hipStream_t)stream);
}
