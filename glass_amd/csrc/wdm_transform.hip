// wdm_transform.hip -- the WDM wavelet transforms of utils/src/glass_wavelet.c
// on gfx950:
//   forward  wavelet_transform_timefreq (:293-303) = real FFT of ND = NF*NT
//            samples, then per layer m a Meyer-windowed band of NT bins
//            around m*NT/2 and an inverse FFT of NT (wavelet_transform_freq,
//            :410-460);
//   inverse  wavelet_transform_inverse_freq (:351-405).
//
// The real FFT is a complex FFT of ND/2 points (the sample array read as
// interleaved complex pairs) in the four-step form M = M1*M2: column FFTs
// of length M1 over TC adjacent columns per workgroup (TC*16-byte
// contiguous loads), a twiddle, row FFTs of length M2 over TR adjacent rows
// with the output written in natural order (TR*16-byte contiguous stores).
// The layer kernel then forms the needed bins of the real spectrum from Z[k]
// and Z[M-k] on the fly.  All FFTs run in LDS (fft_lds.h); one workgroup
// per layer writes its NT pixels contiguously.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/glass_hip_wdm.h"
#include "fastmath.h"
#include "fft_lds.h"
#include "reduce.h"
#include "runtime.h"

namespace glass {
namespace wdm {

void wdm_window(int Nf, int Nt, double A, double* phif);  // wdm_ucb.hip

constexpr int kTfThreads = 256;  // the gather kernel and the PSD kernels

// Threads of a workgroup that transforms count series of length n in LDS:
// one radix-4 butterfly per thread and stage (n count / 4), a power of two
// from 64 to 1024.  With fewer threads than butterflies the workgroups of
// the column and row passes held one wave per SIMD.
__host__ __device__ constexpr int fft_threads(int n, int count)
{
    int b = n * count / 4, t = 64;
    while (t * 2 <= b && t < 1024) t *= 2;
    return t;
}
constexpr double kPI2 = 6.283185307179586;

// twiddle table (and its stride) for an FFT of length n
struct TwSel
{
    const double2* tw;
    int stride;
};

__host__ __device__ constexpr int fold_count(int n)  // transforms per workgroup: n * count <= 4096
{
    return n > 2048 ? 1 : (n > 1024 ? 2 : 4);
}

// Work index of workgroup b of nb: workgroups are dealt to the 8 XCDs in
// turn (b mod 8), so consecutive indices go to one XCD's blocks in order
// and neighbouring columns / rows / layers, which share cache lines, meet
// in the same L2.  The last nb mod 8 workgroups keep their own index.
__device__ inline int xcd_order(int b, int nb)
{
    const int full = nb & ~7;
    return b < full ? (b & 7) * (full >> 3) + (b >> 3) : b;
}

__device__ inline double2 twiddle_of(long long p, long long M)
{
    // exp(-2 pi i p / M), 0 <= p < M
    double s, c;
    sincos_red(-kPI2 * ((double)p / (double)M), &s, &c);
    return make_double2(c, s);
}

template <int M1, int NTH = fft_threads(M1, fold_count(M1))>
__global__ void __launch_bounds__(NTH) col_kernel(const double2* __restrict__ z, double2* __restrict__ y, int M2,
                                                  TwSel t)
{
    constexpr int TC = fold_count(M1);
    extern __shared__ double2 buf[];
    const int col0 = xcd_order(blockIdx.x, gridDim.x) * TC;
    const long long M = (long long)M1 * M2;
    for (int i = threadIdx.x; i < TC * M1; i += NTH)
    {
        const int n1 = i / TC, c = i % TC;
        buf[c * M1 + n1] = z[(size_t)n1 * M2 + col0 + c];
    }
    __syncthreads();
    lds_fft<M1, TC, NTH>(buf, t.tw, t.stride);
    for (int i = threadIdx.x; i < TC * M1; i += NTH)
    {
        const int k1 = i / TC, c = i % TC;
        const double2 v = buf[c * M1 + k1];
        const double2 w = twiddle_of(((long long)(col0 + c) * k1) % M, M);
        y[(size_t)k1 * M2 + col0 + c] = cmul(v, w);
    }
}

template <int M2, int NTH = fft_threads(M2, fold_count(M2))>
__global__ void __launch_bounds__(NTH) row_kernel(const double2* __restrict__ y, double2* __restrict__ Z, int M1,
                                                  TwSel t)
{
    constexpr int TR = fold_count(M2);
    extern __shared__ double2 buf[];
    const int row0 = xcd_order(blockIdx.x, gridDim.x) * TR;
    for (int i = threadIdx.x; i < TR * M2; i += NTH)
    {
        const int r = i / M2, n2 = i % M2;
        buf[r * M2 + n2] = y[(size_t)(row0 + r) * M2 + n2];
    }
    __syncthreads();
    lds_fft<M2, TR, NTH>(buf, t.tw, t.stride);
    for (int i = threadIdx.x; i < TR * M2; i += NTH)
    {
        const int k2 = i / TR, r = i % TR;
        Z[(size_t)(row0 + r) + (size_t)M1 * k2] = buf[r * M2 + k2];
    }
}

// Bin k (0 <= k <= M) of the real FFT of 2M samples from the complex FFT Z
// of their interleaved pairs: X = E + exp(-2 pi i k / 2M) O.
__device__ inline double2 real_bin(const double2* __restrict__ Z, long long k, long long M)
{
    const double2 a = Z[k % M], b = Z[(M - k % M) % M];
    const double2 E = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    const double2 O = make_double2(0.5 * (a.y + b.y), -0.5 * (a.x - b.x));
    const double2 w = twiddle_of(k, 2 * M);
    return make_double2(E.x + (w.x * O.x - w.y * O.y), E.y + (w.x * O.y + w.y * O.x));
}

// One layer m of wavelet_transform_freq.  FROM_Z: the spectrum is formed
// from the complex FFT Z of the interleaved samples; otherwise src holds
// the ND/2+1 one-sided bins.
template <int NT, bool FROM_Z, int NTH = fft_threads(NT, 1)>
__global__ void __launch_bounds__(NTH) layer_fwd_kernel(const double2* __restrict__ src, double* __restrict__ wdm,
                                                        int NF, const double* __restrict__ win, TwSel t)
{
    extern __shared__ double2 buf[];
    const int m = xcd_order(blockIdx.x, gridDim.x);
    const int center = m * NT / 2;
    const long long M = (long long)NF * NT / 2;
    for (int jj = threadIdx.x; jj < NT; jj += NTH)
    {
        const int j = jj - NT / 2;
        const int idx = center + j;
        double2 v = make_double2(0.0, 0.0);
        const bool skip = jj == 0 || (m == 0 && idx < 0) || (m == NF && idx > center);
        if (!skip)
        {
            double wt = win[j < 0 ? -j : j];
            if (j == 0 && (m == 0 || m == NF)) wt /= 2.0;
            const double2 S = FROM_Z ? real_bin(src, idx, M) : src[idx];
            // conjugate: the forward FFT then yields the conjugated inverse
            v = make_double2(wt * S.x, -(wt * S.y));
        }
        buf[jj] = v;
    }
    __syncthreads();
    lds_fft<NT, 1, NTH>(buf, t.tw, t.stride);
    for (int n = threadIdx.x; n < NT; n += NTH)
    {
        const double re = buf[n].x / NT, im = -buf[n].y / NT;
        if (m == 0)
        {
            if ((n & 1) == 0) wdm[n] = re * M_SQRT2;
        }
        else if (m == NF)
        {
            if ((n & 1) == 0) wdm[n + 1] = re * M_SQRT2;
        }
        else
        {
            const int sgn = (m % 2 == 0) ? 1 : -1;
            wdm[n + (size_t)NT * m] = ((n + m) % 2 == 0) ? re : sgn * im;
        }
    }
}

// wavelet_transform_inverse_freq, first half: per layer the prefactors and
// their forward FFT of NT.
template <int NT, int NTH = fft_threads(NT, 1)>
__global__ void __launch_bounds__(NTH) layer_inv_kernel(const double* __restrict__ wdm, double2* __restrict__ F,
                                                        int NF, TwSel t)
{
    extern __shared__ double2 buf[];
    const int m = blockIdx.x;
    for (int n = threadIdx.x; n < NT; n += NTH)
    {
        double2 v = make_double2(0.0, 0.0);
        if (m == 0) v.x = wdm[(2 * n) % NT] * M_SQRT1_2;
        else if (m == NF) v.x = wdm[(2 * n) % NT + 1] * M_SQRT1_2;
        else if ((n + m) % 2 == 0) v.x = wdm[n + (size_t)NT * m];
        else v.y = -wdm[n + (size_t)NT * m];
        buf[n] = v;
    }
    __syncthreads();
    lds_fft<NT, 1, NTH>(buf, t.tw, t.stride);
    for (int n = threadIdx.x; n < NT; n += NTH) F[(size_t)m * NT + n] = buf[n];
}

// Second half: each output bin i sums the windowed layer spectra that
// cover it, in increasing m (the reference's accumulation order).
__global__ void __launch_bounds__(kTfThreads) gather_inv_kernel(const double2* __restrict__ F,
                                                                double2* __restrict__ freq, int NF, int NT,
                                                                const double* __restrict__ win)
{
    const long long ND = (long long)NF * NT;
    const long long i = (long long)blockIdx.x * kTfThreads + threadIdx.x;
    if (i > ND / 2) return;
    const int q0 = (int)((2 * i) / NT);
    double2 acc = make_double2(0.0, 0.0);
    for (int m = q0 - 1; m <= q0 + 2; m++)
    {
        if (m < 0 || m > NF) continue;
        const long long center = (long long)m * NT / 2;
        long long lo = center - NT / 2, hi = center + NT / 2;
        if (lo < 0) lo = 0;
        if (hi > ND / 2 + 1) hi = ND / 2 + 1;
        if (i < lo || i >= hi) continue;
        const long long d = i > center ? i - center : center - i;
        if (d > NT / 2) continue;
        const int s = (m == 0 || m == NF) ? (int)((2 * i) % NT) : (int)(i % NT);
        const double2 f = F[(size_t)m * NT + s];
        acc.x += f.x * win[d] * NF / 2;
        acc.y += f.y * win[d] * NF / 2;
    }
    freq[i] = acc;
}

// ------------------------------------------------------------ dispatch --

static bool pow2(long long v) { return v > 0 && (v & (v - 1)) == 0; }

struct Plan
{
    int NF, NT;
    long long ND, M;
    int M1, M2;
};

static int make_plan(const glass_hip_wdm* w, Plan* p)
{
    if (!w) return set_error("NULL wdm"), GLASS_HIP_EINVAL;
    p->NF = w->NF;
    p->NT = w->NT;
    p->ND = (long long)w->NF * w->NT;
    if (!pow2(w->NT) || w->NT < 32 || w->NT > 4096)
        return set_error("WDM transform: NT = %d (a power of two in 32..4096 is required)", w->NT), GLASS_HIP_EINVAL;
    const int nf3 = (w->NF % 3 == 0) ? w->NF / 3 : w->NF;
    if (w->NF < 2 || !pow2(nf3))
        return set_error("WDM transform: NF = %d (2^a or 3*2^a is required)", w->NF), GLASS_HIP_EINVAL;
    if (p->ND < 2048) return set_error("WDM transform: ND = NF*NT = %lld < 2048", p->ND), GLASS_HIP_EINVAL;
    p->M = p->ND / 2;
    const bool three = (w->NF % 3 == 0);
    const long long P = three ? p->M / 3 : p->M;
    int c = 0;
    while ((1LL << c) < P) c++;
    const int c2 = (c + 1) / 2, c1 = c - c2;
    p->M2 = 1 << c2;
    p->M1 = (three ? 3 : 1) << c1;
    if (p->M2 > 4096 || p->M1 > 4096)
        return set_error("WDM transform: ND = %lld too large for the two-pass FFT", p->ND), GLASS_HIP_EINVAL;
    return GLASS_HIP_OK;
}

static TwSel tw_for(int n, const double2* tw, const double2* tw3)
{
    return (n % 3 == 0) ? TwSel{tw3, kTw3N / n} : TwSel{tw, kTwN / n};
}

#define GH_COL(N)                                                                                        \
    case N:                                                                                              \
        hipLaunchKernelGGL(col_kernel<N>, dim3(p.M2 / fold_count(N)), dim3(fft_threads(N, fold_count(N))), \
                           sizeof(double2) * N * fold_count(N), s, z, y, p.M2, tw_for(N, tw, tw3));      \
        break;
#define GH_ROW(N)                                                                                        \
    case N:                                                                                              \
        hipLaunchKernelGGL(row_kernel<N>, dim3(p.M1 / fold_count(N)), dim3(fft_threads(N, fold_count(N))), \
                           sizeof(double2) * N * fold_count(N), s, y, Z, p.M1, tw_for(N, tw, tw3));      \
        break;
#define GH_LAYER(N)                                                                                      \
    case N:                                                                                              \
        if (fromZ)                                                                                       \
            hipLaunchKernelGGL((layer_fwd_kernel<N, true>), dim3(p.NF + 1), dim3(fft_threads(N, 1)),     \
                               sizeof(double2) * N, s, src, out, p.NF, win, tw_for(N, tw, tw3));         \
        else                                                                                             \
            hipLaunchKernelGGL((layer_fwd_kernel<N, false>), dim3(p.NF + 1), dim3(fft_threads(N, 1)),    \
                               sizeof(double2) * N, s, src, out, p.NF, win, tw_for(N, tw, tw3));         \
        break;
#define GH_INV(N)                                                                                        \
    case N:                                                                                              \
        hipLaunchKernelGGL(layer_inv_kernel<N>, dim3(p.NF + 1), dim3(fft_threads(N, 1)), sizeof(double2) * N, s, \
                           pix, F, p.NF, tw_for(N, tw, tw3));                                            \
        break;

static int fft_2pass(const Plan& p, const double2* z, double2* y, double2* Z, hipStream_t s)
{
    const double2 *tw, *tw3;
    int rc;
    if ((rc = twiddles(&tw)) || (rc = twiddles3(&tw3))) return rc;
    switch (p.M1)
    {
        GH_COL(32) GH_COL(64) GH_COL(128) GH_COL(256) GH_COL(512) GH_COL(1024) GH_COL(2048) GH_COL(4096)
        GH_COL(48) GH_COL(96) GH_COL(192) GH_COL(384) GH_COL(768) GH_COL(1536) GH_COL(3072)
        default: return set_error("WDM transform: column size %d", p.M1), GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipGetLastError());
    switch (p.M2)
    {
        GH_ROW(32) GH_ROW(64) GH_ROW(128) GH_ROW(256) GH_ROW(512) GH_ROW(1024) GH_ROW(2048) GH_ROW(4096)
        default: return set_error("WDM transform: row size %d", p.M2), GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

static int layers_fwd(const Plan& p, const double2* src, bool fromZ, const double* win, double* out, hipStream_t s)
{
    const double2 *tw, *tw3;
    int rc;
    if ((rc = twiddles(&tw)) || (rc = twiddles3(&tw3))) return rc;
    switch (p.NT)
    {
        GH_LAYER(32) GH_LAYER(64) GH_LAYER(128) GH_LAYER(256) GH_LAYER(512) GH_LAYER(1024) GH_LAYER(2048)
        GH_LAYER(4096)
        default: return set_error("WDM transform: NT %d", p.NT), GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

static int layers_inv(const Plan& p, const double* pix, double2* F, const double* win, double2* freq,
                      hipStream_t s)
{
    const double2 *tw, *tw3;
    int rc;
    if ((rc = twiddles(&tw)) || (rc = twiddles3(&tw3))) return rc;
    switch (p.NT)
    {
        GH_INV(32) GH_INV(64) GH_INV(128) GH_INV(256) GH_INV(512) GH_INV(1024) GH_INV(2048) GH_INV(4096)
        default: return set_error("WDM transform: NT %d", p.NT), GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipGetLastError());
    const long long nb = (p.ND / 2 + 1 + kTfThreads - 1) / kTfThreads;
    hipLaunchKernelGGL(gather_inv_kernel, dim3((unsigned)nb), dim3(kTfThreads), 0, s, F, freq, p.NF, p.NT, win);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

// forward window of a grid, cached per (device, NF, NT, A)
static int window_dev(const Plan& p, double A, const double** out)
{
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, double>, double*> cache;
    int dev = 0;
    GH_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(dev, p.NF, p.NT, A);
    auto it = cache.find(key);
    if (it != cache.end())
    {
        *out = it->second;
        return GLASS_HIP_OK;
    }
    std::vector<double> h(p.NT / 2 + 1);
    wdm_window(p.NF, p.NT, A, h.data());
    double* d = nullptr;
    GH_CHECK(hipMalloc(&d, sizeof(double) * h.size()));
    GH_CHECK(hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    cache[key] = d;
    *out = d;
    return GLASS_HIP_OK;
}

// Device-pointer forward transform; scratch: 2 x ND doubles (y and Z).
static int forward_dev(const Plan& p, double A, const double* d_time, double* d_pix, double2* scratch,
                       hipStream_t s)
{
    const double* win;
    int rc;
    if ((rc = window_dev(p, A, &win))) return rc;
    double2* y = scratch;
    double2* Z = scratch + p.M;
    if ((rc = fft_2pass(p, (const double2*)d_time, y, Z, s))) return rc;
    return layers_fwd(p, Z, true, win, d_pix, s);
}

// ------------------------------------- stationary PSD -> WDM pixel PSD --
//
// stationary_dft_psd_to_wdm_layer_var (utils/src/glass_wavelet.c:865-899):
// the variance of a WDM pixel of layer m for stationary noise whose DFT
// variance is dft_psd[k] (k = 0..ND/2), as the window-weighted sum over the
// NT - 1 DFT bins of the layer's band, k = m NT/2 + l folded into [0, ND/2];
// the edge layers 0 and NF see one side of the band and their l = 0 weight
// is quartered.  One workgroup per layer (a fixed-order tree sum of
// non-negative terms: within rounding of the reference's running sum), then
// stationary_dft_psd_to_wdm_psd (:901-911) spreads the layer variances over
// the NF x NT pixels (row 0: even n layer 0, odd n layer NF).
constexpr int kPsdThreads = 256;

__global__ void __launch_bounds__(kPsdThreads) layer_var_kernel(int NF, int NT, const double* __restrict__ win,
                                                                const double* __restrict__ psd,
                                                                double* __restrict__ lv)
{
    const int m = blockIdx.x;
    const int half = NT / 2;
    const long long nmax = (long long)NF * NT / 2;
    double acc = 0.0;
    for (int l = -(half - 1) + (int)threadIdx.x; l < half; l += kPsdThreads)
    {
        if ((m == 0 && l < 0) || (m == NF && l > 0)) continue;
        const int al = l < 0 ? -l : l;
        double wt = win[al] * win[al];
        if (l == 0 && (m == 0 || m == NF)) wt *= 0.25;
        long long k = (long long)m * half + l;
        k = (k < 0 ? -k : k) % (2 * nmax);
        if (k > nmax) k = 2 * nmax - k;
        acc += wt * psd[k];
    }
    acc = block_sum<kPsdThreads>(acc);
    if (threadIdx.x == 0) lv[m] = (m > 0 && m < NF) ? acc / (2.0 * NT * NT) : acc / ((double)NT * NT);
}

__global__ void wdm_psd_fill_kernel(int NF, int NT, const double* __restrict__ lv, double* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)NF * NT) return;
    const int m = (int)(i / NT), n = (int)(i % NT);
    out[i] = m > 0 ? lv[m] : ((n & 1) ? lv[NF] : lv[0]);
}

static int psd_check(const glass_hip_wdm* w)
{
    if (!w) return set_error("NULL wdm"), GLASS_HIP_EINVAL;
    if (w->NF < 1 || w->NT < 2 || (w->NT & 1))
        return set_error("wdm psd: NF = %d, NT = %d (NT even >= 2)", w->NF, w->NT), GLASS_HIP_EINVAL;
    return GLASS_HIP_OK;
}

static int wdm_psd_dev(const glass_hip_wdm* w, const double* d_psd, double* d_lv, double* d_pix, hipStream_t s)
{
    Plan p{};
    p.NF = w->NF;
    p.NT = w->NT;
    const double* win;
    if (int rc = window_dev(p, w->A, &win)) return rc;
    layer_var_kernel<<<w->NF + 1, kPsdThreads, 0, s>>>(w->NF, w->NT, win, d_psd, d_lv);
    if (d_pix)
    {
        const long long n = (long long)w->NF * w->NT;
        wdm_psd_fill_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(w->NF, w->NT, d_lv, d_pix);
    }
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

}  // namespace wdm
}  // namespace glass

using namespace glass;
using namespace glass::wdm;

extern "C" int glass_hip_wdm_psd(const glass_hip_wdm* w, const double* d_dft_psd, double* d_layer_var,
                                 double* d_wdm_psd, void* stream)
{
    if (int rc = psd_check(w)) return rc;
    if (!d_dft_psd || !d_layer_var) return set_error("wdm_psd: NULL pointer"), GLASS_HIP_EINVAL;
    return wdm_psd_dev(w, d_dft_psd, d_layer_var, d_wdm_psd, (hipStream_t)stream);
}

// host arrays: dft_psd ND/2 + 1 in; layer_var NF + 1 and/or wdm_psd ND out
static int wdm_psd_host(const glass_hip_wdm* w, const double* dft_psd, double* layer_var, double* wdm_psd)
{
    if (int rc = psd_check(w)) return rc;
    if (!dft_psd) return set_error("wdm psd: NULL dft_psd"), GLASS_HIP_EINVAL;
    ThreadCtx* c;
    if (int rc = ctx_get(&c)) return rc;
    const size_t np = (size_t)w->NF * w->NT / 2 + 1, nl = (size_t)w->NF + 1;
    const size_t npix = wdm_psd ? (size_t)w->NF * w->NT : 0;
    if (int rc = c->dbuf.ensure(sizeof(double) * (np + nl + npix))) return rc;
    double* d_psd = (double*)c->dbuf.p;
    double* d_lv = d_psd + np;
    double* d_pix = wdm_psd ? d_lv + nl : nullptr;
    GH_CHECK(hipMemcpyAsync(d_psd, dft_psd, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    if (int rc = wdm_psd_dev(w, d_psd, d_lv, d_pix, c->stream)) return rc;
    if (layer_var) GH_CHECK(hipMemcpyAsync(layer_var, d_lv, sizeof(double) * nl, hipMemcpyDeviceToHost, c->stream));
    if (wdm_psd) GH_CHECK(hipMemcpyAsync(wdm_psd, d_pix, sizeof(double) * npix, hipMemcpyDeviceToHost, c->stream));
    GH_CHECK(hipStreamSynchronize(c->stream));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_stationary_dft_psd_to_wdm_layer_var(const glass_hip_wdm* w, const double* dft_psd,
                                                             double* layer_var)
{
    if (!layer_var) return set_error("wdm psd: NULL layer_var"), GLASS_HIP_EINVAL;
    return wdm_psd_host(w, dft_psd, layer_var, nullptr);
}

extern "C" int glass_hip_stationary_dft_psd_to_wdm_psd(const glass_hip_wdm* w, const double* dft_psd,
                                                       double* wdm_psd)
{
    if (!wdm_psd) return set_error("wdm psd: NULL wdm_psd"), GLASS_HIP_EINVAL;
    return wdm_psd_host(w, dft_psd, nullptr, wdm_psd);
}

extern "C" int glass_hip_wdm_transform(const glass_hip_wdm* w, const double* d_time, double* d_pixels,
                                       void* stream)
{
    Plan p;
    int rc = make_plan(w, &p);
    if (rc) return rc;
    if (!d_time || !d_pixels) return set_error("wdm_transform: NULL pointer"), GLASS_HIP_EINVAL;
    if (((uintptr_t)d_time & 15) != 0)
        return set_error("wdm_transform: d_time must be 16-byte aligned (read as complex pairs)"), GLASS_HIP_EINVAL;
    ThreadCtx* c;
    if ((rc = ctx_get(&c))) return rc;
    if ((rc = c->dpart.ensure(sizeof(double) * 2 * p.ND))) return rc;
    return forward_dev(p, w->A, d_time, d_pixels, (double2*)c->dpart.p, (hipStream_t)stream);
}

extern "C" int glass_hip_wavelet_transform_timefreq(const glass_hip_wdm* w, double* timedata)
{
    Plan p;
    int rc = make_plan(w, &p);
    if (rc) return rc;
    if (!timedata) return set_error("wavelet_transform_timefreq: NULL data"), GLASS_HIP_EINVAL;
    ThreadCtx* c;
    if ((rc = ctx_get(&c))) return rc;
    const size_t bytes = sizeof(double) * p.ND;
    if ((rc = c->dbuf.ensure(2 * bytes)) || (rc = c->dpart.ensure(2 * bytes))) return rc;
    double* d_time = (double*)c->dbuf.p;
    double* d_pix = d_time + p.ND;
    GH_CHECK(hipMemcpyAsync(d_time, timedata, bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = forward_dev(p, w->A, d_time, d_pix, (double2*)c->dpart.p, c->stream))) return rc;
    GH_CHECK(hipMemcpyAsync(timedata, d_pix, bytes, hipMemcpyDeviceToHost, c->stream));
    GH_CHECK(hipStreamSynchronize(c->stream));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_wavelet_transform_freq(const glass_hip_wdm* w, const double* freqdata, double* wdmdata)
{
    Plan p;
    int rc = make_plan(w, &p);
    if (rc) return rc;
    if (!freqdata || !wdmdata) return set_error("wavelet_transform_freq: NULL data"), GLASS_HIP_EINVAL;
    ThreadCtx* c;
    if ((rc = ctx_get(&c))) return rc;
    const double* win;
    if ((rc = window_dev(p, w->A, &win))) return rc;
    const size_t fb = sizeof(double) * (p.ND + 2), pb = sizeof(double) * p.ND;
    if ((rc = c->dbuf.ensure(fb + pb))) return rc;
    double* d_f = (double*)c->dbuf.p;
    double* d_pix = d_f + p.ND + 2;
    GH_CHECK(hipMemcpyAsync(d_f, freqdata, fb, hipMemcpyHostToDevice, c->stream));
    if ((rc = layers_fwd(p, (const double2*)d_f, false, win, d_pix, c->stream))) return rc;
    GH_CHECK(hipMemcpyAsync(wdmdata, d_pix, pb, hipMemcpyDeviceToHost, c->stream));
    GH_CHECK(hipStreamSynchronize(c->stream));
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_wavelet_transform_inverse_freq(const glass_hip_wdm* w, const double* wdmdata,
                                                        double* freqdata)
{
    Plan p;
    int rc = make_plan(w, &p);
    if (rc) return rc;
    if (!freqdata || !wdmdata) return set_error("wavelet_transform_inverse_freq: NULL data"), GLASS_HIP_EINVAL;
    ThreadCtx* c;
    if ((rc = ctx_get(&c))) return rc;
    const double* win;
    if ((rc = window_dev(p, w->A, &win))) return rc;
    const size_t pb = sizeof(double) * p.ND, fb = sizeof(double) * (p.ND + 2);
    const size_t Fb = sizeof(double2) * (size_t)(p.NF + 1) * p.NT;
    if ((rc = c->dbuf.ensure(pb + fb)) || (rc = c->dpart.ensure(Fb))) return rc;
    double* d_pix = (double*)c->dbuf.p;
    double* d_f = d_pix + p.ND;
    GH_CHECK(hipMemcpyAsync(d_pix, wdmdata, pb, hipMemcpyHostToDevice, c->stream));
    if ((rc = layers_inv(p, d_pix, (double2*)c->dpart.p, win, (double2*)d_f, c->stream))) return rc;
    GH_CHECK(hipMemcpyAsync(freqdata, d_f, fb, hipMemcpyDeviceToHost, c->stream));
    GH_CHECK(hipStreamSynchronize(c->stream));
    return GLASS_HIP_OK;
}

// ------------------------------------------------------------ by layers --
//
// wavelet_transform_timefreq_by_layers (utils/src/glass_wavelet.c:544-611):
// N = NT (Nlayers + 1) heterodyned samples -> rFFT(N) -> per layer m =
// jmin .. jmin + Nlayers - 1 the window-weighted band of NT bins around
// (m - jmin + 1) NT / 2, inverse FFT(NT) / NT, re or +-im by the (n + m)
// parity, scaled by sqrt(NF) / (Nlayers + 1); pixel (n, m) at n Nlayers +
// m - jmin, the rest of the N samples zero.  One workgroup; the real input
// is transformed as a complex series with zero imaginary part.

namespace glass {
namespace {

constexpr int kBlThreads = 256;

template <int NT, int LP1>
__global__ void __launch_bounds__(kBlThreads) by_layers_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                               const double* __restrict__ win, int jmin, int NF,
                                                               const double2* __restrict__ tw,
                                                               const double2* __restrict__ tw3)
{
    constexpr int N = NT * LP1, L = LP1 - 1;
    __shared__ double2 F[N];
    __shared__ double2 B[L * NT];
    const int tid = threadIdx.x;
    for (int i = tid; i < N; i += kBlThreads) F[i] = make_double2(in[i], 0.0);
    __syncthreads();
    constexpr bool three = N % 3 == 0;
    lds_fft<N, 1, kBlThreads>(F, three ? tw3 : tw, (three ? kTw3N : kTwN) / N);
    for (int it = tid; it < L * NT; it += kBlThreads)
    {
        const int off = it / NT, jj = it % NT;
        const int j = jj - NT / 2;
        const int m = jmin + off;
        const int idx = (off + 1) * NT / 2 + j;
        double2 v = make_double2(0.0, 0.0);
        if (jj > 0 && idx >= 0 && idx <= N / 2)
        {
            double wt = win[j < 0 ? -j : j];
            if (j == 0 && (m == 0 || m == NF)) wt /= 2.0;
            // conjugated: the forward FFT then gives the conjugate of the inverse
            v = make_double2(wt * F[idx].x, -(wt * F[idx].y));
        }
        B[off * NT + jj] = v;
    }
    __syncthreads();
    lds_fft<NT, L, kBlThreads>(B, tw, kTwN / NT);
    const double norm = sqrt((double)NF) / (L + 1.0);
    for (int k = tid; k < N; k += kBlThreads)
    {
        double val = 0.0;
        if (k < NT * L)
        {
            const int n = k / L, off = k % L;
            const int m = jmin + off;
            const double2 r = B[off * NT + n];
            const double re = r.x / NT, im = -r.y / NT;
            if (m == 0 || m == NF) val = (n % 2 == 0) ? re * M_SQRT2 * norm : 0.0;
            else
            {
                const int sgn = (m % 2 == 0) ? 1 : -1;
                val = ((n + m) % 2 == 0) ? re * norm : sgn * im * norm;
            }
        }
        out[k] = val;
    }
}

}  // namespace
}  // namespace glass

extern "C" int glass_hip_wavelet_transform_timefreq_by_layers(const glass_hip_wdm* w, int jmin, int Nlayers,
                                                              const double* window, double* timedata)
{
    using namespace glass;
    if (!w || !window || !timedata) return set_error("wavelet_transform_timefreq_by_layers: NULL argument"), GLASS_HIP_EINVAL;
    const int NT = w->NT;
    if ((NT != 256 && NT != 512) || Nlayers < 1 || Nlayers > 3 || jmin < 0 || jmin + Nlayers - 1 > w->NF)
        return set_error("wavelet_transform_timefreq_by_layers: NT = %d (256 or 512), Nlayers = %d (1..3), "
                         "jmin = %d (layers within 0..NF = %d)", NT, Nlayers, jmin, w->NF),
               GLASS_HIP_EINVAL;
    ThreadCtx* c;
    int rc = ctx_get(&c);
    if (rc) return rc;
    const double2 *tw, *tw3;
    if ((rc = twiddles(&tw)) || (rc = twiddles3(&tw3))) return rc;
    const int N = NT * (Nlayers + 1), nw = NT / 2 + 1;
    const size_t bytes = sizeof(double) * (2 * (size_t)N + nw);
    if ((rc = c->dbuf.ensure(bytes)) || (rc = c->hbuf.ensure(bytes))) return rc;
    double* d_in = (double*)c->dbuf.p;
    double* d_out = d_in + N;
    double* d_win = d_out + N;
    double* h = (double*)c->hbuf.p;
    std::memcpy(h, timedata, sizeof(double) * N);
    std::memcpy(h + 2 * (size_t)N, window, sizeof(double) * nw);
    GH_CHECK(hipMemcpyAsync(d_in, h, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
    GH_CHECK(hipMemcpyAsync(d_win, h + 2 * (size_t)N, sizeof(double) * nw, hipMemcpyHostToDevice, c->stream));
#define GH_BL(nt, lp1)                                                                                     \
    if (NT == nt && Nlayers + 1 == lp1)                                                                    \
        hipLaunchKernelGGL((by_layers_kernel<nt, lp1>), dim3(1), dim3(kBlThreads), 0, c->stream, d_in, d_out, \
                           d_win, jmin, w->NF, tw, tw3);
    GH_BL(256, 2) GH_BL(256, 3) GH_BL(256, 4) GH_BL(512, 2) GH_BL(512, 3) GH_BL(512, 4)
#undef GH_BL
    GH_CHECK(hipGetLastError());
    GH_CHECK(hipMemcpyAsync(h, d_out, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    GH_CHECK(hipStreamSynchronize(c->stream));
    std::memcpy(timedata, h, sizeof(double) * N);
    return GLASS_HIP_OK;
}
