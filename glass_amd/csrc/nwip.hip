// nwip.hip -- noise-weighted inner products and the Fourier-domain
// Gaussian likelihood (utils/src/glass_math.c:362-426,
// utils/src/glass_model.c:726-774), and the 3x3 inverse noise covariance
// (utils/src/glass_math.c:240-289).
//
// Streaming, HBM-bound kernels: each block reduces a contiguous range of
// bins with coalesced loads, writes one partial; a single-block pass sums
// the partials in block order (deterministic).
#include <cmath>
#include <cstring>

#include "reduce.h"
#include "runtime.h"
#include "../../include/glass_hip_wdm.h"

namespace glass {

constexpr int kRedThr = 256;
constexpr int kMaxPart = 1024;

// Generic weighted sum of products: term t of bin n is
// a_t[n] . b_t[n] * w_t[n] * coef_t  with complex a,b (Re a conj b).
// Used for fourier_nwip (1 term) and snr (6 terms).
struct NwipTerms
{
    const double2* a[6];
    const double2* b[6];
    const double* w[6];
    double coef[6];
    int nterm;
};

__global__ __launch_bounds__(kRedThr) void nwip_kernel(NwipTerms T, int N, int chunk, double* part)
{
    const int b0 = blockIdx.x * chunk, b1 = min(N, b0 + chunk);
    double acc = 0.0;
    for (int t = 0; t < T.nterm; t++)
    {
        double s = 0.0;
        for (int n = b0 + threadIdx.x; n < b1; n += kRedThr)
        {
            double2 a = T.a[t][n], b = T.b[t][n];
            s += (a.x * b.x + a.y * b.y) * T.w[t][n];
        }
        acc += T.coef[t] * s;
    }
    acc = block_sum<kRedThr>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kRedThr) void sum_partials(const double* part, int n, double scale, double* out)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kRedThr) s += part[i];
    s = block_sum<kRedThr>(s);
    if (threadIdx.x == 0) out[0] = scale * s;
}

// chi^2 of the 3-channel residual r = d - m over NFFT bins:
// sum_n 2*[ XX iC00 + YY iC11 + ZZ iC22 + 2(XY iC01 + XZ iC02 + YZ iC12) ]
// (glass_model.c:758-767 with fourier_nwip's factor 2)
__global__ __launch_bounds__(kRedThr) void logl3_kernel(const double2* __restrict__ dX, const double2* __restrict__ dY,
                                                        const double2* __restrict__ dZ, const double2* __restrict__ mX,
                                                        const double2* __restrict__ mY, const double2* __restrict__ mZ,
                                                        const double* __restrict__ invC, int N, int chunk, double* part)
{
    const int b0 = blockIdx.x * chunk, b1 = min(N, b0 + chunk);
    double acc = 0.0;
    for (int n = b0 + threadIdx.x; n < b1; n += kRedThr)
    {
        double2 x = dX[n], y = dY[n], z = dZ[n];
        double2 a = mX[n], b = mY[n], c = mZ[n];
        x.x -= a.x; x.y -= a.y;
        y.x -= b.x; y.y -= b.y;
        z.x -= c.x; z.y -= c.y;
        double diag = (x.x * x.x + x.y * x.y) * invC[0 * (size_t)N + n] +
                      (y.x * y.x + y.y * y.y) * invC[4 * (size_t)N + n] +
                      (z.x * z.x + z.y * z.y) * invC[8 * (size_t)N + n];
        double off = (x.x * y.x + x.y * y.y) * invC[1 * (size_t)N + n] +
                     (x.x * z.x + x.y * z.y) * invC[2 * (size_t)N + n] +
                     (y.x * z.x + y.y * z.y) * invC[5 * (size_t)N + n];
        acc += diag + 2.0 * off;
    }
    acc = block_sum<kRedThr>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kRedThr) void wnwip_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                        const double* __restrict__ w, const int* __restrict__ list,
                                                        int N, int chunk, double* part)
{
    const int b0 = blockIdx.x * chunk, b1 = min(N, b0 + chunk);
    double acc = 0.0;
    for (int n = b0 + threadIdx.x; n < b1; n += kRedThr)
    {
        int k = list[n];
        acc += a[k] * b[k] * w[k];
    }
    acc = block_sum<kRedThr>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Per-bin analytic 3x3 inverse with Cholesky-like pivots and signed
// log-domain division (glass_math.c:236-289).
__global__ void invert3_kernel(const double* __restrict__ C, double* __restrict__ iC, double* __restrict__ ld, int N)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const size_t P = (size_t)N;
    double cxx = C[0 * P + n], cyy = C[4 * P + n], czz = C[8 * P + n];
    double cxy = C[1 * P + n], cxz = C[2 * P + n], cyz = C[5 * P + n];
    double d1 = cxx;
    double d2 = cyy - cxy * cxy / d1;
    double t = cyz - cxy * cxz / d1;
    double d3 = (czz - cxz * cxz / d1) - (t * t) / d2;
    double l = log(d1) + log(d2) + log(d3);
    if (ld) ld[n] = l;
    auto sld = [l](double num) { return copysign(exp(log(fabs(num)) - l), num); };
    double ixx = sld(cyy * czz - cyz * cyz), iyy = sld(czz * cxx - cxz * cxz), izz = sld(cxx * cyy - cxy * cxy);
    double ixy = sld(cxz * cyz - czz * cxy), ixz = sld(cxy * cyz - cxz * cyy), iyz = sld(cxy * cxz - cxx * cyz);
    iC[0 * P + n] = ixx; iC[1 * P + n] = ixy; iC[2 * P + n] = ixz;
    iC[3 * P + n] = ixy; iC[4 * P + n] = iyy; iC[5 * P + n] = iyz;
    iC[6 * P + n] = ixz; iC[7 * P + n] = iyz; iC[8 * P + n] = izz;
}

__global__ void sub_kernel(double* a, const double* b, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] -= b[i];
}

static int grid_for(int N, int* chunk)
{
    int nb = (N + 4 * kRedThr - 1) / (4 * kRedThr);
    if (nb < 1) nb = 1;
    if (nb > kMaxPart) nb = kMaxPart;
    *chunk = (N + nb - 1) / nb;
    return nb;
}

// Device-pointer reductions (stream-ordered); result to *d_out.
int nwip_terms_dev(const NwipTerms& T, int N, double scale, double* d_out, ThreadCtx* cx)
{
    int rc, chunk;
    if ((rc = cx->dpart.ensure(sizeof(double) * kMaxPart))) return rc;
    int nb = grid_for(N, &chunk);
    nwip_kernel<<<nb, kRedThr, 0, cx->stream>>>(T, N, chunk, (double*)cx->dpart.p);
    sum_partials<<<1, kRedThr, 0, cx->stream>>>((double*)cx->dpart.p, nb, scale, d_out);
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

// Stage host arrays into the thread's device buffer; returns device ptrs.
static int stage(ThreadCtx* cx, int n, const void* const* src, const size_t* bytes, void** dst, double** d_res)
{
    size_t tot = 256;
    for (int i = 0; i < n; i++) tot += (bytes[i] + 255) & ~(size_t)255;
    int rc;
    if ((rc = cx->dbuf.ensure(tot))) return rc;
    if ((rc = cx->hbuf.ensure(tot))) return rc;
    size_t off = 0;
    for (int i = 0; i < n; i++)
    {
        memcpy((char*)cx->hbuf.p + off, src[i], bytes[i]);
        dst[i] = (char*)cx->dbuf.p + off;
        off += (bytes[i] + 255) & ~(size_t)255;
    }
    GH_CHECK(hipMemcpyAsync(cx->dbuf.p, cx->hbuf.p, off, hipMemcpyHostToDevice, cx->stream));
    *d_res = (double*)((char*)cx->dbuf.p + off);
    return GLASS_HIP_OK;
}

static double fetch_result(ThreadCtx* cx, double* d_res)
{
    double* h = (double*)((char*)cx->hbuf.p);
    if (hipMemcpyAsync(h, d_res, sizeof(double), hipMemcpyDeviceToHost, cx->stream) != hipSuccess ||
        hipStreamSynchronize(cx->stream) != hipSuccess)
    {
        set_error("result copy failed");
        return NAN;
    }
    return h[0];
}

}  // namespace glass

using namespace glass;

extern "C" double glass_hip_fourier_nwip(const double* a, const double* b, const double* invC, int N)
{
    ThreadCtx* cx;
    if (N < 0 || (N > 0 && (!a || !b || !invC)))
    {
        set_error("fourier_nwip: bad arguments");
        return NAN;
    }
    if (N == 0) return 0.0;
    if (ctx_get(&cx)) return NAN;
    const void* src[3] = {a, b, invC};
    size_t by[3] = {sizeof(double) * 2 * N, sizeof(double) * 2 * N, sizeof(double) * N};
    void* d[3];
    double* dres;
    if (stage(cx, 3, src, by, d, &dres)) return NAN;
    NwipTerms T{};
    T.nterm = 1;
    T.a[0] = (const double2*)d[0];
    T.b[0] = (const double2*)d[1];
    T.w[0] = (const double*)d[2];
    T.coef[0] = 1.0;
    if (nwip_terms_dev(T, N, 2.0, dres, cx)) return NAN;
    return fetch_result(cx, dres);
}

extern "C" double glass_hip_snr(const double* X, const double* Y, const double* Z, const double* invC, int Ninv,
                                int N)
{
    ThreadCtx* cx;
    if (N < 0 || N > Ninv || !X || !Y || !Z || !invC)
    {
        set_error("snr: bad arguments");
        return NAN;
    }
    if (N == 0) return 0.0;
    if (ctx_get(&cx)) return NAN;
    // only the bins that are read: invC planes are taken from window bin 0
    // for N bins (glass_math.c:409-421)
    const void* src[9] = {X, Y, Z, invC, invC + 4 * (size_t)Ninv, invC + 8 * (size_t)Ninv, invC + 1 * (size_t)Ninv,
                          invC + 2 * (size_t)Ninv, invC + 5 * (size_t)Ninv};
    size_t by[9];
    for (int i = 0; i < 3; i++) by[i] = sizeof(double) * 2 * N;
    for (int i = 3; i < 9; i++) by[i] = sizeof(double) * N;
    void* d[9];
    double* dres;
    if (stage(cx, 9, src, by, d, &dres)) return NAN;
    NwipTerms T{};
    T.nterm = 6;
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int t = 0; t < 6; t++)
    {
        T.a[t] = (const double2*)d[ia[t]];
        T.b[t] = (const double2*)d[ib[t]];
        T.w[t] = (const double*)d[3 + t];
        T.coef[t] = t < 3 ? 1.0 : 2.0;
    }
    if (nwip_terms_dev(T, N, 2.0, dres, cx)) return NAN;
    return sqrt(fetch_result(cx, dres));
}

extern "C" double glass_hip_wavelet_nwip(const double* a, const double* b, const double* invC, const int* list,
                                         int N, int nmax)
{
    ThreadCtx* cx;
    if (N < 0 || nmax < 0 || (N > 0 && (!a || !b || !invC || !list)))
    {
        set_error("wavelet_nwip: bad arguments");
        return NAN;
    }
    for (int i = 0; i < N; i++)
        if (list[i] < 0 || list[i] >= nmax)
        {
            set_error("wavelet_nwip: list[%d]=%d outside [0,%d)", i, list[i], nmax);
            return NAN;
        }
    if (N == 0) return 0.0;
    if (ctx_get(&cx)) return NAN;
    const void* src[4] = {a, b, invC, list};
    size_t by[4] = {sizeof(double) * nmax, sizeof(double) * nmax, sizeof(double) * nmax, sizeof(int) * N};
    void* d[4];
    double* dres;
    if (stage(cx, 4, src, by, d, &dres)) return NAN;
    int rc, chunk;
    if ((rc = cx->dpart.ensure(sizeof(double) * kMaxPart))) return NAN;
    int nb = grid_for(N, &chunk);
    wnwip_kernel<<<nb, kRedThr, 0, cx->stream>>>((const double*)d[0], (const double*)d[1], (const double*)d[2],
                                                  (const int*)d[3], N, chunk, (double*)cx->dpart.p);
    sum_partials<<<1, kRedThr, 0, cx->stream>>>((double*)cx->dpart.p, nb, 1.0, dres);
    if (hipGetLastError() != hipSuccess)
    {
        set_error("wavelet_nwip launch failed");
        return NAN;
    }
    return fetch_result(cx, dres);
}

extern "C" double glass_hip_gaussian_log_likelihood(const double* dX, const double* dY, const double* dZ,
                                                    const double* dA, const double* dE, const double* mX,
                                                    const double* mY, const double* mZ, const double* mA,
                                                    const double* mE, const double* invC, int NFFT, int Nchannel)
{
    ThreadCtx* cx;
    if (NFFT <= 0 || !invC || Nchannel < 1 || Nchannel > 3)
    {
        set_error("gaussian_log_likelihood: bad arguments");
        return NAN;
    }
    if (ctx_get(&cx)) return NAN;
    const size_t cb = sizeof(double) * 2 * NFFT, pb = sizeof(double) * NFFT;
    double* dres;
    if (Nchannel == 3)
    {
        if (!dX || !dY || !dZ || !mX || !mY || !mZ)
        {
            set_error("gaussian_log_likelihood: X,Y,Z required");
            return NAN;
        }
        const void* src[7] = {dX, dY, dZ, mX, mY, mZ, invC};
        size_t by[7] = {cb, cb, cb, cb, cb, cb, 9 * pb};
        void* d[7];
        if (stage(cx, 7, src, by, d, &dres)) return NAN;
        int chunk;
        if (cx->dpart.ensure(sizeof(double) * kMaxPart)) return NAN;
        int nb = grid_for(NFFT, &chunk);
        logl3_kernel<<<nb, kRedThr, 0, cx->stream>>>((const double2*)d[0], (const double2*)d[1], (const double2*)d[2],
                                                      (const double2*)d[3], (const double2*)d[4], (const double2*)d[5],
                                                      (const double*)d[6], NFFT, chunk, (double*)cx->dpart.p);
        // chi2 = 2 * sum; logL = -chi2/2
        sum_partials<<<1, kRedThr, 0, cx->stream>>>((double*)cx->dpart.p, nb, -1.0, dres);
        if (hipGetLastError() != hipSuccess)
        {
            set_error("logL launch failed");
            return NAN;
        }
        return fetch_result(cx, dres);
    }
    // 1 channel (X) or 2 channels (A,E): diagonal terms only (glass_model.c:749-756)
    const double* dd[2] = {Nchannel == 1 ? dX : dA, dE};
    const double* mm[2] = {Nchannel == 1 ? mX : mA, mE};
    const int nc = Nchannel;
    for (int c = 0; c < nc; c++)
        if (!dd[c] || !mm[c])
        {
            set_error("gaussian_log_likelihood: channel arrays required");
            return NAN;
        }
    // residuals formed on the host side of the copy would change rounding
    // order only at the ulp level; form them on the device instead
    const void* src[6] = {dd[0], mm[0], invC, dd[nc - 1], mm[nc - 1], invC + 4 * (size_t)NFFT};
    size_t by[6] = {cb, cb, pb, cb, cb, pb};
    void* d[6];
    if (stage(cx, 6, src, by, d, &dres)) return NAN;
    // residual in place on the device copies of the data arrays
    // r = d - m via a 2-term expansion: (d-m).(d-m) = d.d - 2 d.m + m.m
    // would lose precision; use a dedicated pass instead.
    int n2 = 2 * NFFT;
    for (int c = 0; c < nc; c++)
        sub_kernel<<<(n2 + 255) / 256, 256, 0, cx->stream>>>((double*)d[3 * c], (const double*)d[3 * c + 1], n2);
    NwipTerms T{};
    T.nterm = nc;
    for (int c = 0; c < nc; c++)
    {
        T.a[c] = T.b[c] = (const double2*)d[3 * c];
        T.w[c] = (const double*)d[3 * c + 2];
        T.coef[c] = 1.0;
    }
    if (nwip_terms_dev(T, NFFT, -1.0, dres, cx)) return NAN;
    return fetch_result(cx, dres);
}

extern "C" int glass_hip_invert_noise_covariance(const double* C, double* invC, double* logdetC, int N)
{
    ThreadCtx* cx;
    if (N <= 0 || !C || !invC)
    {
        set_error("invert_noise_covariance: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    int rc;
    if ((rc = ctx_get(&cx))) return rc;
    const size_t pb = sizeof(double) * N;
    if ((rc = cx->dbuf.ensure(19 * pb + 512))) return rc;
    if ((rc = cx->hbuf.ensure(19 * pb + 512))) return rc;
    double* dC = (double*)cx->dbuf.p;
    double* diC = dC + 9 * (size_t)N;
    double* dld = diC + 9 * (size_t)N;
    memcpy(cx->hbuf.p, C, 9 * pb);
    GH_CHECK(hipMemcpyAsync(dC, cx->hbuf.p, 9 * pb, hipMemcpyHostToDevice, cx->stream));
    invert3_kernel<<<(N + 255) / 256, 256, 0, cx->stream>>>(dC, diC, dld, N);
    GH_CHECK(hipGetLastError());
    GH_CHECK(hipMemcpyAsync(cx->hbuf.p, diC, 10 * pb, hipMemcpyDeviceToHost, cx->stream));
    GH_CHECK(hipStreamSynchronize(cx->stream));
    memcpy(invC, cx->hbuf.p, 9 * pb);
    if (logdetC) memcpy(logdetC, (double*)cx->hbuf.p + 9 * (size_t)N, pb);
    return GLASS_HIP_OK;
}

// ------------------------------------------------ wavelet domain -------

namespace glass {

// sum_k r^T C r over n pixels: r 3 planes, C 6 planes (XX,YY,ZZ,XY,XZ,YZ);
// per pixel XX + 2 XY as snr_wavelet / gaussian_log_likelhood_wavelet
// group it (utils/src/glass_math.c:428-438, utils/src/glass_model.c:839-853)
__global__ __launch_bounds__(kRedThr) void wquad_kernel(const double* __restrict__ r, const double* __restrict__ C,
                                                        int n, int chunk, double* part)
{
    const int b0 = blockIdx.x * chunk, b1 = min(n, b0 + chunk);
    double acc = 0.0;
    for (int k = b0 + threadIdx.x; k < b1; k += kRedThr)
    {
        const double x = r[k], y = r[n + k], z = r[2 * n + k];
        const double xx = x * x * C[k] + y * y * C[n + k] + z * z * C[2 * n + k];
        const double xy = x * y * C[3 * n + k] + x * z * C[4 * n + k] + y * z * C[5 * n + k];
        acc += xx + 2.0 * xy;
    }
    acc = block_sum<kRedThr>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Stages n pixels of 9 planes (r: 3, C: 6) already laid out in the pinned
// buffer, reduces, returns the sum.
static double wquad_staged(ThreadCtx* cx, int n)
{
    const size_t planes = sizeof(double) * 9 * (size_t)n;
    if (hipMemcpyAsync(cx->dbuf.p, cx->hbuf.p, planes, hipMemcpyHostToDevice, cx->stream) != hipSuccess)
    {
        set_error("wavelet staging copy failed");
        return NAN;
    }
    int chunk;
    if (cx->dpart.ensure(sizeof(double) * kMaxPart)) return NAN;
    const int nb = grid_for(n, &chunk);
    const double* d = (const double*)cx->dbuf.p;
    double* dres = (double*)((char*)cx->dbuf.p + ((planes + 255) & ~(size_t)255));
    wquad_kernel<<<nb, kRedThr, 0, cx->stream>>>(d, d + 3 * (size_t)n, n, chunk, (double*)cx->dpart.p);
    sum_partials<<<1, kRedThr, 0, cx->stream>>>((double*)cx->dpart.p, nb, 1.0, dres);
    if (hipGetLastError() != hipSuccess)
    {
        set_error("wavelet reduction launch failed");
        return NAN;
    }
    return fetch_result(cx, dres);
}

static const int kPlane6[6] = {0, 4, 8, 1, 2, 5};  // XX YY ZZ XY XZ YZ in a row-major 3x3

}  // namespace glass

extern "C" double glass_hip_snr_wavelet(const double* X, const double* Y, const double* Z, const double* invC, int N,
                                        const int* list, int Nlist)
{
    if (N < 0 || Nlist < 0 || (Nlist > 0 && (!X || !Y || !Z || !invC || !list)))
    {
        set_error("snr_wavelet: bad arguments");
        return NAN;
    }
    for (int t = 0; t < Nlist; t++)
        if (list[t] < 0 || list[t] >= N)
        {
            set_error("snr_wavelet: list[%d]=%d outside [0,%d)", t, list[t], N);
            return NAN;
        }
    if (Nlist == 0) return 0.0;
    ThreadCtx* cx;
    if (ctx_get(&cx)) return NAN;
    const size_t bytes = sizeof(double) * 9 * (size_t)Nlist + 512;
    if (cx->dbuf.ensure(bytes) || cx->hbuf.ensure(bytes)) return NAN;
    // the listed pixels only: 3 template planes and 6 invC planes
    double* h = (double*)cx->hbuf.p;
    const double* ch[3] = {X, Y, Z};
    for (int t = 0; t < Nlist; t++)
    {
        const int k = list[t];
        for (int a = 0; a < 3; a++) h[a * (size_t)Nlist + t] = ch[a][k];
        for (int p = 0; p < 6; p++) h[(3 + p) * (size_t)Nlist + t] = invC[kPlane6[p] * (size_t)N + k];
    }
    const double s2 = wquad_staged(cx, Nlist);
    return sqrt(s2);
}

extern "C" double glass_hip_gaussian_log_likelihood_wavelet(const glass_hip_wdm* w, const double* dX,
                                                            const double* dY, const double* dZ, const double* mX,
                                                            const double* mY, const double* mZ, const int* mlist,
                                                            int Nlist, const double* invC, int N)
{
    if (!w || N <= 0 || Nlist < 0 || !dX || !dY || !dZ || !invC || (Nlist > 0 && (!mX || !mY || !mZ || !mlist)) ||
        w->NT <= 0)
    {
        set_error("gaussian_log_likelihood_wavelet: bad arguments");
        return NAN;
    }
    ThreadCtx* cx;
    if (ctx_get(&cx)) return NAN;
    const size_t bytes = sizeof(double) * 9 * (size_t)N + 512;
    if (cx->dbuf.ensure(bytes) || cx->hbuf.ensure(bytes)) return NAN;
    // residual with the reference's list rule, formed in the staging buffer
    double* r = (double*)cx->hbuf.p;
    memcpy(r, dX, sizeof(double) * N);
    memcpy(r + N, dY, sizeof(double) * N);
    memcpy(r + 2 * (size_t)N, dZ, sizeof(double) * N);
    for (int t = 0; t < Nlist; t++)
    {
        const int k = mlist[t];
        if (k >= 0 && k < N)
        {
            const int i = (k + w->kmin) % w->NT;
            r[k] -= mX[k];
            r[N + k] -= mY[k];
            r[2 * (size_t)N + k] -= mZ[k];
            if (i < w->imin || i > w->imax) r[k] = r[N + k] = r[2 * (size_t)N + k] = 0.0;
        }
    }
    for (int p = 0; p < 6; p++) memcpy(r + (3 + p) * (size_t)N, invC + kPlane6[p] * (size_t)N, sizeof(double) * N);
    return -0.5 * wquad_staged(cx, N);
}
