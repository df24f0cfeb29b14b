// fd_plain.hip -- the per-call drop-in for ucb_waveform
// (ucb/src/glass_ucb_waveform.c:738-1001): one workgroup builds one template
// (slow series -> 6 LDS FFTs -> TDI) and writes X,Y,Z,A,E.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "fd_waveform.h"
#include "runtime.h"
#include "wdm_common.h"

namespace glass {

// out: [ntmpl][5][N] complex (X,Y,Z,A,E), template t reads params + t*np
template <int N, int NTHR>
__global__ __launch_bounds__(NTHR) void fd_waveform_kernel(const double* __restrict__ params, int np,
                                                           double T, double t0, double L, double fstar,
                                                           const double* __restrict__ pos, int posstride,
                                                           const double2* __restrict__ tw, int twstride,
                                                           int fmt, double2* __restrict__ out)
{
    using S = FdShape<N, NTHR>;
    __shared__ double2 lds[S::LDS];
    const UcbConst& c = ucb_setup_shared(params + (size_t)blockIdx.x * np, np, T, t0, L, fstar);
    double2 d[S::PER][6];
    fd_link_spectra<N, NTHR>(c, lds, pos, posstride, tw, twstride, d);
    double2* o = out + (size_t)blockIdx.x * 5 * N;
#pragma unroll
    for (int k = 0; k < S::PER; k++)
    {
        const int m = threadIdx.x + k * NTHR;
        if (m >= N) continue;
        double2 v[5];
        fd_tdi<N>(c, m, fmt, d[k], v);
#pragma unroll
        for (int ch = 0; ch < 5; ch++) o[ch * N + m] = v[ch];
    }
}

template <int N>
static void launch_fd(int ntmpl, const double* params, int np, double T, double t0, double L,
                      double fstar, const double* pos, int posstride, const double2* tw, int fmt,
                      double2* out, hipStream_t s)
{
    constexpr int NTHR = fd_threads(N);
    fd_waveform_kernel<N, NTHR><<<ntmpl, NTHR, 0, s>>>(params, np, T, t0, L, fstar, pos, posstride, tw,
                                                        kTwN / N, fmt, out);
}

// The template kernel over ntmpl templates with the positions table pos
// (9 doubles per slow sample, BW samples).
static int fd_waveform_launch_pos(int ntmpl, const double* d_params, int np, double T, double t0, double L,
                                  double fstar, const double* pos, int BW, int fmt, double2* d_out, hipStream_t s)
{
    const double2* tw = nullptr;
    int rc;
    if ((rc = twiddles(&tw))) return rc;
    switch (BW)
    {
#define GH_CASE(n) \
    case n: launch_fd<n>(ntmpl, d_params, np, T, t0, L, fstar, pos, 1, tw, fmt, d_out, s); break;
        GH_CASE(16) GH_CASE(32) GH_CASE(64) GH_CASE(128) GH_CASE(256) GH_CASE(512) GH_CASE(1024) GH_CASE(2048) GH_CASE(4096)
#undef GH_CASE
        default: set_error("BW=%d unsupported (power of two 16..4096)", BW); return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

int fd_waveform_launch(int ntmpl, const double* d_params, int np, double T, double t0,
                       const glass_hip_orbit& o, int BW, int fmt, double2* d_out, hipStream_t s)
{
    const double* pos = nullptr;
    if (int rc = positions(o, T, t0, BW, &pos)) return rc;
    return fd_waveform_launch_pos(ntmpl, d_params, np, T, t0, o.L, o.fstar, pos, BW, fmt, d_out, s);
}

namespace wdm {
void spline_host(int N, const double* x, const double* y, double* c);  // wdm_ucb.hip
}

// interpolate_orbits (utils/src/glass_lisa.c:62-72) at the npos slow-sample
// times t0 + T n / npos: the natural cubic spline of every coordinate over
// the orbit's knots (spline_coefficients, utils/src/glass_math.c:98-178) and
// spline_interpolation's binary search and cubic (:180-208), on the host in
// the reference's operation order; the table is cached on the device per
// (device, orbit data, T, t0, npos).
static int numeric_positions(const glass_hip_numeric_orbit& o, double T, double t0, int npos, const double** out)
{
    static std::mutex mu;
    static std::map<std::vector<double>, double*> cache;
    const int N = o.Norb;
    int dev = 0;
    GH_CHECK(hipGetDevice(&dev));
    std::vector<double> key;
    key.reserve(5 + 10 * (size_t)N);
    key.push_back(dev);
    key.push_back(N);
    key.push_back(T);
    key.push_back(t0);
    key.push_back(npos);
    key.insert(key.end(), o.t, o.t + N);
    for (const double* a : {o.x, o.y, o.z}) key.insert(key.end(), a, a + 3 * (size_t)N);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end())
    {
        *out = it->second;
        return GLASS_HIP_OK;
    }
    for (int n = 1; n < N; n++)
        if (!(o.t[n] > o.t[n - 1]))
        {
            set_error("numeric orbit: knots must increase (t[%d] = %g after %g)", n, o.t[n], o.t[n - 1]);
            return GLASS_HIP_EINVAL;
        }
    // coefficients per (spacecraft, coordinate): c[(3 sc + d)][4 N]
    std::vector<double> coef(36 * (size_t)N);
    for (int sc = 0; sc < 3; sc++)
        for (int d = 0; d < 3; d++)
        {
            const double* y = (d == 0 ? o.x : (d == 1 ? o.y : o.z)) + (size_t)sc * N;
            wdm::spline_host(N, o.t, y, coef.data() + 4 * (size_t)N * (3 * sc + d));
        }
    std::vector<double> h(9 * (size_t)npos);
    for (int n = 0; n < npos; n++)
    {
        const double t = ucb_sample_time(T, t0, n, npos);
        if (t < o.t[0] || t > o.t[N - 1])
        {
            set_error("numeric orbit: slow-sample time %g outside the knots [%g, %g]", t, o.t[0], o.t[N - 1]);
            return GLASS_HIP_EINVAL;
        }
        const int k = wdm::knot_index(o.t, N, t);
        const double dx = t - o.t[k];
        // table layout x1 x2 x3 y1 y2 y3 z1 z2 z3
        for (int sc = 0; sc < 3; sc++)
            for (int d = 0; d < 3; d++)
                h[9 * (size_t)n + 3 * d + sc] = wdm::spline_at(coef.data() + 4 * (size_t)N * (3 * sc + d), N, k, dx);
    }
    double* dptr = nullptr;
    GH_CHECK(hipMalloc(&dptr, sizeof(double) * h.size()));
    GH_CHECK(hipMemcpy(dptr, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    cache[key] = dptr;
    *out = dptr;
    return GLASS_HIP_OK;
}

int parse_format(const char* format)
{
    if (!format) return -1;
    if (strcmp(format, "phase") == 0) return TDI_PHASE;
    if (strcmp(format, "frequency") == 0) return TDI_FREQUENCY;
    if (strcmp(format, "sangria") == 0) return TDI_SANGRIA;
    return -1;
}

}  // namespace glass

namespace glass {

// The drop-in's host side for either orbit kind (one of ao / no non-NULL).
static int ucb_waveform_dropin(const glass_hip_orbit* ao, const glass_hip_numeric_orbit* no, const char* format,
                               double T, double t0, const double* params, int NParams, double* X, double* Y,
                               double* Z, double* A, double* E, int BW, int NI)
{
    int fmt = parse_format(format);
    if (fmt < 0)
    {
        set_error("Unsupported data format %s", format ? format : "(null)");
        return GLASS_HIP_EINVAL;
    }
    if ((!ao && !no) || !params || !X || NParams < 7 || NParams > 9)
    {
        set_error("ucb_waveform: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    if (no && (no->Norb < 4 || !no->t || !no->x || !no->y || !no->z || !(no->L > 0) || !(no->fstar > 0)))
    {
        set_error("ucb_waveform_numeric: bad orbit (Norb %d, L %g, fstar %g)", no->Norb, no->L, no->fstar);
        return GLASS_HIP_EINVAL;
    }
    const bool all = (fmt == TDI_SANGRIA) || NI > 1;
    if (all && (!Y || !Z || !A || !E))
    {
        set_error("ucb_waveform: Y,Z,A,E required");
        return GLASS_HIP_EINVAL;
    }
    ThreadCtx* cx;
    int rc;
    if ((rc = ctx_get(&cx))) return rc;
    const size_t outb = sizeof(double2) * 5 * (size_t)BW;
    if ((rc = cx->dbuf.ensure(outb + 128))) return rc;
    if ((rc = cx->hbuf.ensure(outb + 128))) return rc;
    double* hp = (double*)cx->hbuf.p;
    double* dp = (double*)((char*)cx->dbuf.p + outb);
    memcpy(hp + 10 * BW, params, sizeof(double) * NParams);
    GH_CHECK(hipMemcpyAsync(dp, hp + 10 * BW, sizeof(double) * NParams, hipMemcpyHostToDevice, cx->stream));
    if (ao)
        rc = fd_waveform_launch(1, dp, NParams, T, t0, *ao, BW, fmt, (double2*)cx->dbuf.p, cx->stream);
    else
    {
        const double* pos = nullptr;
        if (!(rc = numeric_positions(*no, T, t0, BW, &pos)))
            rc = fd_waveform_launch_pos(1, dp, NParams, T, t0, no->L, no->fstar, pos, BW, fmt, (double2*)cx->dbuf.p,
                                        cx->stream);
    }
    if (rc) return rc;
    GH_CHECK(hipMemcpyAsync(hp, cx->dbuf.p, outb, hipMemcpyDeviceToHost, cx->stream));
    GH_CHECK(hipStreamSynchronize(cx->stream));
    double* outs[5] = {X, Y, Z, A, E};
    for (int ch = 0; ch < (all ? 5 : 1); ch++) memcpy(outs[ch], hp + (size_t)ch * 2 * BW, sizeof(double) * 2 * BW);
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_ucb_bandwidth(double L, double fstar, double f, double fdot, double costheta, double A,
                                       double T, int N)
{
    return glass::ucb_bandwidth(L, fstar, f, fdot, costheta, A, T, N);
}

extern "C" int glass_hip_ucb_alignment(const glass_hip_orbit* orbit, double T, int NFFT, int data_qmin,
                                       const double* params, int* BW, int* qmin, int* imin)
{
    if (!orbit || !params || !BW || !qmin || !imin || NFFT < 2)
    {
        glass::set_error("ucb_alignment: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    // map_array_to_ucb_params (ucb/src/glass_ucb_model.c:25-38)
    double f0 = params[0] / T;
    double costheta = params[1];
    double amp = exp(params[3]);
    double dfdt = params[7] / (T * T);
    *BW = 2 * glass::ucb_bandwidth(orbit->L, orbit->fstar, f0, dfdt, costheta, amp, T, NFFT);
    *qmin = (int)(f0 * T) - *BW / 2;
    *imin = *qmin - data_qmin;
    return GLASS_HIP_OK;
}

}  // namespace glass

extern "C" int glass_hip_ucb_waveform(const glass_hip_orbit* orbit, const char* format, double T,
                                      double t0, const double* params, int NParams, double* X,
                                      double* Y, double* Z, double* A, double* E, int BW, int NI)
{
    return glass::ucb_waveform_dropin(orbit, nullptr, format, T, t0, params, NParams, X, Y, Z, A, E, BW, NI);
}

extern "C" int glass_hip_ucb_waveform_numeric(const glass_hip_numeric_orbit* orbit, const char* format, double T,
                                              double t0, const double* params, int NParams, double* X, double* Y,
                                              double* Z, double* A, double* E, int BW, int NI)
{
    if (!orbit)
    {
        glass::set_error("ucb_waveform_numeric: NULL orbit");
        return GLASS_HIP_EINVAL;
    }
    return glass::ucb_waveform_dropin(nullptr, orbit, format, T, t0, params, NParams, X, Y, Z, A, E, BW, NI);
}

extern "C" int glass_hip_numeric_orbit_armlength(int Norb, const double* x, const double* y, const double* z,
                                                 double* L, double* fstar)
{
    if (Norb < 1 || !x || !y || !z || !L || !fstar)
    {
        glass::set_error("numeric_orbit_armlength: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    // initialize_numeric_orbit (utils/src/glass_lisa.c:207-259)
    double L12 = 0.0, L23 = 0.0, L31 = 0.0;
    for (int n = 0; n < Norb; n++)
    {
        const double x1 = x[n], x2 = x[Norb + n], x3 = x[2 * Norb + n];
        const double y1 = y[n], y2 = y[Norb + n], y3 = y[2 * Norb + n];
        const double z1 = z[n], z2 = z[Norb + n], z3 = z[2 * Norb + n];
        L12 += sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
        L23 += sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2) + (z3 - z2) * (z3 - z2));
        L31 += sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
    }
    L12 /= (double)Norb;
    L31 /= (double)Norb;
    L23 /= (double)Norb;
    *L = (L12 + L31 + L23) / 3.;
    *fstar = 299792458. / (2.0 * M_PI * *L);
    return GLASS_HIP_OK;
}
