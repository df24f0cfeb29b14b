// fd_plain.hip -- the per-call drop-in for ucb_waveform
// (ucb/src/glass_ucb_waveform.c:738-1001): one workgroup builds one template
// (slow series -> 6 LDS FFTs -> TDI) and writes X,Y,Z,A,E.
#include <cstring>

#include "fd_waveform.h"
#include "runtime.h"

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

int fd_waveform_launch(int ntmpl, const double* d_params, int np, double T, double t0,
                       const glass_hip_orbit& o, int BW, int fmt, double2* d_out, hipStream_t s)
{
    const double* pos = nullptr;
    const double2* tw = nullptr;
    int rc;
    if ((rc = twiddles(&tw))) return rc;
    if ((rc = positions(o, T, t0, BW, &pos))) return rc;
    switch (BW)
    {
#define GH_CASE(n) \
    case n: launch_fd<n>(ntmpl, d_params, np, T, t0, o.L, o.fstar, pos, 1, tw, fmt, d_out, s); break;
        GH_CASE(16) GH_CASE(32) GH_CASE(64) GH_CASE(128) GH_CASE(256) GH_CASE(512) GH_CASE(1024) GH_CASE(2048) GH_CASE(4096)
#undef GH_CASE
        default: set_error("BW=%d unsupported (power of two 16..4096)", BW); return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipGetLastError());
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

extern "C" int glass_hip_ucb_waveform(const glass_hip_orbit* orbit, const char* format, double T,
                                      double t0, const double* params, int NParams, double* X,
                                      double* Y, double* Z, double* A, double* E, int BW, int NI)
{
    using namespace glass;
    int fmt = parse_format(format);
    if (fmt < 0)
    {
        set_error("Unsupported data format %s", format ? format : "(null)");
        return GLASS_HIP_EINVAL;
    }
    if (!orbit || !params || !X || NParams < 7 || NParams > 9)
    {
        set_error("ucb_waveform: bad arguments");
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
    if ((rc = fd_waveform_launch(1, dp, NParams, T, t0, *orbit, BW, fmt, (double2*)cx->dbuf.p, cx->stream)))
        return rc;
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
