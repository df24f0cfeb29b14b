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
    __shared__ double2 lds[6 * N];
    UcbConst c;
    ucb_setup(params + (size_t)blockIdx.x * np, np, T, t0, L, fstar, c);
    fd_slow_fft<N, NTHR>(c, lds, pos, posstride, tw, twstride);
    double2* o = out + (size_t)blockIdx.x * 5 * N;
    for (int m = threadIdx.x; m < N; m += NTHR)
    {
        double2 v[5];
        fd_tdi_bin<N>(c, lds, m, fmt, v);
#pragma unroll
        for (int ch = 0; ch < 5; ch++) o[ch * N + m] = v[ch];
    }
}

template <int N>
static void launch_fd(int ntmpl, const double* params, int np, double T, double t0, double L,
                      double fstar, const double* pos, int posstride, const double2* tw, int fmt,
                      double2* out, hipStream_t s)
{
    constexpr int NTHR = N < 64 ? 64 : (N > 1024 ? 1024 : N);
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
        GH_CASE(16) GH_CASE(32) GH_CASE(64) GH_CASE(128) GH_CASE(256) GH_CASE(512) GH_CASE(1024)
#undef GH_CASE
        default: set_error("BW=%d unsupported (power of two 16..1024)", BW); return GLASS_HIP_EINVAL;
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
