// fisher_fd.hip -- drop-in for ucb_fisher (ucb/src/glass_ucb_waveform.c:
// 364-522, up to the LAPACKE eigensolve): the reference's forward-difference
// Fisher matrix, reproduced with its conventions so a caller sees the same
// matrix:
//   * step eps = 1e-6 on each entry of the parameter array (:369, :397);
//   * cos(theta) and cos(iota) clipped at +1 after the step (:411-415);
//   * each perturbed template re-aligned with ucb_alignment, so its BW may
//     differ from the unperturbed one (:437-441);
//   * dh_i over 2*BW_i doubles against the unperturbed template (zero beyond
//     its own BW) (:448-466);
//   * Gram sum of the nine channel-pair fourier_nwip terms over the BW of
//     the LAST perturbed template, with invC read from window bin 0 while
//     dh is template-local (:473-496);
//   * NaN entries replaced by 10 (:498-507).
// The nine templates are built by the plain waveform kernel; one workgroup
// per matrix entry reduces the Gram sum.
#include <cmath>
#include <cstring>
#include <vector>

#include "fd_waveform.h"
#include "reduce.h"
#include "runtime.h"

namespace glass {

int fd_waveform_launch(int ntmpl, const double* d_params, int np, double T, double t0, const glass_hip_orbit& o,
                       int BW, int fmt, double2* d_out, hipStream_t s);
int parse_format(const char* format);

// h: template t at h + t*5*bwcap, channel a at + a*bw[t] (plain kernel
// layout).  Entry (i,j), i <= j, of the 8x8 matrix per workgroup.
__global__ __launch_bounds__(256) void fisher_fd_gram_kernel(const double2* __restrict__ h, const int* __restrict__ bw,
                                                             int bwcap, const double* __restrict__ invC, int NFFT,
                                                             double inv_eps, double* __restrict__ F)
{
    int pair = blockIdx.x, i = 0;
    while (pair >= 8 - i)
    {
        pair -= 8 - i;
        i++;
    }
    const int j = i + pair;
    const int bw0 = bw[0], bwi = bw[1 + i], bwj = bw[1 + j], bwp = bw[8];
    const double2* h0 = h;
    const double2* hi = h + (size_t)(1 + i) * 5 * bwcap;
    const double2* hj = h + (size_t)(1 + j) * 5 * bwcap;
    double acc = 0.0;
    for (int n = threadIdx.x; n < bwp; n += 256)
    {
        double2 di[3], dj[3];
        for (int a = 0; a < 3; a++)
        {
            double2 z0 = n < bw0 ? h0[a * bw0 + n] : make_double2(0.0, 0.0);
            double2 zi = n < bwi ? hi[a * bwi + n] : make_double2(0.0, 0.0);
            double2 zj = n < bwj ? hj[a * bwj + n] : make_double2(0.0, 0.0);
            // dh_i is written over 2*BW_i doubles only (:460-466)
            di[a] = n < bwi ? make_double2((zi.x - z0.x) * inv_eps, (zi.y - z0.y) * inv_eps) : make_double2(0.0, 0.0);
            dj[a] = n < bwj ? make_double2((zj.x - z0.x) * inv_eps, (zj.y - z0.y) * inv_eps) : make_double2(0.0, 0.0);
        }
        double s = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                s += (di[a].x * dj[b].x + di[a].y * dj[b].y) * invC[(size_t)(3 * a + b) * NFFT + n];
        acc += 2.0 * s;  // fourier_nwip's factor 2
    }
    acc = block_sum<256>(acc);
    if (threadIdx.x == 0)
    {
        if (acc != acc) acc = 10.0;
        F[i * 8 + j] = acc;
        F[j * 8 + i] = acc;
    }
}

}  // namespace glass

using namespace glass;

extern "C" int glass_hip_ucb_fisher(const glass_hip_orbit* orbit, const char* format, double T, double t0,
                                    int NFFT, int data_qmin, const double* params, const double* invC,
                                    double* fisher)
{
    int fmt = parse_format(format);
    if (fmt < 0)
    {
        set_error("Unsupported data format %s", format ? format : "(null)");
        return GLASS_HIP_EINVAL;
    }
    if (!orbit || !params || !invC || !fisher || NFFT < 32)
    {
        set_error("ucb_fisher: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    const double eps = 1.0e-6;
    double P[9][8];
    int bw[9];
    for (int t = 0; t < 9; t++)
    {
        for (int k = 0; k < 8; k++) P[t][k] = params[k];
        if (t > 0)
        {
            const int i = t - 1;
            P[t][i] += eps;
            if ((i == 1 || i == 4) && P[t][i] > 1.0) P[t][i] = 1.0;
        }
        int qmin, imin, rc;
        if ((rc = glass_hip_ucb_alignment(orbit, T, NFFT, data_qmin, P[t], &bw[t], &qmin, &imin))) return rc;
        if (bw[t] > NFFT)
        {
            set_error("ucb_fisher: BW %d exceeds NFFT %d", bw[t], NFFT);
            return GLASS_HIP_EINVAL;
        }
    }
    int bwcap = 0;
    for (int t = 0; t < 9; t++) bwcap = bw[t] > bwcap ? bw[t] : bwcap;
    ThreadCtx* cx;
    int rc;
    if ((rc = ctx_get(&cx))) return rc;
    const size_t hbytes = sizeof(double2) * 9 * 5 * (size_t)bwcap;
    const size_t ibytes = sizeof(double) * 9 * (size_t)NFFT;
    const size_t tot = hbytes + ibytes + 1024 + sizeof(P) + 64 * sizeof(double);
    if ((rc = cx->dbuf.ensure(tot)) || (rc = cx->hbuf.ensure(tot))) return rc;
    char* db = (char*)cx->dbuf.p;
    char* hb = (char*)cx->hbuf.p;
    double2* d_h = (double2*)db;
    double* d_ic = (double*)(db + hbytes);
    double* d_P = (double*)(db + hbytes + ibytes);
    int* d_bw = (int*)(db + hbytes + ibytes + sizeof(P));
    double* d_F = (double*)(db + hbytes + ibytes + sizeof(P) + 512);
    memcpy(hb + hbytes, invC, ibytes);
    memcpy(hb + hbytes + ibytes, P, sizeof(P));
    memcpy(hb + hbytes + ibytes + sizeof(P), bw, sizeof(bw));
    GH_CHECK(hipMemcpyAsync(d_ic, hb + hbytes, ibytes + sizeof(P) + sizeof(bw), hipMemcpyHostToDevice, cx->stream));
    for (int t = 0; t < 9; t++)
        if ((rc = fd_waveform_launch(1, d_P + 8 * t, 8, T, t0, *orbit, bw[t], fmt, d_h + (size_t)t * 5 * bwcap,
                                     cx->stream)))
            return rc;
    fisher_fd_gram_kernel<<<36, 256, 0, cx->stream>>>(d_h, d_bw, bwcap, d_ic, NFFT, 1.0 / eps, d_F);
    GH_CHECK(hipGetLastError());
    GH_CHECK(hipMemcpyAsync(hb, d_F, 64 * sizeof(double), hipMemcpyDeviceToHost, cx->stream));
    GH_CHECK(hipStreamSynchronize(cx->stream));
    memcpy(fisher, hb, 64 * sizeof(double));
    return GLASS_HIP_OK;
}
