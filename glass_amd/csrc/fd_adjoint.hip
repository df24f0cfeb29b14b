// fd_adjoint.hip -- reverse-mode (adjoint) gradient of one galactic binary's
// likelihood contribution, for many sources at once (no reference
// counterpart: the reference differences whole waveforms in ucb_fisher,
// ucb/src/glass_ucb_waveform.c:364-522).
//
// With L = (R|h) - (h|h)/2 (R the residual without the source) and h the
// TDI of the six links' spectra d_l = FFT(s_l)/(2N) (fftshifted), the
// gradient is
//   dL/dtheta = (R - h | dh/dtheta) = 2 Re sum_n sum_l conj(ds_l[n]/dtheta) Sb_l[n]
// where the adjoint of the slow series Sb is obtained from the residual in
// three linear steps run backwards:
//   G_b[m]  = sum_a w_ab[j] (R - h)_a[m]            (weighted residual, per bin)
//   Db_l[m] = sum_b conj(T_bl[m]) G_b[m]            (TDI transposed)
//   Sb_l[n] = (1/2N) sum_m exp(+2 pi i n s(m)/N) Db_l[m]
//           = conj(FFT(conj(Db_l at index s(m))))[n] / (2N)   (FFT transposed)
// One workgroup per source: the forward pass (slow series, six LDS FFTs,
// TDI) gives h and L; the six adjoint series go through the same LDS
// buffer and FFT; a last pass recomputes each slow sample with its four
// tangents (f0 T, cos theta, phi, fdot T^2) and plus / cross basis
// (fd_tangent.h) and contracts them with Sb.  The four parameters that
// enter linearly (ln A, cos iota, psi, phi0) follow from the basis sums.
// Cost: one waveform, six more FFTs and one tangent pass over the samples
// (the forward-mode Fisher kernel runs six series sets through FFT and TDI).
#include <cmath>

#include "../../include/glass_hip_batch.h"
#include "fd_derivs.h"
#include "fd_tangent.h"
#include "reduce.h"
#include "runtime.h"

namespace glass {

int parse_format(const char* format);

namespace {

constexpr int kNAdj = 9;  // L, four tangent gradients, basis sums P0, P1 (complex)

// TDI of bin m as channel = alpha a + beta b with a = d_ij - d_ik and
// b = d_ji - d_ki (fd_tdi's combinations): Sangria pf [a e^{-2ix} + b e^{-ix}],
// phase / frequency pf [a (e^{-3ix} - e^{-ix}) + b (e^{-2ix} - 1)].
template <int N>
__device__ inline void tdi_coef(const UcbConst& c, int m, int fmt, double2& alpha, double2& beta)
{
    const double f = ((double)(c.q + m - N / 2)) * c.invT;
    const double fonfs = f * c.invfstar;
    double s1, c1;
    sincos_red(fonfs, &s1, &c1);
    const double c2 = 2. * c1 * c1 - 1.0, s2 = 2. * s1 * c1;
    if (fmt == TDI_SANGRIA)
    {
        const double pf = f * (4.0 * c.invfstar * c.sqT) * s1;
        alpha = make_double2(pf * c2, -pf * s2);
        beta = make_double2(pf * c1, -pf * s1);
    }
    else
    {
        const double c3 = 2. * c1 * c2 - c1, s3 = 2. * s1 * c2 + s1;
        const double pf = (fmt == TDI_PHASE) ? c.sqT : c.sqT * (2. * fonfs);
        alpha = make_double2(pf * (c3 - c1), -pf * (s3 - s1));
        beta = make_double2(pf * (c2 - 1.0), -pf * s2);
    }
}

__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// conj(a) b
__device__ inline double2 cjmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

template <int N, int NTHR>
__global__ __launch_bounds__(NTHR) void adjoint_kernel(DerivArgs A, int want_grad)
{
    using S = FdShape<N, NTHR>;
    __shared__ double2 lds[6 * N];
    __shared__ UcbDiff sdd;
    const int e = blockIdx.x;
    const int t = threadIdx.x;
    const double* p = A.params + (size_t)(A.pidx ? A.pidx[e] : e) * 8;
    // the template constants and their derivatives on lanes of two waves,
    // so the two serial setups overlap (one wave at 32/64 bins: in turn)
    __shared__ UcbConst sc;
    if (t == 0) ucb_setup(p, 8, A.T, A.t0, A.L, A.fstar, sc);
    if (t == (NTHR >= 128 ? 64 : 0)) ucb_diff_setup(p, A.T, sdd);
    __syncthreads();
    UcbConst cr = sc;  // into scalar registers for the passes
    make_uniform(cr);
    const UcbConst& c = cr;
    const int posstride = 4096 / N;
    // forward: the six link spectra of this thread's bins (ends with a
    // barrier)
    double2 d[S::PER][6];
    fd_link_spectra<N, NTHR>(c, lds, A.pos, posstride, A.tw, kTwN / N, d);
    const UcbDiff& dd = sdd;
    const double2* R = A.iroff ? A.resid + A.iroff[e] : A.resid + (size_t)e * (A.rstride / 2);
    const double* iC = A.iioff ? A.invC + A.iioff[e] : A.invC + (size_t)e * A.istride;
    const int W = A.iW ? A.iW[e] : A.W, imin = A.imin[e];
    double acc[kNAdj];
#pragma unroll
    for (int q = 0; q < kNAdj; q++) acc[q] = 0.0;
#pragma unroll
    for (int k = 0; k < S::PER; k++)
    {
        const int m = t + k * NTHR;
        if (m >= N) continue;
        const int j = imin + m;
        const bool used = (A.mode == 0 || m < W) && j >= 0 && j < W;
        double2 G[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
        if (used)
        {
            double2 h[5];
            fd_tdi<N>(c, m, A.fmt, d[k], h);
            double w[6];
#pragma unroll
            for (int q = 0; q < 6; q++) w[q] = iC[(size_t)q * W + j];
            double2 r[3];
#pragma unroll
            for (int ch = 0; ch < 3; ch++) r[ch] = A.resid ? R[(size_t)ch * W + j] : make_double2(0.0, 0.0);
            auto re = [](double2 a, double2 b) { return a.x * b.x + a.y * b.y; };
            auto bf3 = [&](const double2* u, const double2* v) {
                return 2.0 * (w[0] * re(u[0], v[0]) + w[1] * re(u[1], v[1]) + w[2] * re(u[2], v[2]) +
                              w[3] * (re(u[0], v[1]) + re(u[1], v[0])) + w[4] * (re(u[0], v[2]) + re(u[2], v[0])) +
                              w[5] * (re(u[1], v[2]) + re(u[2], v[1])));
            };
            acc[0] += bf3(r, h) - 0.5 * bf3(h, h);
            const double2 x0 = csub(r[0], h[0]), x1 = csub(r[1], h[1]), x2 = csub(r[2], h[2]);
            // G_b = sum_a w_ab (R - h)_a, w symmetric: XX,YY,ZZ,XY,XZ,YZ
            G[0] = make_double2(w[0] * x0.x + w[3] * x1.x + w[4] * x2.x, w[0] * x0.y + w[3] * x1.y + w[4] * x2.y);
            G[1] = make_double2(w[3] * x0.x + w[1] * x1.x + w[5] * x2.x, w[3] * x0.y + w[1] * x1.y + w[5] * x2.y);
            G[2] = make_double2(w[4] * x0.x + w[5] * x1.x + w[2] * x2.x, w[4] * x0.y + w[5] * x1.y + w[2] * x2.y);
        }
        if (want_grad)
        {
            double2 al, be;
            tdi_coef<N>(c, m, A.fmt, al, be);
            // X = al (d12 - d13) + be (d21 - d31), Y = al (d23 - d21) + be (d32 - d12),
            // Z = al (d31 - d32) + be (d13 - d23); Db_l = sum_b conj(T_bl) G_b
            const double2 aX = cjmul(al, G[0]), bX = cjmul(be, G[0]);
            const double2 aY = cjmul(al, G[1]), bY = cjmul(be, G[1]);
            const double2 aZ = cjmul(al, G[2]), bZ = cjmul(be, G[2]);
            double2 Db[6];
            Db[0] = csub(aX, bY);  // 12
            Db[1] = csub(bX, aY);  // 21
            Db[2] = csub(aZ, bX);  // 31
            Db[3] = csub(bZ, aX);  // 13
            Db[4] = csub(aY, bZ);  // 23
            Db[5] = csub(bY, aZ);  // 32
            const int s = (m + N / 2) & (N - 1);
#pragma unroll
            for (int l = 0; l < 6; l++) lds[l * N + s] = cconj(Db[l]);
        }
    }
    if (want_grad)
    {
        __syncthreads();
        lds_fft<N, 6, NTHR>(lds, A.tw, kTwN / N);
        const double inv2N = 1. / (double)(2 * N);
        for (int n = t; n < N; n += NTHR)
        {
            const double tn = ucb_sample_time(c.T, c.t0, n, N);
            ucb_sample_sets_to<0, 6>(c, dd, tn, A.pos + (size_t)n * posstride * 9, [&](int q, int link, double2 v) {
                // Sb_link[n] = conj(FFT value) / (2N), read where it is used
                const double2 f = lds[link * N + n];
                const double2 sb = make_double2(f.x * inv2N, -f.y * inv2N);
                const double2 pr = cjmul(v, sb);  // conj(ds) Sb
                if (q == 0)
                {
                    acc[5] += pr.x;
                    acc[6] += pr.y;
                }
                else if (q == 1)
                {
                    acc[7] += pr.x;
                    acc[8] += pr.y;
                }
                else
                    acc[q - 1] += pr.x;  // q = 2..5 -> acc[1..4]
            });
        }
    }
    // workgroup sums of the nine partials
#pragma unroll
    for (int q = 0; q < kNAdj; q++) acc[q] = block_sum<NTHR < 64 ? 64 : NTHR>(acc[q]);
    if (t == 0)
    {
        A.outL[e] = acc[0];
        if (want_grad)
        {
            const double2 P0 = make_double2(acc[5], acc[6]), P1 = make_double2(acc[7], acc[8]);
            // s = DP s0 + DC s1: conj(ds) Sb summed = conj(dDP) P0 + conj(dDC) P1
            auto lin = [&](double2 a, double2 b) { return cadd(cjmul(a, P0), cjmul(b, P1)); };
            const double2 sP = lin(make_double2(c.DPr, c.DPi), make_double2(c.DCr, c.DCi));
            double* g = A.grad + (size_t)e * 8;
            g[0] = 2.0 * acc[1];                                                    // f0 T
            g[1] = 2.0 * acc[2];                                                    // cos(theta)
            g[2] = 2.0 * acc[3];                                                    // phi
            g[3] = 2.0 * sP.x;                                                      // ln A: ds = s
            g[4] = 2.0 * lin(make_double2(dd.DPr4, dd.DPi4), make_double2(dd.DCr4, dd.DCi4)).x;  // cos(iota)
            g[5] = 2.0 * lin(make_double2(dd.DPr5, dd.DPi5), make_double2(dd.DCr5, dd.DCi5)).x;  // psi
            g[6] = -2.0 * sP.y;                                                     // phi0: ds = -i s
            g[7] = 2.0 * acc[4];                                                    // fdot T^2
        }
    }
}

}  // namespace

int launch_adjoint(int N, const DerivArgs& A, int n, bool want_grad, hipStream_t s)
{
    if (n <= 0) return GLASS_HIP_OK;
    switch (N)
    {
#define GH_A(NN)                                                                                   \
    case NN:                                                                                       \
        adjoint_kernel<NN, fd_threads(NN)><<<n, fd_threads(NN), 0, s>>>(A, want_grad ? 1 : 0);    \
        break;
        GH_A(32) GH_A(64) GH_A(128) GH_A(256) GH_A(512) GH_A(1024)
#undef GH_A
        default: set_error("fd adjoint: BW=%d unsupported (power of two 32..1024)", N); return GLASS_HIP_EINVAL;
    }
    GH_CHECK(hipGetLastError());
    return GLASS_HIP_OK;
}

}  // namespace glass

using namespace glass;

extern "C" int glass_hip_fd_source_grad(const glass_hip_orbit* orbit, const char* format, double T, double t0, int n,
                                        const double* d_params, int BW, const int* d_imin, int W,
                                        const double* d_resid, long long resid_stride, const double* d_invC,
                                        long long invC_stride, int mode, double* d_logL, double* d_grad,
                                        void* stream)
{
    int fmt = parse_format(format);
    if (fmt < 0)
    {
        set_error("Unsupported data format %s", format ? format : "(null)");
        return GLASS_HIP_EINVAL;
    }
    if (!orbit || n < 0 || !d_params || !d_imin || W < 1 || !d_invC || !d_logL || (mode != 0 && mode != 1) ||
        resid_stride < 0 || invC_stride < 0)
    {
        set_error("fd_source_grad: bad arguments");
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
    return launch_adjoint(BW, A, n, d_grad != nullptr, (hipStream_t)stream);
}
