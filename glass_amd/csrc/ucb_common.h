// ucb_common.h -- shared device code for the UCB frequency-domain path on
// gfx950 (CDNA4).  FP64 throughout.
//
// The slow-series sample restates ucb/src/glass_ucb_waveform.c:840-937 with
// the LISA response helpers of utils/src/glass_lisa.c:1521-1596.  The phase
// chain t -> k.x -> xi -> arg2 is evaluated with contraction disabled and in
// the reference's operation order, because arg2 is a difference of ~1e6-1e7
// rad products whose rounding (~1e-10 rad) would otherwise move the result
// by far more than the 1e-12 parity budget.  Amplitude-only quantities
// (detector tensors, transfer-function sinc) are free to contract.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fastmath.h"

namespace glass {

// utils/src/glass_constants.h:31-89, utils/src/glass_lisa.h:31-49
constexpr double kPI2 = 6.283185307179586;
constexpr double kPI = 3.14159265358979323846;
constexpr double kCLIGHT = 299792458.;
constexpr double kInvCLIGHT = 1.0 / kCLIGHT;  // correctly rounded (compile time)
constexpr double kYEAR = 3.15581497632e7;
constexpr double kAU = 1.49597870700e11;
constexpr double kSQ3 = 1.73205080757;
constexpr double kLARM = 2.5e9;
constexpr double kSQRT1_2 = 0.70710678118654752440;
constexpr double kINVSQ6 = 0.408248290463863;  // glass_lisa.c:436

// Per-template constants (glass_ucb_waveform.c:804-837).
struct UcbConst
{
    double T, t0;
    double f0, phi0, dfdt, d2fdt2;
    double df;      // PI2*(q/T)
    double pi2f0;   // PI2*f0, first factor of arg2
    double pidfdt;  // M_PI*dfdt
    double pid2;    // (M_PI/3)*d2fdt2
    double aevc;    // 0.666...*dfdt/f0
    double halfd2;  // 0.5*d2fdt2
    double DPr, DPi, DCr, DCi;
    double k[3];
    double ep[6], ec[6];  // symmetric tensors: xx, yy, zz, xy, xz, yz
    double invfstar, fstar, invL;
    double sqT, invT;  // sqrt(T), 1/T (TDI prefactor and bin frequency)
    long q;
    int np;
};

// glass_ucb_waveform.c:804-837 + LISA_polarization_tensor (glass_lisa.c:1566-1596)
__host__ __device__ inline void ucb_setup(const double* p, int np, double T, double t0,
                                          double L, double fstar, UcbConst& c)
{
    // k feeds the phase through k.x: no contraction anywhere in the setup
#pragma clang fp contract(off)
    c.T = T;
    c.t0 = t0;
    c.np = np;
    c.f0 = p[0] / T;
    double costh = p[1], phi = p[2], amp = exp(p[3]), cosi = p[4], psi = p[5];
    c.phi0 = p[6];
    c.dfdt = 0.0;
    c.d2fdt2 = 0.0;
    if (np > 7) c.dfdt = p[7] / (T * T);
    if (np > 8) c.d2fdt2 = p[8] / (T * T * T);
    c.q = (long)(c.f0 * T);
    double cosps = cos(2. * psi), sinps = sin(2. * psi);
    double Aplus = amp * (1. + cosi * cosi) * 0.5;
    double Across = -amp * (cosi);
    c.df = kPI2 * (((double)c.q) / T);
    c.pi2f0 = kPI2 * c.f0;
    c.pidfdt = kPI * c.dfdt;
    c.pid2 = (kPI / 3.0) * c.d2fdt2;
    c.aevc = 0.66666666666666666666 * c.dfdt / c.f0;
    c.halfd2 = 0.5 * c.d2fdt2;
    c.DPr = Aplus * cosps;
    c.DPi = -Across * sinps;
    c.DCr = -Aplus * sinps;
    c.DCi = -Across * cosps;
    double sinth = sqrt(1.0 - costh * costh);
    double cosph = cos(phi), sinph = sin(phi);
    double u0 = costh * cosph, u1 = costh * sinph, u2 = -sinth;
    double v0 = sinph, v1 = -cosph, v2 = 0.;
    c.k[0] = -sinth * cosph;
    c.k[1] = -sinth * sinph;
    c.k[2] = -costh;
    c.ep[0] = v0 * v0 - u0 * u0;  c.ec[0] = u0 * v0 + v0 * u0;
    c.ep[1] = v1 * v1 - u1 * u1;  c.ec[1] = u1 * v1 + v1 * u1;
    c.ep[2] = v2 * v2 - u2 * u2;  c.ec[2] = u2 * v2 + v2 * u2;
    c.ep[3] = v0 * v1 - u0 * u1;  c.ec[3] = u0 * v1 + v0 * u1;
    c.ep[4] = v0 * v2 - u0 * u2;  c.ec[4] = u0 * v2 + v0 * u2;
    c.ep[5] = v1 * v2 - u1 * u2;  c.ec[5] = u1 * v2 + v1 * u2;
    c.fstar = fstar;
    c.invfstar = 1. / fstar;
    c.invL = 1. / L;
    c.sqT = sqrt(T);
    c.invT = 1. / T;
}

// r^T E r for a symmetric 3x3 tensor stored as xx,yy,zz,xy,xz,yz
__device__ inline double quad3(const double* E, double a, double b, double cc)
{
    return E[0] * a * a + E[1] * b * b + E[2] * cc * cc + 2.0 * (E[3] * a * b + E[4] * a * cc + E[5] * b * cc);
}

// One slow-series sample of the six links at time t: out(l, value) receives
// link l in the order 12, 21, 31, 13, 23, 32.
// pos: spacecraft positions x1,x2,x3,y1,y2,y3,z1,z2,z3 at t (metres), KS
// doubles apart (1 in a sample record, N in a structure-of-arrays table).
// glass_ucb_waveform.c:843-936, LISA_detector_tensor glass_lisa.c:1521-1564.
// The loop runs over the sending spacecraft i; its two links (i,j) share
// xi_i, arg2_i and the frequency at spacecraft i.  UNROLL unrolls it so that
// l is a compile-time index (values may then stay in registers).
template <bool UNROLL, int KS = 1, typename Out>
__device__ inline void ucb_slow_sample_to(const UcbConst& c, double t, const double* __restrict__ pos, Out&& out)
{
    const double x1 = pos[0], x2 = pos[KS], x3 = pos[2 * KS];
    const double y1 = pos[3 * KS], y2 = pos[4 * KS], y3 = pos[5 * KS];
    const double z1 = pos[6 * KS], z2 = pos[7 * KS], z3 = pos[8 * KS];

    // unit arm vectors (amplitude path)
    const double r12a = (x2 - x1) * c.invL, r12b = (y2 - y1) * c.invL, r12c = (z2 - z1) * c.invL;
    const double r13a = (x3 - x1) * c.invL, r13b = (y3 - y1) * c.invL, r13c = (z3 - z1) * c.invL;
    const double r23a = (x3 - x2) * c.invL, r23b = (y3 - y2) * c.invL, r23c = (z3 - z2) * c.invL;
    const double dp12 = quad3(c.ep, r12a, r12b, r12c), dc12 = quad3(c.ec, r12a, r12b, r12c);
    const double dp13 = quad3(c.ep, r13a, r13b, r13c), dc13 = quad3(c.ec, r13a, r13b, r13c);
    const double dp23 = quad3(c.ep, r23a, r23b, r23c), dc23 = quad3(c.ec, r23a, r23b, r23c);
    const double k12 = c.k[0] * r12a + c.k[1] * r12b + c.k[2] * r12c;
    const double k13 = c.k[0] * r13a + c.k[1] * r13b + c.k[2] * r13c;
    const double k23 = c.k[0] * r23a + c.k[1] * r23b + c.k[2] * r23c;
    double dft;
    {
#pragma clang fp contract(off)
        dft = c.df * t;
    }
#pragma unroll(UNROLL ? 3 : 1)
    for (int i = 0; i < 3; i++)
    {
        const double px = i == 0 ? x1 : (i == 1 ? x2 : x3);
        const double py = i == 0 ? y1 : (i == 1 ? y2 : y3);
        const double pz = i == 0 ? z1 : (i == 1 ? z2 : z3);
        double a2, fon, ae;
        {
#pragma clang fp contract(off)
            // the quotient by c, as IEEE division gives it (div_by_recip)
            double kdotx = div_by_recip(px * c.k[0] + py * c.k[1] + pz * c.k[2], kCLIGHT, kInvCLIGHT);
            double xi = t - kdotx;
            double fi = c.f0;
            a2 = c.pi2f0 * xi - c.phi0 - dft;
            ae = 1.0;
            if (c.np > 7)
            {
                fi += c.dfdt * xi;
                a2 += c.pidfdt * xi * xi;
                ae += c.aevc * xi;
            }
            if (c.np > 8)
            {
                fi += c.halfd2 * xi * xi;
                a2 += c.pid2 * xi * xi * xi;
            }
            fon = fi * c.invfstar;
        }
        // links (i, j) of spacecraft i: kdotr antisymmetric, d+/dx symmetric
        const double kra = i == 0 ? k12 : (i == 1 ? -k12 : -k13);
        const double krb = i == 0 ? k13 : (i == 1 ? k23 : -k23);
        const double dpa = i == 0 ? dp12 : (i == 1 ? dp12 : dp13);
        const double dpb = i == 0 ? dp13 : (i == 1 ? dp23 : dp23);
        const double dca = i == 0 ? dc12 : (i == 1 ? dc12 : dc13);
        const double dcb = i == 0 ? dc13 : (i == 1 ? dc23 : dc23);
#pragma unroll
        for (int h = 0; h < 2; h++)
        {
            const double kr = h ? krb : kra, dp = h ? dpb : dpa, dc = h ? dcb : dca;
            const double arg1 = 0.5 * fon * (1.0 + kr);
            const double sinc = 0.25 * sinc_ratio(arg1);
            const double t1r = ae * (dp * c.DPr + dc * c.DCr);
            const double t1i = ae * (dp * c.DPi + dc * c.DCi);
            double s2, c2;
            sincos_red(arg1 - a2, &s2, &c2);
            out(i + 3 * h, make_double2(sinc * (t1r * c2 + t1i * s2), sinc * (-t1r * s2 + t1i * c2)));
        }
        if constexpr (UNROLL) __builtin_amdgcn_sched_barrier(0);
    }
}

// The sample written to lds[l*stride + j].
template <int KS = 1>
__device__ inline void ucb_slow_sample(const UcbConst& c, double t, const double* __restrict__ pos,
                                       double2* lds, int stride, int j)
{
    ucb_slow_sample_to<false, KS>(c, t, pos, [&](int l, double2 v) { lds[l * stride + j] = v; });
}

// Time of slow sample n (0-based) of a BW-sample series:
// t = t0 + T*(double)(n)/(double)BW  (glass_ucb_waveform.c:843)
__host__ __device__ inline double ucb_sample_time(double T, double t0, int n, int BW)
{
#pragma clang fp contract(off)
    return t0 + T * (double)(n) / (double)BW;
}

// TDI format codes (glass_ucb_waveform.c:965-975)
enum TdiFormat : int { TDI_PHASE = 0, TDI_FREQUENCY = 1, TDI_SANGRIA = 2 };

// utils/src/glass_lisa.h:43-49
constexpr double kSPS = 8.321000e-23;
constexpr double kSACC = 9.000000e-30;
constexpr double kSLOC = 2.89e-24;

__host__ __device__ inline double ipow_(double x, int n)
{
    double xn = 1.0;
    for (int i = 0; i < n; i++) xn *= x;
    return xn;
}

// AEnoise with rednoise (utils/src/glass_lisa.c:1183-1198)
__host__ __device__ inline double ae_noise(double L, double fstar, double f)
{
    double fon = f / fstar;
    double trans = ipow_(sin(fon), 2);
    double red = 1. + 16.0 * (pow((2.0e-5 / f), 10.0) + ipow_(1.0e-4 / f, 2));
    return (16.0 / 3.0) * trans *
           ((2.0 + cos(fon)) * (kSPS + kSLOC) +
            2.0 * (3.0 + 2.0 * cos(fon) + cos(2.0 * fon)) * (kSLOC / 2.0 + kSACC / ipow_(kPI2 * f, 4) * red)) /
           ipow_(2.0 * L, 2);
}

// 2^((int)log2(x) + 1) clamped to [nmin, nmax], as ucb_bandwidth forms its
// Doppler and SNR spreads (ucb/src/glass_ucb_waveform.c:709-722).  Values
// below 8 (including negative or NaN x) land on the floor nmin = 16.
__host__ __device__ inline int pow2_spread(double x, int nmin, int nmax)
{
    if (!(x >= 8.0)) return nmin;
    int e = (int)log2(x) + 1;
    int v = e >= 30 ? (1 << 30) : (1 << e);
    if (v > nmax) v = nmax;
    if (v < nmin) v = nmin;
    return v;
}

// ucb_bandwidth (ucb/src/glass_ucb_waveform.c:698-725)
__host__ __device__ inline int ucb_bandwidth(double L, double fstar, double f, double fdot, double costheta,
                                             double A, double T, int N)
{
#pragma clang fp contract(off)
    const int Nmin = 16;
    int half = N / 2, lg = 0;
    while ((2 << lg) <= half) lg++;  // (int)log2(N/2)
    const int Nmax = 1 << lg;
    double sqT = sqrt(T);
    double sf = sin(f / fstar);
    double sn = ae_noise(L, fstar, f);
    double sintheta = sin(acos(costheta));
    double bw = 2 * T * ((4. + kPI2 * f * (kAU / kCLIGHT) * sintheta) / kYEAR + fabs(fdot) * T);
    int DS = pow2_spread(bw - 1, Nmin, Nmax);
    double SNRm = A * sqT * sf / sqrt(sn);  // analytic_snr, glass_ucb_waveform.c:22-25
    int SS = pow2_spread(SNRm - 1, Nmin, Nmax);
    return (DS > SS) ? DS : SS;
}

}  // namespace glass
