// fd_tangent.h -- exact parameter derivatives of the frequency-domain UCB
// template (forward mode), built on the same slow-series construction as
// the waveform (ucb/src/glass_ucb_waveform.c:804-962).
//
// One link sample is V = S * T1 * E with
//   S  = sin(arg1)/(4 arg1),  arg1 = f_i/(2 f*) (1 + k.r_ij)
//   T1 = aevol (d+ DP + dx DC),  DP = DPr + i DPi, DC = DCr + i DCi
//   E  = exp(-i Phi),  Phi = arg1 - arg2
// and FFT + TDI are complex-linear with parameter-independent coefficients
// (q and BW held fixed), so d h / d theta = TDI(FFT(d V / d theta)).
// Because V is complex-linear in (DP, DC):
//   h = DP Hp + DC Hc  with Hp = TDI FFT (S aevol d+ E), Hc likewise with dx,
//   d h/d lnA  = h,   d h/d phi0 = -i h,
//   d h/d cos(iota), d h/d psi = DP' Hp + DC' Hc  (DP', DC' closed form),
// leaving genuine tangent series only for f0 T, cos(theta), phi and
// fdot T^2.  The six "series sets" built per sample are
//   0: S aevol d+ E    1: S aevol dx E
//   2: dV/d(f0 T)      3: dV/d cos(theta)   4: dV/d phi   5: dV/d(fdot T^2)
#pragma once
#include "fd_waveform.h"

namespace glass {

// Sky and polarisation pieces needed by the derivatives.
struct UcbDiff
{
    double u[3], v[3];
    double du1[3], du2[3], dv2[3];  // d/dcos(theta), d/dphi (dv/dcos(theta) = 0)
    double dk1[3], dk2[3];
    double DPr4, DPi4, DCr4, DCi4;  // d(DP, DC)/d cos(iota)
    double DPr5, DPi5, DCr5, DCi5;  // d(DP, DC)/d psi
    double invT, invT2, f0;
};

__host__ __device__ inline void ucb_diff_setup(const double* p, double T, UcbDiff& d)
{
    const double costh = p[1], phi = p[2], amp = exp(p[3]), cosi = p[4], psi = p[5];
    const double sinth = sqrt(1.0 - costh * costh);
    const double cosph = cos(phi), sinph = sin(phi);
    // d sin(theta)/d cos(theta) = -cot diverges at the poles (|cos theta| =
    // 1, where clipped proposals land); there the tangent is taken as at
    // |cos theta| = 1 - 1e-6, the reference's Fisher step
    // (ucb/src/glass_ucb_waveform.c:369-415), so the Fisher matrix stays
    // finite (sin theta >= sqrt(2e-6) in the cotangent only)
    const double cot = costh / fmax(sinth, 1.4142135623730951e-3);
    d.u[0] = costh * cosph;  d.u[1] = costh * sinph;  d.u[2] = -sinth;
    d.v[0] = sinph;          d.v[1] = -cosph;         d.v[2] = 0.0;
    d.du1[0] = cosph;        d.du1[1] = sinph;        d.du1[2] = cot;
    d.du2[0] = -costh * sinph; d.du2[1] = costh * cosph; d.du2[2] = 0.0;
    d.dv2[0] = cosph;        d.dv2[1] = sinph;        d.dv2[2] = 0.0;
    d.dk1[0] = cot * cosph;  d.dk1[1] = cot * sinph;  d.dk1[2] = -1.0;
    d.dk2[0] = sinth * sinph; d.dk2[1] = -sinth * cosph; d.dk2[2] = 0.0;
    const double cosps = cos(2. * psi), sinps = sin(2. * psi);
    const double Aplus = amp * (1. + cosi * cosi) * 0.5, Across = -amp * cosi;
    // dAplus/dcos(iota) = amp cos(iota), dAcross/dcos(iota) = -amp
    d.DPr4 = amp * cosi * cosps;  d.DPi4 = amp * sinps;
    d.DCr4 = -amp * cosi * sinps; d.DCi4 = amp * cosps;
    // d cos(2 psi)/d psi = -2 sin(2 psi), d sin(2 psi)/d psi = 2 cos(2 psi)
    d.DPr5 = -2.0 * Aplus * sinps;  d.DPi5 = -2.0 * Across * cosps;
    d.DCr5 = -2.0 * Aplus * cosps;  d.DCi5 = 2.0 * Across * sinps;
    d.invT = 1.0 / T;
    d.invT2 = 1.0 / (T * T);
    d.f0 = p[0] / T;
}

// sin(x)/x and its derivative (cos x - sin(x)/x)/x, by series for |x|<=1.75
__device__ inline void sinc_ratio_d(double x, double* f, double* df)
{
    const double x2 = x * x;
    if (!(x2 <= 3.0625))
    {
        double s, c;
        sincos_red(x, &s, &c);
        *f = s / x;
        *df = (c - *f) / x;
        return;
    }
    // f = sum c_n x^{2n}, c_n = (-1)^n/(2n+1)!; f' = x sum 2n c_n x^{2n-2}
    const double c[13] = {1.0, -1.0 / 6.0, 1.0 / 120.0, -1.0 / 5040.0, 1.0 / 362880.0, -1.0 / 39916800.0,
                          1.0 / 6227020800.0, -1.0 / 1307674368000.0, 1.0 / 355687428096000.0,
                          -1.0 / 121645100408832000.0, 1.0 / 51090942171709440000.0,
                          -1.0 / 25852016738884976640000.0, 1.0 / 15511210043330985984000000.0};
    double p = c[12], q = 24.0 * c[12];
#pragma unroll
    for (int n = 11; n >= 1; n--)
    {
        p = fma(p, x2, c[n]);
        q = fma(q, x2, 2.0 * n * c[n]);
    }
    *f = fma(p, x2, c[0]);
    *df = x * q;
}

__device__ inline double2 cmulc(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// Series sets [S0, S0+NS) of one slow sample at time t: out(set, link,
// value) for every set of the range and link 0..5 (12, 21, 31, 13, 23, 32).
// Sets 0, 1: the plus / cross basis (the sample is DP s0 + DC s1, complex
// products); 2..5: d/d(f0 T), d/dcos(theta), d/dphi, d/d(fdot T^2).
template <int S0, int NS, typename Out>
__device__ __forceinline__ void ucb_sample_sets_to(const UcbConst& c, const UcbDiff& dd, double t,
                                          const double* __restrict__ pos, Out&& out)
{
    constexpr bool W01 = S0 <= 1 && S0 + NS > 0;  // any basis set
    constexpr bool WT = S0 + NS > 2;              // any tangent set
    const double x[3] = {pos[0], pos[1], pos[2]}, y[3] = {pos[3], pos[4], pos[5]}, z[3] = {pos[6], pos[7], pos[8]};
    // arms 12, 13, 23: projections on u, v, k and their sky derivatives
    double ra[3], rb[3], rk[3], ra1[3], ra2[3], rb2[3], rk1[3], rk2[3];
#pragma unroll
    for (int m = 0; m < 3; m++)
    {
        const int i0 = m < 2 ? 0 : 1, i1 = m == 0 ? 1 : 2;
        const double r0 = (x[i1] - x[i0]) * c.invL, r1 = (y[i1] - y[i0]) * c.invL, r2 = (z[i1] - z[i0]) * c.invL;
        ra[m] = r0 * dd.u[0] + r1 * dd.u[1] + r2 * dd.u[2];
        rb[m] = r0 * dd.v[0] + r1 * dd.v[1] + r2 * dd.v[2];
        rk[m] = r0 * c.k[0] + r1 * c.k[1] + r2 * c.k[2];
        ra1[m] = r0 * dd.du1[0] + r1 * dd.du1[1] + r2 * dd.du1[2];
        ra2[m] = r0 * dd.du2[0] + r1 * dd.du2[1] + r2 * dd.du2[2];
        rb2[m] = r0 * dd.dv2[0] + r1 * dd.dv2[1] + r2 * dd.dv2[2];
        rk1[m] = r0 * dd.dk1[0] + r1 * dd.dk1[1] + r2 * dd.dk1[2];
        rk2[m] = r0 * dd.dk2[0] + r1 * dd.dk2[1] + r2 * dd.dk2[2];
    }
    double dft;
    {
#pragma clang fp contract(off)
        dft = c.df * t;
    }
    const double2 DP = make_double2(c.DPr, c.DPi), DC = make_double2(c.DCr, c.DCi);
    const double ca = 0.66666666666666666666;
#pragma unroll 1
    for (int i = 0; i < 3; i++)
    {
        const double px = i == 0 ? x[0] : (i == 1 ? x[1] : x[2]);
        const double py = i == 0 ? y[0] : (i == 1 ? y[1] : y[2]);
        const double pz = i == 0 ? z[0] : (i == 1 ? z[1] : z[2]);
        double xi, a2, fi, ae;
        {
#pragma clang fp contract(off)
            const double kdotx = div_by_recip(px * c.k[0] + py * c.k[1] + pz * c.k[2], kCLIGHT, kInvCLIGHT);
            xi = t - kdotx;
            fi = c.f0;
            a2 = c.pi2f0 * xi - c.phi0 - dft;
            ae = 1.0;
            if (c.np > 7)
            {
                fi += c.dfdt * xi;
                a2 += c.pidfdt * xi * xi;
                ae += c.aevc * xi;
            }
        }
        const double fon = fi * c.invfstar;
        // derivatives of xi, arg2, f_i, aevol: index 0 -> f0 T, 1 -> cos(theta),
        // 2 -> phi, 3 -> fdot T^2
        const double dxi1 = -div_by_recip(px * dd.dk1[0] + py * dd.dk1[1] + pz * dd.dk1[2], kCLIGHT, kInvCLIGHT);
        const double dxi2 = -div_by_recip(px * dd.dk2[0] + py * dd.dk2[1] + pz * dd.dk2[2], kCLIGHT, kInvCLIGHT);
        const double da2dxi = kPI2 * (dd.f0 + c.dfdt * xi);
        const double da2[4] = {kPI2 * xi * dd.invT, da2dxi * dxi1, da2dxi * dxi2, kPI * xi * xi * dd.invT2};
        const double dfo[4] = {dd.invT * c.invfstar, c.dfdt * dxi1 * c.invfstar, c.dfdt * dxi2 * c.invfstar,
                               xi * dd.invT2 * c.invfstar};
        const double dae[4] = {-ca * c.dfdt * xi * dd.invT / (dd.f0 * dd.f0), ca * c.dfdt / dd.f0 * dxi1,
                               ca * c.dfdt / dd.f0 * dxi2, ca * xi * dd.invT2 / dd.f0};
#pragma unroll
        for (int h = 0; h < 2; h++)
        {
            // link (i, j): arm and sign of k.r_ij (antisymmetric)
            const int arm = (i == 0) ? (h ? 1 : 0) : (i == 1 ? (h ? 2 : 0) : (h ? 2 : 1));
            const double sg = (i == 0) ? 1.0 : (i == 1 ? (h ? 1.0 : -1.0) : -1.0);
            const int link = i + 3 * h;  // 12,21,31 then 13,23,32
            const double a = arm == 0 ? ra[0] : (arm == 1 ? ra[1] : ra[2]);
            const double b = arm == 0 ? rb[0] : (arm == 1 ? rb[1] : rb[2]);
            const double kr = sg * (arm == 0 ? rk[0] : (arm == 1 ? rk[1] : rk[2]));
            const double dp = b * b - a * a, dc = 2.0 * a * b;
            const double arg1 = 0.5 * fon * (1.0 + kr);
            double sr, dsr;
            sinc_ratio_d(arg1, &sr, &dsr);
            const double S = 0.25 * sr;
            double sPh, cPh;
            sincos_red(arg1 - a2, &sPh, &cPh);
            const double2 E = make_double2(cPh, -sPh);
            if constexpr (W01)
            {
#pragma unroll
                for (int q = 0; q < 2; q++)
                {
                    if (q < S0 || q >= S0 + NS) continue;
                    const double w = S * ae * (q == 0 ? dp : dc);
                    out(q, link, make_double2(w * E.x, w * E.y));
                }
            }
            if constexpr (WT)
            {
                const double2 T1 = make_double2(ae * (dp * DP.x + dc * DC.x), ae * (dp * DP.y + dc * DC.y));
                const double2 ST1E = cmulc(make_double2(S * T1.x, S * T1.y), E);
                const double a1 = arm == 0 ? ra1[0] : (arm == 1 ? ra1[1] : ra1[2]);
                const double a2v = arm == 0 ? ra2[0] : (arm == 1 ? ra2[1] : ra2[2]);
                const double b2 = arm == 0 ? rb2[0] : (arm == 1 ? rb2[1] : rb2[2]);
                const double k1 = sg * (arm == 0 ? rk1[0] : (arm == 1 ? rk1[1] : rk1[2]));
                const double k2 = sg * (arm == 0 ? rk2[0] : (arm == 1 ? rk2[1] : rk2[2]));
#pragma unroll
                for (int q = 2; q < 6; q++)
                {
                    if (q < S0 || q >= S0 + NS) continue;
                    const int v = q - 2;  // 0: f0 T, 1: cos(theta), 2: phi, 3: fdot T^2
                    const double dkr = v == 1 ? k1 : (v == 2 ? k2 : 0.0);
                    const double ddp = v == 1 ? -2.0 * a * a1 : (v == 2 ? 2.0 * b * b2 - 2.0 * a * a2v : 0.0);
                    const double ddc = v == 1 ? 2.0 * a1 * b : (v == 2 ? 2.0 * (a2v * b + a * b2) : 0.0);
                    const double darg1 = 0.5 * (dfo[v] * (1.0 + kr) + fon * dkr);
                    const double dS = 0.25 * dsr * darg1;
                    const double dPhi = darg1 - da2[v];
                    const double2 dT1 = make_double2(dae[v] * (dp * DP.x + dc * DC.x) + ae * (ddp * DP.x + ddc * DC.x),
                                                     dae[v] * (dp * DP.y + dc * DC.y) + ae * (ddp * DP.y + ddc * DC.y));
                    // E (dS T1 + S dT1) - i dPhi S T1 E
                    const double2 g = cmulc(make_double2(dS * T1.x + S * dT1.x, dS * T1.y + S * dT1.y), E);
                    out(q, link, make_double2(g.x + dPhi * ST1E.y, g.y - dPhi * ST1E.x));
                }
            }
        }
    }
}

// The sets [S0, S0+NS) of one sample into lds[(set - S0)*6*stride + link*stride + j].
template <int S0, int NS>
__device__ __forceinline__ void ucb_sample_sets(const UcbConst& c, const UcbDiff& dd, double t, const double* __restrict__ pos,
                                       double2* lds, int stride, int j)
{
    ucb_sample_sets_to<S0, NS>(c, dd, t, pos,
                               [&](int q, int link, double2 v) { lds[(q - S0) * 6 * stride + link * stride + j] = v; });
}

}  // namespace glass
