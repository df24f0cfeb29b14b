// wdm_common.h -- shared pieces of the wavelet-domain (WDM) path: natural
// cubic splines with the reference's coefficient recurrence and lookups,
// the TDI term table, and the per-template constants.
//
// Everything on the phase chain is evaluated with contraction off and in the
// reference's operation order (utils/src/glass_math.c:98-190 splines,
// ucb/src/glass_ucb_waveform.c:331-362 barycentre phase), so a template's
// phase samples of ~1e6 rad round exactly as on the CPU.
#pragma once
#include <hip/hip_runtime.h>

#include "fastmath.h"

namespace glass {
namespace wdm {

constexpr double PI2 = 6.283185307179586;
constexpr double CLIGHT = 299792458.;
constexpr double VEARTH = 0.00010103671;
constexpr int kMaxLayers = 3;   // ucb_wavelet_layers spans at most j-1..j+1
// A layer count forced on every template of a batched plan (a workload
// knob with no reference counterpart: the reference always takes
// ucb_wavelet_layers' 1-3): L + 1 a power of two from 8 to 64 with
// NT (L + 1) <= 32768, at NT = 256 or 512.  Its tables take the slot after
// the reference counts'.
constexpr int kForcedSlot = kMaxLayers + 1;
constexpr int kMaxForcedSamples = 32768;
GH_HD inline bool forced_layers_ok(int NT, int L)
{
    const int n = L + 1;
    return (NT == 256 || NT == 512) && n >= 8 && n <= 64 && (n & (n - 1)) == 0 && NT * n <= kMaxForcedSamples;
}
constexpr int kMaxKnots = 512;  // orbit knots per template (200 per year: 2.5 yr)

// Index lookups of spline_interpolation (utils/src/glass_math.c:393-467).
// even: floor((x - a0)/(a1 - a0)) clamped to [0, N-1].
GH_HD inline int even_index(double x, double a0, double ds, int N)
{
#pragma clang fp contract(off)
    int n = (int)floor((x - a0) / ds);
    if (n < 0) n = 0;
    if (n > N - 1) n = N - 1;
    return n;
}

// binary_search on strictly increasing knots: the largest index with
// a[i] <= x for a[0] <= x <= a[N-1] (x equal to the last knot returns N-1,
// as the reference's recursion does).  Callers keep x inside the knots.
GH_HD inline int knot_index(const double* a, int N, double x)
{
    if (x >= a[N - 1]) return N - 1;
    int lo = 0, hi = N - 1;
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// y0 + y1 dx + y2 dx dx + y3 dx dx dx, coefficients c[k*N + n]
GH_HD inline double spline_at(const double* c, int N, int n, double dx)
{
#pragma clang fp contract(off)
    return c[n] + c[N + n] * dx + c[2 * N + n] * dx * dx + c[3 * N + n] * dx * dx * dx;
}

// Barycentre phase and amplitude (ucb_phase / ucb_amplitude); fddot = 0.
GH_HD inline double ucb_phase(double t, double f0, double phi0, double fdot)
{
#pragma clang fp contract(off)
    const double fddot = 0.0;
    return -phi0 + PI2 * (f0 * t + 0.5 * fdot * t * t + 1.0 / 6.0 * fddot * t * t * t);
}
GH_HD inline double ucb_amplitude(double t, double A0, double f0, double fdot)
{
#pragma clang fp contract(off)
    return A0 * (1.0 + 2.0 / 3.0 * fdot / f0 * t);
}

// The eight delayed one-arm terms of a Michelson TDI channel with vertices
// (a,b,c) (LISA_TDI_spline, utils/src/glass_lisa.c:769-887):
//   t_e = t - kr[v] - m1*L[i1] - m2*L[i2]
//   M  += s1 (h+ P1[j1] + hx Q1[j1]) ;  M += s2 (h+ P2[j2] + hx Q2[j2])
// (P,Q) = (App,Acp) for pm = 1, (Apm,Acm) for pm = 0; indices 0=a,1=b,2=c.
// Multiplying an arm length by 1 or 0 is exact, so each delay rounds as the
// written-out expression.
struct TdiTerm
{
    signed char v, i1, i2, s1, pm1, j1, s2, pm2, j2;
    double m1, m2;
};

// ucb_wavelet_layers (ucb/src/glass_ucb_waveform.c:1003-1037)
GH_HD inline void wavelet_layers(double p0, double p7, double Tobs, double df, double A, int* jstart,
                                 int* jwidth)
{
#pragma clang fp contract(off)
    const double fstart = p0 / Tobs;
    const double fstop = (p0 / Tobs + p7 / Tobs / Tobs);
    double fmin, fmax;
    if (fstart < fstop)
    {
        fmin = fstart * (1.0 - VEARTH);
        fmax = fstop * (1.0 + VEARTH);
    }
    else
    {
        fmin = fstop * (1.0 - VEARTH);
        fmax = fstart * (1.0 + VEARTH);
    }
    const int j = (int)rint(fmin / df);
    const double dfmin = fmin - j * df;
    const double dfmax = fmax - j * df;
    int jmin = j, jmax = j;
    if (dfmin < 0.0 && fabs(dfmin) > A / PI2) jmin = j - 1;
    if (dfmax > 0.0 && fabs(dfmax) > A / PI2) jmax = j + 1;
    *jstart = jmin;
    *jwidth = jmax - jmin + 1;
}

// Per-template constants shared by the two template kernels.
struct Hdr
{
    double f0, phi0, fdot, A0;     // barycentre phase / amplitude
    double k[3];                   // propagation vector
    double Ap, Ac, c2p, s2p;       // polarisation amplitudes, cos/sin 2 psi
    double ep[9], ec[9];           // polarisation tensors (3x3 row-major)
    double fhet;                   // heterodyne frequency (jmin-1) df
    int jmin, L, pad0, pad1;
};

// Fill the constants from params[8] (map_array_to_ucb_params ordering):
// [f0 T, cos theta, phi, ln A, cos iota, psi, phi0, fdot T^2].
GH_HD inline void make_hdr(const double* p, double Tobs, double df, double A, Hdr* h)
{
#pragma clang fp contract(off)
    h->f0 = p[0] / Tobs;
    h->phi0 = p[6];
    h->fdot = p[7] / Tobs / Tobs;
    h->A0 = exp_cr(p[3]);
    const double costh = p[1], phi = p[2];
    const double sinth = sqrt(1.0 - costh * costh);
    double sinph, cosph;
    sincos_cr(phi, &sinph, &cosph);
    h->k[0] = -sinth * cosph;
    h->k[1] = -sinth * sinph;
    h->k[2] = -costh;
    // LISA_polarization_tensor (utils/src/glass_lisa.c:1566-1596)
    const double u[3] = {costh * cosph, costh * sinph, -sinth};
    const double v[3] = {sinph, -cosph, 0.};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
        {
            h->ep[3 * i + j] = v[i] * v[j] - u[i] * u[j];
            h->ec[3 * i + j] = u[i] * v[j] + v[i] * u[j];
        }
    const double cosi = p[4], psi = p[5];
    sincos_cr(2. * psi, &h->s2p, &h->c2p);
    h->Ap = 0.5 * (1. + cosi * cosi);
    h->Ac = -cosi;
    wavelet_layers(p[0], p[7], Tobs, df, A, &h->jmin, &h->L);
    h->fhet = (h->jmin - 1) * df;
    h->pad0 = h->pad1 = 0;
}

}  // namespace wdm
}  // namespace glass
