/*
 * orc_wdm.c -- oracle for the wavelet-domain (WDM) UCB path (test
 * infrastructure only: the parity checker and CPU baseline of the HIP
 * library).
 *
 * Restates, in FP64 with the reference's operation order:
 *   hypergeometric_function, beta_function, incomplete_beta_function
 *                                 utils/src/glass_math.c:1214-1242
 *   setup_wdm_basis (A = 0)       utils/src/glass_wavelet.c:39-49
 *   my_phitilde, build_wdm_filter_freq utils/src/glass_wavelet.c:305-349
 *   wavelet_transform_timefreq    utils/src/glass_wavelet.c:293-303
 *   wavelet_transform_inverse_freq utils/src/glass_wavelet.c:351-405
 *   wavelet_transform_freq        utils/src/glass_wavelet.c:410-460
 *   wavelet_transform_timefreq_by_layers utils/src/glass_wavelet.c:544-611
 *   spline_coefficients, spline_interpolation, even_sampled_search,
 *   binary_search                 utils/src/glass_math.c:98-190,393-473
 *   initialize_interpolated_analytic_orbits (knots + position splines)
 *                                 utils/src/glass_lisa.c:271-380
 *   LISA_spacecraft_to_barycenter_time utils/src/glass_lisa.c:712-754
 *   hplus_and_hcross, LISA_TDI_spline, LISA_spline_response
 *                                 utils/src/glass_lisa.c:757-1074
 *   extract_amplitude_and_phase   utils/src/glass_math.c:1244-1308
 *   unwrap_phase                  utils/src/glass_math.c:1034-1052
 *   ucb_phase, ucb_amplitude, ucb_barycenter_waveform
 *                                 ucb/src/glass_ucb_waveform.c:331-362
 *   ucb_wavelet_layers, build_interpolated_waveform, ucb_waveform_wavelet
 *                                 ucb/src/glass_ucb_waveform.c:1003-1240
 *   snr_wavelet                   utils/src/glass_math.c:428-438
 *   gaussian_log_likelhood_wavelet utils/src/glass_model.c:805-855
 *   stationary_dft_psd_to_wdm_layer_var / _psd utils/src/glass_wavelet.c:865-921
 * Constants: utils/src/glass_constants.h (PI2, CLIGHT, YEAR, AU, VEARTH),
 * utils/src/glass_lisa.h:37 (200 orbit samples per year),
 * utils/src/glass_data.h:29 (edge buffer), utils/src/glass_wavelet.h:21
 * (filter constant 6).
 */
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#include "glass_oracle.h"

#define W_PI2    6.283185307179586
#define W_CLIGHT 299792458.
#define W_YEAR   3.15581497632e7
#define W_VEARTH 0.00010103671
#define W_SAMPLES_PER_YEAR 200
#define W_EDGE   100
#define W_FILTER 6.0

/* Elementary functions of the template path.  Default: the platform libm
 * (glibc), as the reference calls it.  With orc_set_cr_trig(1) sin and cos
 * (and atan2, exp) are the correctly rounded values (quad-precision sinq,
 * cosq, atan2q, expq rounded to double): glibc misrounds ~0.1% of these arguments, and the TDI
 * combination of eight delayed terms cancels by (f/f*)^2, so the reference's
 * own output moves by up to ~1e-11 between conforming libms; the switch
 * separates that libm dependence from the algorithm in parity checks. */
static int g_cr_trig = 0;
void orc_set_cr_trig(int on) { g_cr_trig = on; }
static double w_sin(double x) { return g_cr_trig ? (double)sinq((__float128)x) : sin(x); }
static double w_cos(double x) { return g_cr_trig ? (double)cosq((__float128)x) : cos(x); }
static double w_exp(double x) { return g_cr_trig ? (double)expq((__float128)x) : exp(x); }
static double w_atan2(double y, double x)
{
    return g_cr_trig ? (double)atan2q((__float128)y, (__float128)x) : atan2(y, x);
}

/* LISA_polarization_tensor (utils/src/glass_lisa.c:1566-1596) with the
 * template path's sin/cos */
static void pol_tensor(double costh, double phi, double ep[4][4], double ec[4][4], double k[4])
{
    double u[4], v[4];
    double sinth = sqrt(1.0 - costh * costh);
    double cosph = w_cos(phi), sinph = w_sin(phi);
    u[1] = costh * cosph;  u[2] = costh * sinph;  u[3] = -sinth;
    v[1] = sinph;          v[2] = -cosph;         v[3] = 0.;
    k[1] = -sinth * cosph; k[2] = -sinth * sinph; k[3] = -costh;
    for (int i = 1; i <= 3; i++)
        for (int j = 1; j <= 3; j++)
        {
            ep[i][j] = v[i] * v[j] - u[i] * u[j];
            ec[i][j] = u[i] * v[j] + v[i] * u[j];
        }
}

/* ------------------------------------------------------------ window ---- */

static double hypergeometric(double a, double b, double c, double x)
{
    /* Gauss series 2F1(a,b;c;x), terms until |term| <= 1e-8 */
    double term = a * b * x / c;
    double value = 1.0 + term;
    int n = 1;
    while (fabs(term) > 1e-8)
    {
        a++, b++, c++, n++;
        term *= a * b * x / c / n;
        value += term;
    }
    return value;
}

double orc_incomplete_beta(double a, double b, double x)
{
    double F = hypergeometric(a, 1 - b, a + 1, x);
    double B = exp(lgamma(a) + lgamma(b) - lgamma(a + b));
    return pow(x, a) * F / B / a;
}

static double meyer(double omega, double DeltaOmega, double A)
{
    double B = DeltaOmega - 2 * A;
    double insDOM = 1.0 / sqrt(DeltaOmega);
    double z = 0.0;
    if (fabs(omega) >= A && fabs(omega) < A + B)
    {
        double x = (fabs(omega) - A) / B;
        double y = orc_incomplete_beta(W_FILTER, W_FILTER, x);
        z = insDOM * cos(y * M_PI_2);
    }
    if (fabs(omega) < A) z = insDOM;
    return z;
}

void orc_wdm_filter(double *phif, int Nf, int Nt, double A, int forward)
{
    double DeltaOmega = M_PI / Nf;
    double domega = 2 * DeltaOmega / (double)Nt;
    for (int i = 0; i < Nt / 2 + 1; i++) phif[i] = meyer(domega * i, DeltaOmega, A);
    double norm = 0.0;
    for (int i = 0; i < Nt / 2 + 1; i++)
        norm += (i == 0 ? 1 : 2) * phif[i] * phif[i];
    norm *= domega * M_1_PI;
    norm = sqrt(norm);
    if (forward) norm *= Nf / 2.0;
    for (int i = 0; i < Nt / 2 + 1; i++) phif[i] /= norm;
}

void orc_wdm_init(orc_wdm *w, int NF, int NT, double dt, double *window)
{
    w->NF = NF;
    w->NT = NT;
    w->dt = dt;
    w->df = 1.0 / (2.0 * dt);
    /* setup_wdm_basis: B = dOmega, A = (dOmega - B)/2 */
    w->A = 0.0;
    w->imin = W_EDGE;
    w->imax = NT - W_EDGE;
    w->kmin = NT;              /* first pixel of layer 1        */
    w->kmax = NT * (NF - 1);   /* first pixel of layer NF - 1   */
    w->window = window;
    orc_wdm_filter(window, NF, NT, w->A, 1);
}

/* --------------------------------------------------------- transforms --- */

static void ifft_norm(const double *in, double *out, int n)
{
    orc_fft(in, out, n, 1);
    for (int i = 0; i < 2 * n; i++) out[i] /= n;
}

/* one layer m of the forward transform from the one-sided spectrum */
static void layer_forward(const orc_wdm *w, const double *freq, int m, double *DX, double *DXt,
                          double *wdm)
{
    int Nf = w->NF, Nt = w->NT;
    int center = m * Nt / 2;
    memset(DX, 0, 2 * Nt * sizeof(double));
    for (int j = -Nt / 2 + 1; j < Nt / 2; j++)
    {
        int idx = center + j;
        if (m == 0 && idx < 0) continue;
        if (m == Nf && idx > center) continue;
        double wt = w->window[abs(j)];
        if (j == 0 && (m == 0 || m == Nf)) wt /= 2.0;
        DX[2 * (Nt / 2 + j)] = wt * freq[2 * idx];
        DX[2 * (Nt / 2 + j) + 1] = wt * freq[2 * idx + 1];
    }
    ifft_norm(DX, DXt, Nt);
    if (m == 0)
        for (int n = 0; n < Nt; n += 2) wdm[n] = DXt[2 * n] * M_SQRT2;
    else if (m == Nf)
        for (int n = 0; n < Nt; n += 2) wdm[n + 1] = DXt[2 * n] * M_SQRT2;
    else
    {
        int sgn = (m % 2 == 0) ? 1 : -1;
        for (int n = 0; n < Nt; n++)
            wdm[n + Nt * m] = ((n + m) % 2 == 0) ? DXt[2 * n] : sgn * DXt[2 * n + 1];
    }
}

void orc_wavelet_transform_freq(const orc_wdm *w, const double *freq, double *wdm)
{
    int Nt = w->NT;
    double *DX = malloc(4 * Nt * sizeof(double));
    for (int m = 0; m <= w->NF; m++) layer_forward(w, freq, m, DX, DX + 2 * Nt, wdm);
    free(DX);
}

void orc_wavelet_transform_timefreq(const orc_wdm *w, double *timedata)
{
    int ND = w->NT * w->NF;
    double *freq = malloc((ND + 2) * sizeof(double));
    orc_fftr(timedata, freq, ND);
    orc_wavelet_transform_freq(w, freq, timedata);
    free(freq);
}

void orc_wavelet_transform_inverse_freq(const orc_wdm *w, const double *wdm, double *freq)
{
    int Nf = w->NF, Nt = w->NT, ND = Nf * Nt;
    memset(freq, 0, (ND + 2) * sizeof(double));
    double *pre = malloc(4 * Nt * sizeof(double));
    double *res = pre + 2 * Nt;
    for (int m = 0; m < Nf + 1; m++)
    {
        int center = m * Nt / 2;
        memset(pre, 0, 2 * Nt * sizeof(double));
        if (m == 0)
            for (int n = 0; n < Nt; n++) pre[2 * n] = wdm[(2 * n) % Nt] * M_SQRT1_2;
        else if (m == Nf)
            for (int n = 0; n < Nt; n++) pre[2 * n] = wdm[(2 * n) % Nt + 1] * M_SQRT1_2;
        else
            for (int n = 0; n < Nt; n++)
            {
                if ((n + m) % 2 == 0) pre[2 * n] = wdm[n + Nt * m];
                else pre[2 * n + 1] = -wdm[n + Nt * m];
            }
        orc_fft(pre, res, Nt, 0);
        int lo = center - Nt / 2, hi = center + Nt / 2;
        if (lo < 0) lo = 0;
        if (hi > ND / 2 + 1) hi = ND / 2 + 1;
        for (int i = lo; i < hi; i++)
        {
            int d = abs(i - center);
            if (d > Nt / 2) continue;
            int src = (m == 0 || m == Nf) ? (2 * i) % Nt : i % Nt;
            freq[2 * i] += res[2 * src] * w->window[d] * Nf / 2;
            freq[2 * i + 1] += res[2 * src + 1] * w->window[d] * Nf / 2;
        }
    }
    free(pre);
}

void orc_wavelet_transform_timefreq_by_layers(const orc_wdm *w, int jmin, int Nlayers,
                                              const double *phif, double *timedata)
{
    int Nf = w->NF, Nt = w->NT;
    int N = Nt * (Nlayers + 1);
    double normalization = sqrt((double)Nf) / (Nlayers + 1.0);
    double *freq = malloc((N + 2) * sizeof(double));
    double *DX = malloc(4 * Nt * sizeof(double));
    double *DXt = DX + 2 * Nt;
    orc_fftr(timedata, freq, N);
    memset(timedata, 0, N * sizeof(double));
    for (int m = jmin; m < jmin + Nlayers; m++)
    {
        int off = m - jmin;
        int center = (off + 1) * Nt / 2;
        memset(DX, 0, 2 * Nt * sizeof(double));
        for (int j = -Nt / 2 + 1; j < Nt / 2; j++)
        {
            int idx = center + j;
            if (idx < 0 || idx > N / 2) continue;
            double wt = phif[abs(j)];
            if (j == 0 && (m == 0 || m == Nf)) wt /= 2.0;
            DX[2 * (Nt / 2 + j)] = wt * freq[2 * idx];
            DX[2 * (Nt / 2 + j) + 1] = wt * freq[2 * idx + 1];
        }
        ifft_norm(DX, DXt, Nt);
        if (m == 0 || m == Nf)
            for (int n = 0; n < Nt; n += 2)
                timedata[n * Nlayers + off] = DXt[2 * n] * M_SQRT2 * normalization;
        else
        {
            int sgn = (m % 2 == 0) ? 1 : -1;
            for (int n = 0; n < Nt; n++)
                timedata[n * Nlayers + off] = ((n + m) % 2 == 0) ? DXt[2 * n] * normalization
                                                                 : sgn * DXt[2 * n + 1] * normalization;
        }
    }
    free(DX);
    free(freq);
}

/* ------------------------------------------------------------ splines --- */

void orc_spline_coefficients(int N, const double *x, const double *y, double *c)
{
    double *y0 = c, *y1 = c + N, *y2 = c + 2 * N, *y3 = c + 3 * N;
    double *dx = calloc(N, sizeof(double)), *alpha = calloc(N, sizeof(double));
    double *l = calloc(N, sizeof(double)), *z = calloc(N, sizeof(double));
    double *mu = calloc(N, sizeof(double));
    dx[0] = x[1] - x[0];
    for (int i = 1; i < N - 1; i++)
    {
        dx[i] = x[i + 1] - x[i];
        alpha[i] = 3.0 * ((y[i + 1] - y[i]) / dx[i] - (y[i] - y[i - 1]) / dx[i - 1]);
    }
    l[0] = l[N - 1] = 1.0;
    for (int i = 1; i < N - 1; i++)
    {
        l[i] = 2.0 * (x[i + 1] - x[i - 1]) - dx[i - 1] * mu[i - 1];
        mu[i] = dx[i] / l[i];
        z[i] = (alpha[i] - dx[i - 1] * z[i - 1]) / l[i];
    }
    y0[N - 1] = y[N - 1];
    y1[N - 1] = y2[N - 1] = y3[N - 1] = 0.0;
    for (int i = N - 2; i >= 0; i--)
    {
        y0[i] = y[i];
        y2[i] = z[i] - mu[i] * y2[i + 1];
        y1[i] = (y0[i + 1] - y0[i]) / dx[i] - dx[i] * (y2[i + 1] + 2.0 * y2[i]) / 3.0;
        y3[i] = (y2[i + 1] - y2[i]) / (3.0 * dx[i]);
    }
    free(dx); free(alpha); free(l); free(z); free(mu);
}

int orc_binary_search(const double *a, int lo, int hi, double x)
{
    /* largest index whose knot is <= x (recursive bisection that steps over
     * repeated knots); -1 when x is below the range */
    if (x == a[lo]) return lo;
    if (hi <= lo) return -1;
    int mid = lo + (hi - lo) / 2;
    int next = mid;
    while (a[mid] == a[next]) next++;
    if (x > a[mid] && x < a[next]) return mid;
    if (a[mid] >= x) return orc_binary_search(a, lo, mid, x);
    return orc_binary_search(a, next, hi, x);
}

int orc_even_search(const double *a, int lo, int hi, double x)
{
    double dx = a[1] - a[0];
    int n = (int)floor((x - a[0]) / dx);
    if (n < lo) n = lo;
    if (n > hi - 1) n = hi - 1;
    return n;
}

typedef struct
{
    int N;
    const double *x;
    const double *c;  /* y0 | y1 | y2 | y3, N each */
    int even;
} spline_t;

static double spline_eval(const spline_t *s, double x)
{
    int n = s->even ? orc_even_search(s->x, 0, s->N, x) : orc_binary_search(s->x, 0, s->N, x);
    const double *c = s->c;
    int N = s->N;
    double dx = x - s->x[n];
    return c[n] + c[N + n] * dx + c[2 * N + n] * dx * dx + c[3 * N + n] * dx * dx * dx;
}

double orc_spline_eval(int N, const double *x, const double *c, int even, double xv)
{
    spline_t s = {N, x, c, even};
    return spline_eval(&s, xv);
}

/* -------------------------------------------------- interpolated orbit -- */

int orc_iorbit_size(double Tobs)
{
    int N = (int)(W_SAMPLES_PER_YEAR * Tobs / W_YEAR);
    return N < 20 ? 20 : N;
}

void orc_iorbit_fill(const orc_orbit *o, double Tobs, double t0, int N, double *t, double *pos,
                     double *coef)
{
    double buffer = Tobs / (double)(N - 1);
    double cadence = (Tobs + 2.0 * buffer) / (double)(N - 1);
    double x[4], y[4], z[4];
    for (int n = 0; n < N; n++)
    {
        t[n] = t0 + n * cadence - buffer;
        orc_analytic_orbits(o, t[n], x, y, z);
        for (int i = 0; i < 3; i++)
        {
            pos[(3 * i + 0) * N + n] = x[i + 1];
            pos[(3 * i + 1) * N + n] = y[i + 1];
            pos[(3 * i + 2) * N + n] = z[i + 1];
        }
    }
    for (int s = 0; s < 9; s++) orc_spline_coefficients(N, t, pos + s * N, coef + 4 * s * N);
}

/* spacecraft sc, coordinate d (0 x, 1 y, 2 z) at time t */
static double orbit_pos(int Norb, const double *ot, const double *oc, int sc, int d, double t)
{
    spline_t s = {Norb, ot, oc + 4 * (3 * sc + d) * Norb, 0};
    return spline_eval(&s, t);
}

static void sc_to_ssb_time(int Norb, const double *ot, const double *oc, double costh, double phi,
                           const double *time, double *shifted, int N)
{
    double sinth = sqrt(1.0 - costh * costh);
    double cosph = w_cos(phi), sinph = w_sin(phi);
    double k[3] = {-sinth * cosph, -sinth * sinph, -costh};
    for (int n = 0; n < N; n++)
    {
        double r[3];
        for (int d = 0; d < 3; d++) r[d] = orbit_pos(Norb, ot, oc, 0, d, time[n]) / W_CLIGHT;
        double kdotr = 0.0;
        for (int i = 0; i < 3; i++) kdotr += k[i] * r[i];
        shifted[n] = time[n] - kdotr;
    }
}

/* ------------------------------------------------------ TDI response ---- */

/* The eight delayed one-arm terms of a Michelson-type TDI channel with
 * vertices (a,b,c) (LISA_TDI_spline, utils/src/glass_lisa.c:769-887):
 *   t_e = t - kr[v] - m1*L[i1] - m2*L[i2]
 *   M  += s1 * (h+ P1[j1] + hx Q1[j1]),  M += s2 * (h+ P2[j2] + hx Q2[j2])
 * with (P,Q) = (App,Acp) ("p") or (Apm,Acm) ("m").  Vertex / arm indices
 * are 0=a, 1=b, 2=c.  Multiplying an arm length by 1 or 0 is exact, so the
 * delays round exactly as the written-out expressions. */
typedef struct
{
    signed char v, i1, i2;
    double m1, m2;
    signed char s1, pm1, j1, s2, pm2, j2;   /* pm: 1 = plus-arm (App/Acp), 0 = minus */
} tdi_term;

static const tdi_term TDI8[8] = {
    {0, 2, 1, 2.0, 2.0, +1, 0, 1, -1, 1, 2},
    {1, 2, 1, 1.0, 2.0, -1, 0, 2, +1, 1, 2},
    {2, 1, 2, 1.0, 2.0, +1, 1, 1, -1, 0, 1},
    {0, 1, 0, 2.0, 0.0, -1, 0, 1, +1, 0, 2},
    {0, 2, 0, 2.0, 0.0, +1, 1, 2, -1, 1, 1},
    {2, 1, 0, 1.0, 0.0, -1, 1, 1, +1, 0, 1},
    {1, 2, 0, 1.0, 0.0, +1, 0, 2, -1, 1, 2},
    {0, 0, 0, 0.0, 0.0, -1, 0, 2, +1, 1, 1},
};

const void *orc_tdi_term_table(void) { return TDI8; }

static void accumulate(double *M, int s, double h1, double h2, double P, double Q)
{
    double v = h1 * P + h2 * Q;
    if (s > 0) *M += v;
    else *M -= v;
}

static void tdi_channel(double *M, double *Mf, const int vtx[3], double t, const spline_t *amp,
                        const spline_t *ph, double Ap, double Ac, double c2p, double s2p,
                        const double *App, const double *Apm, const double *Acp, const double *Acm,
                        const double *kr, const double *L)
{
    *M = 0.0;
    *Mf = 0.0;
    for (int e = 0; e < 8; e++)
    {
        const tdi_term *q = &TDI8[e];
        double te = t - kr[vtx[q->v]] - q->m1 * L[vtx[q->i1]] - q->m2 * L[vtx[q->i2]];
        double a = spline_eval(amp, te);
        double phase = spline_eval(ph, te);
        double cp = w_cos(phase), sp = w_sin(phase);
        double hp = a * (Ap * c2p * cp + Ac * s2p * sp);
        double hc = a * (Ac * c2p * sp - Ap * s2p * cp);
        double hpf = a * (-Ap * c2p * sp + Ac * s2p * cp);
        double hcf = a * (Ac * c2p * cp + Ap * s2p * sp);
        int j1 = vtx[q->j1], j2 = vtx[q->j2];
        double P1 = q->pm1 ? App[j1] : Apm[j1], Q1 = q->pm1 ? Acp[j1] : Acm[j1];
        double P2 = q->pm2 ? App[j2] : Apm[j2], Q2 = q->pm2 ? Acp[j2] : Acm[j2];
        accumulate(M, q->s1, hp, hc, P1, Q1);
        accumulate(M, q->s2, hp, hc, P2, Q2);
        accumulate(Mf, q->s1, hpf, hcf, P1, Q1);
        accumulate(Mf, q->s2, hpf, hcf, P2, Q2);
    }
}

void orc_extract_amplitude_and_phase(int Ns, double *As, double *Dphi, const double *M,
                                     const double *Mf, const double *phiR)
{
    double *flip = calloc(Ns, sizeof(double)), *pjump = calloc(Ns, sizeof(double));
    for (int i = 0; i < Ns; i++) As[i] = sqrt(M[i] * M[i] + Mf[i] * Mf[i]);
    flip[0] = 1.0;
    pjump[0] = 0.0;
    int i = 1;
    do
    {
        flip[i] = flip[i - 1];
        pjump[i] = pjump[i - 1];
        if (As[i] < As[i - 1] && As[i] < As[i + 1])
        {
            double d1 = As[i + 1] + As[i - 1] - 2.0 * As[i];
            double d2 = -As[i + 1] + As[i - 1] - 2.0 * As[i];
            double d3 = -As[i + 1] + As[i - 1] + 2.0 * As[i];
            if (fabs(d2 / d1) < 0.1)
            {
                flip[i + 1] = -1.0 * flip[i];
                pjump[i + 1] = pjump[i] + M_PI;
                i++;
            }
            if (fabs(d3 / d1) < 0.1)
            {
                flip[i] = -1.0 * flip[i - 1];
                pjump[i] = pjump[i - 1] + M_PI;
            }
        }
        i++;
    } while (i < Ns - 1);
    flip[Ns - 1] = flip[Ns - 2];
    pjump[Ns - 1] = pjump[Ns - 2];
    for (i = 0; i < Ns; i++)
    {
        As[i] = flip[i] * As[i];
        double v = remainder(phiR[i], W_PI2);
        Dphi[i] = -w_atan2(Mf[i], M[i]) + pjump[i] - v;
    }
    free(flip);
    free(pjump);
}

void orc_unwrap_phase(int N, double *phase)
{
    double v = phase[0];
    for (int i = 0; i < N; i++)
    {
        double u = phase[i];
        double q = rint(fabs(u - v) / W_PI2);
        if (q > 0.0)
        {
            if (v > u) u += q * W_PI2;
            else u -= q * W_PI2;
        }
        v = u;
        phase[i] = u;
    }
}

/* LISA_spline_response for a UCB (amp/phase splines, no frequency spline):
 * M, Mf of X,Y,Z at times tarray[0..N), then amplitude/phase extraction and
 * unwrapping.  amp3, ph3: 3 x N outputs. */
static void spline_response(int Norb, const double *ot, const double *oc, const double *tarray,
                            int N, double costh, double phi, double cosi, double psi,
                            const spline_t *amp, const spline_t *ph, const double *phase_ref,
                            double *amp3, double *ph3)
{
    double *R = calloc(6 * N, sizeof(double));   /* M: X,Y,Z then Mf: X,Y,Z */
    double c2p = w_cos(2. * psi), s2p = w_sin(2. * psi);
    double Ap = 0.5 * (1. + cosi * cosi), Ac = -cosi;
    double ep[4][4], ec[4][4], k[4];
    pol_tensor(costh, phi, ep, ec, k);
    static const int VTX[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
    for (int m = 0; m < N; m++)
    {
        double t = tarray[m];
        double x[3][3], n[3][3], L[3], kr[3], kn[3], dp[3], dc[3];
        double App[3], Apm[3], Acp[3], Acm[3];
        for (int i = 0; i < 3; i++)
            for (int d = 0; d < 3; d++) x[i][d] = orbit_pos(Norb, ot, oc, i, d, t) / W_CLIGHT;
        for (int i = 0; i < 3; i++)
        {
            n[0][i] = x[1][i] - x[2][i];
            n[1][i] = x[2][i] - x[0][i];
            n[2][i] = x[0][i] - x[1][i];
        }
        for (int i = 0; i < 3; i++)
        {
            L[i] = 0.0;
            for (int j = 0; j < 3; j++) L[i] += n[i][j] * n[i][j];
            L[i] = sqrt(L[i]);
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) n[i][j] /= L[i];
        for (int i = 0; i < 3; i++)
        {
            kr[i] = 0.0;
            for (int j = 0; j < 3; j++) kr[i] += k[j + 1] * x[i][j];
        }
        for (int i = 0; i < 3; i++)
        {
            kn[i] = 0.0;
            for (int j = 0; j < 3; j++) kn[i] += k[j + 1] * n[i][j];
        }
        for (int i = 0; i < 3; i++)
        {
            dp[i] = 0.0;
            dc[i] = 0.0;
            for (int j = 0; j < 3; j++)
                for (int l = 0; l < 3; l++)
                {
                    dp[i] += (n[i][j] * n[i][l]) * ep[j + 1][l + 1];
                    dc[i] += (n[i][j] * n[i][l]) * ec[j + 1][l + 1];
                }
        }
        for (int i = 0; i < 3; i++)
        {
            App[i] = dp[i] / (1.0 + kn[i]);
            Apm[i] = dp[i] / (1.0 - kn[i]);
            Acp[i] = dc[i] / (1.0 + kn[i]);
            Acm[i] = dc[i] / (1.0 - kn[i]);
        }
        for (int ch = 0; ch < 3; ch++)
            tdi_channel(&R[ch * N + m], &R[(3 + ch) * N + m], VTX[ch], t, amp, ph, Ap, Ac, c2p, s2p,
                        App, Apm, Acp, Acm, kr, L);
    }
    for (int ch = 0; ch < 3; ch++)
    {
        orc_extract_amplitude_and_phase(N, amp3 + ch * N, ph3 + ch * N, R + ch * N, R + (3 + ch) * N,
                                        phase_ref);
        orc_unwrap_phase(N, ph3 + ch * N);
    }
    free(R);
}

/* --------------------------------------------------- wavelet waveform --- */

void orc_ucb_wavelet_layers(const orc_wdm *w, double Tobs, const double *params, int *jstart,
                            int *jwidth)
{
    double fstart = params[0] / Tobs;
    double fstop = (params[0] / Tobs + params[7] / Tobs / Tobs);
    double fmin, fmax;
    if (fstart < fstop)
    {
        fmin = fstart * (1.0 - W_VEARTH);
        fmax = fstop * (1.0 + W_VEARTH);
    }
    else
    {
        fmin = fstop * (1.0 - W_VEARTH);
        fmax = fstart * (1.0 + W_VEARTH);
    }
    int j = (int)rint(fmin / w->df);
    double dfmin = fmin - j * w->df;
    double dfmax = fmax - j * w->df;
    int jmin = j, jmax = j;
    if (dfmin < 0.0 && fabs(dfmin) > w->A / W_PI2) jmin = j - 1;
    if (dfmax > 0.0 && fabs(dfmax) > w->A / W_PI2) jmax = j + 1;
    *jstart = jmin;
    *jwidth = jmax - jmin + 1;
}

static double ucb_phase(double t, const double *p, double T)
{
    double f0 = p[0] / T, phi0 = p[6], fdot = p[7] / T / T, fddot = 0.0;
    return -phi0 + W_PI2 * (f0 * t + 0.5 * fdot * t * t + 1.0 / 6.0 * fddot * t * t * t);
}

static double ucb_amplitude(double t, const double *p, double T)
{
    double f0 = p[0] / T, A0 = w_exp(p[3]), fdot = p[7] / T / T;
    return A0 * (1.0 + 2.0 / 3.0 * fdot / f0 * t);
}

/* ucb_barycenter_waveform (ucb/src/glass_ucb_waveform.c:355-362): the SSB
 * phase and amplitude of ucb_phase / ucb_amplitude (:332-353) at N times. */
void orc_ucb_barycenter_waveform(const double *params, int N, const double *times, double *phase,
                                 double *amp, double T)
{
    for (int n = 0; n < N; n++)
    {
        phase[n] = ucb_phase(times[n], params, T);
        amp[n] = ucb_amplitude(times[n], params, T);
    }
}

/* The interpolated TDI time series of ucb_waveform_wavelet
 * (ucb/src/glass_ucb_waveform.c:1064-1178): SSB phase/amplitude splines on
 * the orbit knots, spacecraft-1 phase, TDI amplitude/phase on Norb samples
 * and their splines, then N samples t_i = t0 + i*dt_s of
 * A(t) cos(Phi(t) - 2 pi fhet t_i) into out[3][N].  With N = NT(L+1),
 * dt_s = dt/(L+1), fhet = (jmin-1) df this is the heterodyned series the
 * waveform transforms; with N = NF*NT, dt_s = Tobs/N, fhet = 0 it is the
 * full-cadence series of ucb_waveform_wavelet_slow (:1243-1356). */
void orc_ucb_wavelet_series(int Norb, const double *ot, const double *oc, double Tobs, double t0,
                            const double *params, int N_ds, double dt_ds, double fhet, double *wave)
{
    int Ns = Norb;
    double dt = Tobs / (double)(Ns - 1);
    double *buf = calloc(5 * (size_t)Ns, sizeof(double));
    double *time_ssb = buf, *amp_ssb = buf + Ns, *phase_ssb = buf + 2 * Ns, *phase_sc = buf + 3 * Ns;
    double *time_sc = buf + 4 * Ns;
    for (int i = 0; i < Ns; i++) time_ssb[i] = t0 + i * dt;
    for (int n = 0; n < Ns; n++)
    {
        phase_ssb[n] = ucb_phase(ot[n], params, Tobs);
        amp_ssb[n] = ucb_amplitude(ot[n], params, Tobs);
    }
    double *cssb = malloc(8 * (size_t)Ns * sizeof(double));
    orc_spline_coefficients(Ns, ot, amp_ssb, cssb);
    orc_spline_coefficients(Ns, ot, phase_ssb, cssb + 4 * Ns);
    spline_t amp_s = {Ns, ot, cssb, 1}, ph_s = {Ns, ot, cssb + 4 * Ns, 1};

    double costh = params[1], phi = params[2];
    sc_to_ssb_time(Norb, ot, oc, costh, phi, time_ssb, time_sc, Ns);
    for (int i = 0; i < Ns; i++) phase_sc[i] = spline_eval(&ph_s, time_sc[i]);

    double *dsb = malloc(4 * (size_t)N_ds * sizeof(double));
    double *phase_ds = dsb, *time_ds = dsb + N_ds, *phase_het = dsb + 2 * (size_t)N_ds;
    double *tsc = dsb + 3 * (size_t)N_ds;
    for (int i = 0; i < N_ds; i++)
    {
        time_ds[i] = t0 + i * dt_ds;
        phase_het[i] = W_PI2 * fhet * time_ds[i];
    }
    sc_to_ssb_time(Norb, ot, oc, costh, phi, time_ds, tsc, N_ds);
    for (int i = 0; i < N_ds; i++) phase_ds[i] = spline_eval(&ph_s, tsc[i]);

    double *amp3 = malloc(6 * (size_t)Ns * sizeof(double)), *ph3 = amp3 + 3 * Ns;
    spline_response(Norb, ot, oc, time_ssb, Ns, costh, phi, params[4], params[5], &amp_s, &ph_s,
                    phase_sc, amp3, ph3);

    double *ctdi = malloc(24 * (size_t)Ns * sizeof(double));
    for (int ch = 0; ch < 3; ch++)
    {
        orc_spline_coefficients(Ns, time_ssb, amp3 + ch * Ns, ctdi + (2 * ch) * 4 * Ns);
        orc_spline_coefficients(Ns, time_ssb, ph3 + ch * Ns, ctdi + (2 * ch + 1) * 4 * Ns);
    }
    for (int ch = 0; ch < 3; ch++)
    {
        spline_t as = {Ns, time_ssb, ctdi + (2 * ch) * 4 * Ns, 0};
        spline_t ps = {Ns, time_ssb, ctdi + (2 * ch + 1) * 4 * Ns, 0};
        double *out = wave + (size_t)ch * N_ds;
        for (int n = 0; n < N_ds; n++)
        {
            double Amp = spline_eval(&as, time_ds[n]);
            double Phase = spline_eval(&ps, time_ds[n]);
            Phase += phase_ds[n];
            Phase -= phase_het[n];
            out[n] = Amp * w_cos(Phase);
        }
    }
    free(ctdi); free(amp3); free(dsb); free(cssb); free(buf);
}

/* The response stage alone (for stage-by-stage comparisons): barycentre
 * phase at the knots, phase at spacecraft 1, and the signed TDI amplitude and
 * unwrapped phase of X,Y,Z at the Norb response times (3 x Norb each). */
void orc_ucb_wavelet_response(int Norb, const double *ot, const double *oc, double Tobs, double t0,
                              const double *params, double *phase_ssb, double *phase_sc, double *amp3,
                              double *ph3)
{
    int Ns = Norb;
    double dt = Tobs / (double)(Ns - 1);
    double *time_ssb = malloc(3 * (size_t)Ns * sizeof(double));
    double *amp_ssb = time_ssb + Ns, *time_sc = amp_ssb + Ns;
    for (int i = 0; i < Ns; i++) time_ssb[i] = t0 + i * dt;
    for (int n = 0; n < Ns; n++)
    {
        phase_ssb[n] = ucb_phase(ot[n], params, Tobs);
        amp_ssb[n] = ucb_amplitude(ot[n], params, Tobs);
    }
    double *cssb = malloc(8 * (size_t)Ns * sizeof(double));
    orc_spline_coefficients(Ns, ot, amp_ssb, cssb);
    orc_spline_coefficients(Ns, ot, phase_ssb, cssb + 4 * Ns);
    spline_t amp_s = {Ns, ot, cssb, 1}, ph_s = {Ns, ot, cssb + 4 * Ns, 1};
    sc_to_ssb_time(Norb, ot, oc, params[1], params[2], time_ssb, time_sc, Ns);
    for (int i = 0; i < Ns; i++) phase_sc[i] = spline_eval(&ph_s, time_sc[i]);
    spline_response(Norb, ot, oc, time_ssb, Ns, params[1], params[2], params[4], params[5], &amp_s, &ph_s,
                    phase_sc, amp3, ph3);
    free(cssb);
    free(time_ssb);
}

/* A layer count forced on every template (0: ucb_wavelet_layers' own).  The
 * reference has no such switch; it is the timing harness's override of
 * Nlayers, applied after ucb_wavelet_layers as if that had returned it. */
static int g_force_layers = 0;

void orc_set_wavelet_layers(int L) { g_force_layers = L > 0 ? L : 0; }

int orc_ucb_waveform_wavelet_compact(const orc_wdm *w, int Norb, const double *ot,
                                     const double *oc, double Tobs, double t0, const double *params,
                                     int *list, int *Nwavelet, double *wX, double *wY, double *wZ)
{
    int min_layer, Nlayers;
    orc_ucb_wavelet_layers(w, Tobs, params, &min_layer, &Nlayers);
    if (g_force_layers > 0) Nlayers = g_force_layers;
    int Nt = w->NT;
    int N_ds = Nt * (Nlayers + 1);
    double dt_ds = w->dt / (double)(Nlayers + 1);
    double f0 = (min_layer - 1) * w->df;
    double *wave = malloc(3 * (size_t)N_ds * sizeof(double));
    orc_ucb_wavelet_series(Norb, ot, oc, Tobs, t0, params, N_ds, dt_ds, f0, wave);
    double *window = malloc((Nt / 2 + 1) * sizeof(double));
    orc_wdm_filter(window, Nlayers + 1, Nt, w->A, 1);
    for (int ch = 0; ch < 3; ch++)
        orc_wavelet_transform_timefreq_by_layers(w, min_layer, Nlayers, window, wave + ch * N_ds);
    int N = 0;
    for (int i = 0; i < Nt; i++)
        for (int j = min_layer; j < min_layer + Nlayers; j++)
        {
            int k = i + j * Nt;
            if (k >= w->kmin && k < w->kmax) list[N++] = k - w->kmin;
        }
    *Nwavelet = N;
    for (int n = 0; n < N; n++)
    {
        wX[n] = wave[n];
        wY[n] = wave[N_ds + n];
        wZ[n] = wave[2 * N_ds + n];
    }
    free(window);
    free(wave);
    return Nlayers;
}

int orc_ucb_waveform_wavelet(const orc_wdm *w, int Norb, const double *ot, const double *oc,
                             double Tobs, double t0, const double *params, int *list,
                             int *Nwavelet, double *X, double *Y, double *Z)
{
    int cap = (g_force_layers > 3 ? g_force_layers + 1 : 4) * w->NT;
    double *v = malloc(3 * (size_t)cap * sizeof(double));
    int L = orc_ucb_waveform_wavelet_compact(w, Norb, ot, oc, Tobs, t0, params, list, Nwavelet, v,
                                             v + cap, v + 2 * cap);
    for (int n = 0; n < *Nwavelet; n++)
    {
        X[list[n]] = v[n];
        Y[list[n]] = v[cap + n];
        Z[list[n]] = v[2 * cap + n];
    }
    free(v);
    return L;
}

/* ---------------------------------------------------- wavelet logL ------ */

double orc_snr_wavelet(const double *X, const double *Y, const double *Z, const double *invC,
                       int Npix, const int *list, int Nlist)
{
    /* invC: 9 planes of Npix, row-major 3x3 */
    double s = 0.0;
    s += orc_wavelet_nwip(X, X, invC + 0 * Npix, list, Nlist);
    s += orc_wavelet_nwip(Y, Y, invC + 4 * Npix, list, Nlist);
    s += orc_wavelet_nwip(Z, Z, invC + 8 * Npix, list, Nlist);
    s += orc_wavelet_nwip(X, Y, invC + 1 * Npix, list, Nlist) * 2;
    s += orc_wavelet_nwip(X, Z, invC + 2 * Npix, list, Nlist) * 2;
    s += orc_wavelet_nwip(Y, Z, invC + 5 * Npix, list, Nlist) * 2;
    return sqrt(s);
}

double orc_gaussian_log_likelihood_wavelet(const orc_wdm *w, const double *dX, const double *dY,
                                           const double *dZ, const double *mX, const double *mY,
                                           const double *mZ, const int *mlist, int Nlist,
                                           const double *invC, int Npix)
{
    double *r = malloc(3 * (size_t)Npix * sizeof(double));
    double *rX = r, *rY = r + Npix, *rZ = r + 2 * Npix;
    memcpy(rX, dX, Npix * sizeof(double));
    memcpy(rY, dY, Npix * sizeof(double));
    memcpy(rZ, dZ, Npix * sizeof(double));
    for (int n = 0; n < Nlist; n++)
    {
        int k = mlist[n];
        int i = (k + w->kmin) % w->NT;
        if (k >= 0 && k < Npix)
        {
            rX[k] -= mX[k];
            rY[k] -= mY[k];
            rZ[k] -= mZ[k];
            if (i < w->imin || i > w->imax) rX[k] = rY[k] = rZ[k] = 0.0;
        }
    }
    int *all = malloc(Npix * sizeof(int));
    for (int n = 0; n < Npix; n++) all[n] = n;
    double XX = 0.0, XY = 0.0;
    XX += orc_wavelet_nwip(rX, rX, invC + 0 * Npix, all, Npix);
    XX += orc_wavelet_nwip(rY, rY, invC + 4 * Npix, all, Npix);
    XX += orc_wavelet_nwip(rZ, rZ, invC + 8 * Npix, all, Npix);
    XY += orc_wavelet_nwip(rX, rY, invC + 1 * Npix, all, Npix);
    XY += orc_wavelet_nwip(rX, rZ, invC + 2 * Npix, all, Npix);
    XY += orc_wavelet_nwip(rY, rZ, invC + 5 * Npix, all, Npix);
    free(all);
    free(r);
    return -0.5 * (XX + 2.0 * XY);
}

void orc_wdm_layer_var(const orc_wdm *w, const double *dft_psd, double *layer_var)
{
    int NF = w->NF, NT = w->NT, half = NT / 2, nmax = NF * NT / 2;
    memset(layer_var, 0, (NF + 1) * sizeof(double));
    for (int l = -(half - 1); l < half; l++)
    {
        double w2 = w->window[abs(l)] * w->window[abs(l)];
        for (int m = 0; m <= NF; m++)
        {
            if (m == 0 && l < 0) continue;
            if (m == NF && l > 0) continue;
            double wt = w2;
            if (l == 0 && (m == 0 || m == NF)) wt *= 0.25;
            int k = abs(m * half + l) % (2 * nmax);
            if (k > nmax) k = 2 * nmax - k;
            layer_var[m] += wt * dft_psd[k];
        }
    }
    for (int m = 1; m < NF; m++) layer_var[m] /= 2.0 * NT * NT;
    layer_var[0] /= (double)NT * NT;
    layer_var[NF] /= (double)NT * NT;
}

/* ------------------------------------------------------ CPU baseline ---- */

/* Per template: the compact wavelet waveform, then against data d (3 planes
 * of Npix) and invC (9 planes of Npix):
 *   snr[e]   = snr_wavelet of the template over its pixel list;
 *   dlogL[e] = gaussian_log_likelhood_wavelet(d, model = template)
 *              - gaussian_log_likelhood_wavelet(d, empty model),
 * summed over the template's list only (pixels off the list do not change):
 *   interior pixel k: d_k^T C h_k - h_k^T C h_k / 2;
 *   edge pixel (time index < imin or > imax): the reference zeroes the
 *   residual there, so the change is + d_k^T C d_k / 2. */
double orc_wdm_eval_batch(const orc_wdm *w, int Norb, const double *ot, const double *oc,
                          double Tobs, double t0, int n, const double *params, const double *data,
                          const double *invC, int Npix, double *dlogL, double *snr, int *nwave,
                          int nthreads)
{
    double t_start = 0.0;
#ifdef _OPENMP
    extern double omp_get_wtime(void);
    extern void omp_set_num_threads(int);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    t_start = omp_get_wtime();
#endif
    int cap = (g_force_layers > 3 ? g_force_layers + 1 : 4) * w->NT;
#pragma omp parallel
    {
        int *list = malloc(cap * sizeof(int));
        double *h = malloc(3 * (size_t)cap * sizeof(double));
#pragma omp for schedule(dynamic)
        for (int e = 0; e < n; e++)
        {
            int Nw;
            orc_ucb_waveform_wavelet_compact(w, Norb, ot, oc, Tobs, t0, params + 8 * e, list, &Nw, h,
                                             h + cap, h + 2 * cap);
            double s2 = 0.0, dl = 0.0;
            for (int m = 0; m < Nw; m++)
            {
                int k = list[m];
                double hv[3] = {h[m], h[cap + m], h[2 * cap + m]};
                double hh = 0.0, dh = 0.0, dd = 0.0;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                    {
                        double c = invC[(3 * a + b) * Npix + k];
                        hh += hv[a] * hv[b] * c;
                        dh += data[a * Npix + k] * hv[b] * c;
                        dd += data[a * Npix + k] * data[b * Npix + k] * c;
                    }
                s2 += hh;
                int i = (k + w->kmin) % w->NT;
                if (k >= 0 && k < Npix)
                    dl += (i < w->imin || i > w->imax) ? 0.5 * dd : dh - 0.5 * hh;
            }
            snr[e] = sqrt(s2);
            dlogL[e] = dl;
            nwave[e] = Nw;
        }
        free(h);
        free(list);
    }
#ifdef _OPENMP
    return omp_get_wtime() - t_start;
#else
    return 0.0;
#endif
}

/* ucb_fisher_wavelet (ucb/src/glass_ucb_waveform.c:524-696) up to the
 * eigensolve.  Per parameter i: templates at params +- epsilon/2 (epsilon =
 * 1e-6 on the parameter array; cosines clipped to <= 1 on the plus and >= -1
 * on the minus side, :560-576), derivative (h+ - h-)/epsilon on the PLUS
 * template's pixel list only, at pixels whose time index lies strictly
 * inside (imin, imax) (:592-603); h- there is the minus template's value
 * or 0.  Fisher element (i <= j) = sum of nine wavelet_nwip over every pixel
 * in the order XX, YY, ZZ, XY, XZ, YZ, YX, ZX, ZY (:608-617), NaN -> 10.0.
 *
 * The reference's derivative arrays are zero outside the union of the plus
 * lists, so each of its sums over all data->N pixels adds exact zeros
 * there; summing over the window [klo, khi) of written pixels in ascending
 * pixel order gives the same result to the bit (for finite invC).
 * invC: 9 row-major planes of Npix.  Returns 0, or -1 on allocation
 * failure. */
#ifndef O_NP
#define O_NP 8
#endif

int orc_ucb_fisher_wavelet(const orc_wdm *w, int Norb, const double *ot, const double *oc, double Tobs,
                           double t0, const double *params, const double *invC, int Npix, double *fisher)
{
    const double epsilon = 1.0e-6;
    const double invstep = 1. / (epsilon);
    const int cap = 4 * w->NT;
    int *lp = malloc(16 * (size_t)cap * sizeof(int));
    int nw[16];
    double *vp = malloc(16 * 3 * (size_t)cap * sizeof(double));
    if (!lp || !vp)
    {
        free(lp);
        free(vp);
        return -1;
    }
    /* slot 2i: plus template of parameter i, 2i+1: minus */
    for (int i = 0; i < O_NP; i++)
        for (int sgn = 0; sgn < 2; sgn++)
        {
            double p[O_NP];
            memcpy(p, params, sizeof(p));
            if (sgn == 0) p[i] += epsilon / 2.;
            else p[i] -= epsilon / 2.;
            if (i == 1 || i == 4)
            {
                if (sgn == 0 && p[i] > 1.0) p[i] = 1.0;
                if (sgn == 1 && p[i] < -1.0) p[i] = -1.0;
            }
            int s = 2 * i + sgn;
            double *v = vp + (size_t)s * 3 * cap;
            orc_ucb_waveform_wavelet_compact(w, Norb, ot, oc, Tobs, t0, p, lp + (size_t)s * cap, &nw[s], v, v + cap,
                                             v + 2 * cap);
        }
    int klo = Npix, khi = 0;
    for (int s = 0; s < 16; s++)
        for (int n = 0; n < nw[s]; n++)
        {
            int k = lp[(size_t)s * cap + n];
            if (k < klo) klo = k;
            if (k + 1 > khi) khi = k + 1;
        }
    if (khi <= klo) klo = khi = 0;
    const int nloc = khi - klo;
    double *dh = calloc((size_t)O_NP * 3 * (nloc ? nloc : 1), sizeof(double));
    double *hm = malloc(3 * (size_t)(nloc ? nloc : 1) * sizeof(double));
    if (!dh || !hm)
    {
        free(lp);
        free(vp);
        free(dh);
        free(hm);
        return -1;
    }
    for (int i = 0; i < O_NP; i++)
    {
        const int sp = 2 * i, sm = 2 * i + 1;
        /* the minus template scattered into a zeroed array (:583-590) */
        memset(hm, 0, 3 * (size_t)nloc * sizeof(double));
        for (int n = 0; n < nw[sm]; n++)
        {
            int k = lp[(size_t)sm * cap + n];
            for (int ch = 0; ch < 3; ch++) hm[ch * nloc + (k - klo)] = vp[((size_t)sm * 3 + ch) * cap + n];
        }
        for (int n = 0; n < nw[sp]; n++)
        {
            int k = lp[(size_t)sp * cap + n];
            int i_wdm = (k + w->kmin) % w->NT;
            if (i_wdm > w->imin && i_wdm < w->imax && k >= 0 && k < Npix)
                for (int ch = 0; ch < 3; ch++)
                    dh[((size_t)i * 3 + ch) * nloc + (k - klo)] =
                        (vp[((size_t)sp * 3 + ch) * cap + n] - hm[ch * nloc + (k - klo)]) * invstep;
        }
    }
    /* (a, b) channel pairs of the nine wavelet_nwip terms, in order */
    static const int ta[9] = {0, 1, 2, 0, 0, 1, 1, 2, 2};
    static const int tb[9] = {0, 1, 2, 1, 2, 2, 0, 0, 1};
    for (int i = 0; i < O_NP; i++)
        for (int j = i; j < O_NP; j++)
        {
            double F = 0.0;
            for (int t = 0; t < 9; t++)
            {
                const double *a = dh + ((size_t)i * 3 + ta[t]) * nloc;
                const double *b = dh + ((size_t)j * 3 + tb[t]) * nloc;
                const double *c = invC + (size_t)(3 * ta[t] + tb[t]) * Npix + klo;
                double arg = 0.0;
                for (int q = 0; q < nloc; q++) arg += a[q] * b[q] * c[q];
                F += arg;
            }
            if (F != F) F = 10.0;
            fisher[i * O_NP + j] = F;
            fisher[j * O_NP + i] = F;
        }
    free(lp);
    free(vp);
    free(dh);
    free(hm);
    return 0;
}
