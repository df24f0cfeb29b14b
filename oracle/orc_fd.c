/*
 * orc_fd.c -- oracle for the Fourier-domain UCB path (test infrastructure
 * only: the parity checker and CPU baseline of the HIP library).
 *
 * Restates, in FP64 with the reference's operation order:
 *   analytic_orbits               utils/src/glass_lisa.c:77-124
 *   interpolate_orbits, initialize_numeric_orbit's arm length
 *                                 utils/src/glass_lisa.c:62-72,207-259
 *   initialize_analytic_orbit     utils/src/glass_lisa.c:125-136
 *   LISA_polarization_tensor      utils/src/glass_lisa.c:1566-1596
 *   LISA_detector_tensor          utils/src/glass_lisa.c:1521-1564
 *   ucb_waveform                  ucb/src/glass_ucb_waveform.c:738-1001
 *   LISA_tdi / _FF / _Sangria     utils/src/glass_lisa.c:458-710
 *   XYZ2AE                        utils/src/glass_lisa.c:429-441
 *   ucb_bandwidth / ucb_alignment ucb/src/glass_ucb_waveform.c:698-736
 *   AEnoise, rednoise             utils/src/glass_lisa.c:1183-1198
 *   get_noise_levels("sangria"), XYZnoise_FF, XYZcross_FF
 *                                 utils/src/glass_lisa.c:1082-1135
 *   GetNoiseModel (3 channels)    utils/src/glass_data.c:1156-1242
 *   invert_noise_covariance_matrix utils/src/glass_math.c:236-289
 *   fourier_nwip, wavelet_nwip, snr utils/src/glass_math.c:362-426
 *   gaussian_log_likelihood       utils/src/glass_model.c:726-774
 *   add_ucb_model                 ucb/src/glass_ucb_model.c:192-232
 *   ucb_fisher (before dgeev)     ucb/src/glass_ucb_waveform.c:364-522
 * Constants restate utils/src/glass_constants.h:31-89 and
 * utils/src/glass_lisa.h:31-49.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "glass_oracle.h"

#define O_PI2   6.283185307179586
#define O_SQ3   1.73205080757
#define O_CLIGHT 299792458.
#define O_YEAR  3.15581497632e7
#define O_AU    1.49597870700e11
#define O_LARM  2.5e9
#define O_SPS   8.321000e-23
#define O_SACC  9.000000e-30
#define O_SLOC  2.89e-24
#define O_NP    8

void orc_init_analytic_orbit(orc_orbit *o)
{
    o->L = O_LARM;
    o->fstar = O_CLIGHT / (2.0 * M_PI * O_LARM);
    double ecc = O_LARM / (2.0 * O_SQ3 * O_AU);
    o->R = O_AU * ecc;
}

void orc_analytic_orbits(const orc_orbit *o, double t, double *x, double *y, double *z)
{
    double alpha = O_PI2 * t / O_YEAR;
    double sa = sin(alpha), ca = cos(alpha);
    double sa2 = sa * sa, ca2 = ca * ca, saca = sa * ca;
    double AUca = O_AU * ca, AUsa = O_AU * sa;
    /* constellation angles beta = 0, 2pi/3, 4pi/3 with the reference's
     * literal sin/cos values */
    static const double SB[4] = {0, 0.0, 0.866025403784439, -0.866025403784438};
    static const double CB[4] = {0, 1.0, -0.5, -0.5};
    for (int i = 1; i <= 3; i++)
    {
        double sb = SB[i], cb = CB[i];
        x[i] = AUca + o->R * (saca * sb - (1. + sa2) * cb);
        y[i] = AUsa + o->R * (saca * cb - (1. + ca2) * sb);
        z[i] = -O_SQ3 * o->R * (ca * cb + sa * sb);
    }
}

void orc_polarization_tensor(double costh, double phi, double ep[4][4], double ec[4][4], double k[4])
{
    double u[4], v[4];
    double sinth = sqrt(1.0 - costh * costh);
    double cosph = cos(phi), sinph = sin(phi);
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

static void detector_tensor(double L, double ep[4][4], double ec[4][4], const double *x,
                            const double *y, const double *z, const double k[4],
                            double dp[4][4], double dc[4][4], double kr[4][4])
{
    double r12[4], r13[4], r23[4];
    r12[1] = (x[2] - x[1]) / L;  r13[1] = (x[3] - x[1]) / L;  r23[1] = (x[3] - x[2]) / L;
    r12[2] = (y[2] - y[1]) / L;  r13[2] = (y[3] - y[1]) / L;  r23[2] = (y[3] - y[2]) / L;
    r12[3] = (z[2] - z[1]) / L;  r13[3] = (z[3] - z[1]) / L;  r23[3] = (z[3] - z[2]) / L;
    double p12 = 0, p23 = 0, p13 = 0, c12 = 0, c23 = 0, c13 = 0;
    for (int i = 1; i <= 3; i++)
        for (int j = 1; j <= 3; j++)
        {
            p12 += r12[i] * r12[j] * ep[i][j];  c12 += r12[i] * r12[j] * ec[i][j];
            p23 += r23[i] * r23[j] * ep[i][j];  c23 += r23[i] * r23[j] * ec[i][j];
            p13 += r13[i] * r13[j] * ep[i][j];  c13 += r13[i] * r13[j] * ec[i][j];
        }
    dp[1][2] = dp[2][1] = p12;  dc[1][2] = dc[2][1] = c12;
    dp[2][3] = dp[3][2] = p23;  dc[2][3] = dc[3][2] = c23;
    dp[1][3] = dp[3][1] = p13;  dc[1][3] = dc[3][1] = c13;
    double k12 = 0, k13 = 0, k23 = 0;
    for (int i = 1; i <= 3; i++)
    {
        k12 += k[i] * r12[i];
        k13 += k[i] * r13[i];
        k23 += k[i] * r23[i];
    }
    kr[1][2] = k12;  kr[2][1] = -k12;
    kr[1][3] = k13;  kr[3][1] = -k13;
    kr[2][3] = k23;  kr[3][2] = -k23;
}

/* interpolate_orbits (utils/src/glass_lisa.c:62-72): the spacecraft
 * positions of an orbit read from a file (initialize_numeric_orbit,
 * :138-269), a natural cubic spline per coordinate over the file's knots,
 * evaluated after a binary search.  While a numeric orbit is set (knots t,
 * coefficients [9][4 Norb], row 3 sc + coordinate, as orc_spline_coefficients
 * writes them) the slow series reads it in place of analytic_orbits, as
 * ucb_waveform does through orbit->orbit_function. */
static int g_norb = 0;
static const double *g_ot = NULL, *g_oc = NULL;

void orc_set_numeric_orbit(int Norb, const double *t, const double *coef)
{
    g_norb = (t && coef) ? Norb : 0;
    g_ot = t;
    g_oc = coef;
}

static void orbit_positions(const orc_orbit *o, double t, double *x, double *y, double *z)
{
    if (!g_norb)
    {
        orc_analytic_orbits(o, t, x, y, z);
        return;
    }
    for (int i = 0; i < 3; i++)
    {
        x[i + 1] = orc_spline_eval(g_norb, g_ot, g_oc + 4 * (3 * i + 0) * g_norb, 0, t);
        y[i + 1] = orc_spline_eval(g_norb, g_ot, g_oc + 4 * (3 * i + 1) * g_norb, 0, t);
        z[i + 1] = orc_spline_eval(g_norb, g_ot, g_oc + 4 * (3 * i + 2) * g_norb, 0, t);
    }
}

/* initialize_numeric_orbit's average arm length (glass_lisa.c:207-259):
 * pos [9][N] rows x1 y1 z1 x2 y2 z2 x3 y3 z3 */
double orc_numeric_orbit_armlength(int N, const double *pos)
{
    double L12 = 0.0, L23 = 0.0, L31 = 0.0;
    for (int n = 0; n < N; n++)
    {
        double x1 = pos[0 * N + n], y1 = pos[1 * N + n], z1 = pos[2 * N + n];
        double x2 = pos[3 * N + n], y2 = pos[4 * N + n], z2 = pos[5 * N + n];
        double x3 = pos[6 * N + n], y3 = pos[7 * N + n], z3 = pos[8 * N + n];
        L12 += sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
        L23 += sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2) + (z3 - z2) * (z3 - z2));
        L31 += sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
    }
    L12 /= (double)N;
    L31 /= (double)N;
    L23 /= (double)N;
    return (L12 + L31 + L23) / 3.;
}

/* Slow series of the six links, FFT'd, shifted and normalised.
 * glass_ucb_waveform.c:804-962.  d: [6][2*BW], link order 12,21,31,13,23,32 */
void orc_ucb_slow_series(const orc_orbit *orbit, double T, double t0, const double *params,
                         int NParams, int BW, double *d)
{
    double k[4], ep[4][4], ec[4][4], dp[4][4], dc[4][4], kr[4][4];
    double x[4], y[4], z[4];
    double xi[4], fonfs[4];
    double TR[4][4], TI[4][4];

    double f0 = params[0] / T;
    double costh = params[1], phi = params[2], amp = exp(params[3]);
    double cosi = params[4], psi = params[5], phi0 = params[6];
    double dfdt = 0.0, d2fdt2 = 0.0;
    if (NParams > 7) dfdt = params[7] / (T * T);
    if (NParams > 8) d2fdt2 = params[8] / (T * T * T);
    long q = (long)(f0 * T);
    double cosps = cos(2. * psi), sinps = sin(2. * psi);
    double Aplus = amp * (1. + cosi * cosi) * 0.5;
    double Across = -amp * (cosi);
    double df = O_PI2 * (((double)q) / T);
    double DPr = Aplus * cosps, DPi = -Across * sinps;
    double DCr = -Aplus * sinps, DCi = -Across * cosps;
    orc_polarization_tensor(costh, phi, ep, ec, k);

    double *s = (double *)calloc((size_t)6 * 2 * BW, sizeof(double));
    for (int n = 1; n <= BW; n++)
    {
        double t = t0 + T * (double)(n - 1) / (double)BW;
        orbit_positions(orbit, t, x, y, z);
        detector_tensor(orbit->L, ep, ec, x, y, z, k, dp, dc, kr);
        for (int i = 1; i <= 3; i++)
        {
            double kdotx = (x[i] * k[1] + y[i] * k[2] + z[i] * k[3]) / O_CLIGHT;
            xi[i] = t - kdotx;
            double fi = f0;
            if (NParams > 7) fi += dfdt * xi[i];
            if (NParams > 8) fi += 0.5 * d2fdt2 * xi[i] * xi[i];
            fonfs[i] = fi / orbit->fstar;
            double arg2 = O_PI2 * f0 * xi[i] - phi0 - df * t;
            if (NParams > 7) arg2 += M_PI * dfdt * xi[i] * xi[i];
            if (NParams > 8) arg2 += (M_PI / 3.0) * d2fdt2 * xi[i] * xi[i] * xi[i];
            double aevol = 1.0;
            if (NParams > 7) aevol += 0.66666666666666666666 * dfdt / f0 * xi[i];
            for (int j = 1; j <= 3; j++)
            {
                if (i == j) continue;
                double arg1 = 0.5 * fonfs[i] * (1.0 + kr[i][j]);
                double sinc = 0.25 * sin(arg1) / arg1;
                double t1r = aevol * (dp[i][j] * DPr + dc[i][j] * DCr);
                double t1i = aevol * (dp[i][j] * DPi + dc[i][j] * DCi);
                double t2r = cos(arg1 - arg2);
                double t2i = sin(arg1 - arg2);
                TR[i][j] = sinc * (t1r * t2r + t1i * t2i);
                TI[i][j] = sinc * (-t1r * t2i + t1i * t2r);
            }
        }
        int re = 2 * (n - 1), im = re + 1;
        const int li[6] = {1, 2, 3, 1, 2, 3}, lj[6] = {2, 1, 1, 3, 3, 2};
        for (int l = 0; l < 6; l++)
        {
            s[l * 2 * BW + re] = TR[li[l]][lj[l]];
            s[l * 2 * BW + im] = TI[li[l]][lj[l]];
        }
    }
    double invBW2 = 1. / (double)(2 * BW);
    for (int l = 0; l < 6; l++)
    {
        double *sl = s + (size_t)l * 2 * BW, *dl = d + (size_t)l * 2 * BW;
        orc_forward_complex_fft(sl, BW);
        /* swap halves (negative frequencies first) and normalise */
        for (int i = 0; i < BW; i++)
        {
            dl[i] = sl[i + BW] * invBW2;
            dl[i + BW] = sl[i] * invBW2;
        }
    }
    free(s);
}

static void xyz2ae(double X, double Y, double Z, double *A, double *E)
{
    double invSQ6 = 0.408248290463863;
    *A = (Z - X) * M_SQRT1_2;
    *E = (X - 2 * Y + Z) * invSQ6;
}

/* d: [6][2BW] in link order 12,21,31,13,23,32 ; returns pointer to link (i,j) */
#define D(i, j, idx) (dd[lk[i][j]][idx])

int orc_ucb_waveform(const orc_orbit *orbit, const char *format, double T, double t0,
                     const double *params, int NParams, double *X, double *Y, double *Z,
                     double *A, double *E, int BW, int NI)
{
    int mode;
    if (strcmp("phase", format) == 0) mode = 0;
    else if (strcmp("frequency", format) == 0) mode = 1;
    else if (strcmp("sangria", format) == 0) mode = 2;
    else return -1;

    double *d = (double *)malloc(sizeof(double) * 12 * (size_t)BW);
    orc_ucb_slow_series(orbit, T, t0, params, NParams, BW, d);
    const double *dd[6];
    for (int l = 0; l < 6; l++) dd[l] = d + (size_t)l * 2 * BW;
    int lk[4][4];
    lk[1][2] = 0; lk[2][1] = 1; lk[3][1] = 2; lk[1][3] = 3; lk[2][3] = 4; lk[3][2] = 5;

    double f0 = params[0] / T;
    long q = (long)(f0 * T);
    int BWon2 = BW / 2;
    double sqT = sqrt(T);
    double invfstar = 1. / orbit->fstar;
    double dPhi = invfstar / T;
    double cosdPhi = cos(dPhi), sindPhi = sin(dPhi);
    double norm = 4.0 * invfstar * sqT;

    /* recursion seeded one bin below the band (glass_lisa.c:476-485) */
    double f = ((double)(q + 0 - 1 - BWon2)) / T;
    double fonfs = f * invfstar;
    double c1 = cos(fonfs), s1 = sin(fonfs), c2, s2, c3, s3;

    for (int i = 1; i <= BW; i++)
    {
        int j = 2 * (i - 1), k = j + 1; /* 0-indexed re/im of bin i-1 */
        f = ((double)(q + i - 1 - BWon2)) / T;
        fonfs = f * invfstar;
        double nr = c1 * cosdPhi - s1 * sindPhi;
        double ni = s1 * cosdPhi + c1 * sindPhi;
        c1 = nr;
        s1 = ni;
        c2 = 2. * c1 * c1 - 1.0;
        s2 = 2. * s1 * c1;
        if (mode == 2)
        {
            double pf = f * norm * s1;
            X[j] = pf * ((D(1,2,j) - D(1,3,j)) * c2 + (D(1,2,k) - D(1,3,k)) * s2 +
                         (D(2,1,j) - D(3,1,j)) * c1 + (D(2,1,k) - D(3,1,k)) * s1);
            X[k] = pf * ((D(1,2,k) - D(1,3,k)) * c2 - (D(1,2,j) - D(1,3,j)) * s2 +
                         (D(2,1,k) - D(3,1,k)) * c1 - (D(2,1,j) - D(3,1,j)) * s1);
            Y[j] = pf * ((D(2,3,j) - D(2,1,j)) * c2 + (D(2,3,k) - D(2,1,k)) * s2 +
                         (D(3,2,j) - D(1,2,j)) * c1 + (D(3,2,k) - D(1,2,k)) * s1);
            Y[k] = pf * ((D(2,3,k) - D(2,1,k)) * c2 - (D(2,3,j) - D(2,1,j)) * s2 +
                         (D(3,2,k) - D(1,2,k)) * c1 - (D(3,2,j) - D(1,2,j)) * s1);
            Z[j] = pf * ((D(3,1,j) - D(3,2,j)) * c2 + (D(3,1,k) - D(3,2,k)) * s2 +
                         (D(1,3,j) - D(2,3,j)) * c1 + (D(1,3,k) - D(2,3,k)) * s1);
            Z[k] = pf * ((D(3,1,k) - D(3,2,k)) * c2 - (D(3,1,j) - D(3,2,j)) * s2 +
                         (D(1,3,k) - D(2,3,k)) * c1 - (D(1,3,j) - D(2,3,j)) * s1);
            xyz2ae(X[j], Y[j], Z[j], &A[j], &E[j]);
            xyz2ae(X[k], Y[k], Z[k], &A[k], &E[k]);
            continue;
        }
        c3 = 2. * c1 * c2 - c1;
        s3 = 2. * s1 * c2 + s1;
        double pf = (mode == 0) ? sqT : sqT * (2. * fonfs);
        X[j] = pf * ((D(1,2,j) - D(1,3,j)) * c3 + (D(1,2,k) - D(1,3,k)) * s3 +
                     (D(2,1,j) - D(3,1,j)) * c2 + (D(2,1,k) - D(3,1,k)) * s2 +
                     (D(1,3,j) - D(1,2,j)) * c1 + (D(1,3,k) - D(1,2,k)) * s1 +
                     (D(3,1,j) - D(2,1,j)));
        X[k] = pf * ((D(1,2,k) - D(1,3,k)) * c3 - (D(1,2,j) - D(1,3,j)) * s3 +
                     (D(2,1,k) - D(3,1,k)) * c2 - (D(2,1,j) - D(3,1,j)) * s2 +
                     (D(1,3,k) - D(1,2,k)) * c1 - (D(1,3,j) - D(1,2,j)) * s1 +
                     (D(3,1,k) - D(2,1,k)));
        if (NI > 1)
        {
            Y[j] = pf * ((D(2,3,j) - D(2,1,j)) * c3 + (D(2,3,k) - D(2,1,k)) * s3 +
                         (D(3,2,j) - D(1,2,j)) * c2 + (D(3,2,k) - D(1,2,k)) * s2 +
                         (D(2,1,j) - D(2,3,j)) * c1 + (D(2,1,k) - D(2,3,k)) * s1 +
                         (D(1,2,j) - D(3,2,j)));
            Y[k] = pf * ((D(2,3,k) - D(2,1,k)) * c3 - (D(2,3,j) - D(2,1,j)) * s3 +
                         (D(3,2,k) - D(1,2,k)) * c2 - (D(3,2,j) - D(1,2,j)) * s2 +
                         (D(2,1,k) - D(2,3,k)) * c1 - (D(2,1,j) - D(2,3,j)) * s1 +
                         (D(1,2,k) - D(3,2,k)));
            Z[j] = pf * ((D(3,1,j) - D(3,2,j)) * c3 + (D(3,1,k) - D(3,2,k)) * s3 +
                         (D(1,3,j) - D(2,3,j)) * c2 + (D(1,3,k) - D(2,3,k)) * s2 +
                         (D(3,2,j) - D(3,1,j)) * c1 + (D(3,2,k) - D(3,1,k)) * s1 +
                         (D(2,3,j) - D(1,3,j)));
            Z[k] = pf * ((D(3,1,k) - D(3,2,k)) * c3 - (D(3,1,j) - D(3,2,j)) * s3 +
                         (D(1,3,k) - D(2,3,k)) * c2 - (D(1,3,j) - D(2,3,j)) * s2 +
                         (D(3,2,k) - D(3,1,k)) * c1 - (D(3,2,j) - D(3,1,j)) * s1 +
                         (D(2,3,k) - D(1,3,k)));
            xyz2ae(X[j], Y[j], Z[j], &A[j], &E[j]);
            xyz2ae(X[k], Y[k], Z[k], &A[k], &E[k]);
            if (mode == 0)
            {
                A[k] = -A[k];
                E[k] = -E[k];
            }
        }
    }
    free(d);
    return 0;
}

/* ------------------------------------------------------------ noise ---- */
static double ipow(double x, int n)
{
    double xn = 1.0;
    for (int i = 0; i < n; i++) xn *= x;
    return xn;
}

static double rednoise(double f)
{
    return 1. + 16.0 * (pow((2.0e-5 / f), 10.0) + ipow(1.0e-4 / f, 2));
}

double orc_AEnoise(double L, double fstar, double f)
{
    double fonfstar = f / fstar;
    double trans = ipow(sin(fonfstar), 2);
    return (16.0 / 3.0) * trans *
           ((2.0 + cos(fonfstar)) * (O_SPS + O_SLOC) +
            2.0 * (3.0 + 2.0 * cos(fonfstar) + cos(2.0 * fonfstar)) *
                (O_SLOC / 2.0 + O_SACC / ipow(O_PI2 * f, 4) * rednoise(f))) /
           ipow(2.0 * L, 2);
}

void orc_noise_levels_sangria(double f, double *Spm, double *Sop)
{
    *Spm = 5.76e-30 / (O_PI2 * f * O_CLIGHT) / (O_PI2 * f * O_CLIGHT) * (1.0 + pow(0.4e-3 / f, 2)) *
           (1.0 + pow(f / 8.0e-3, 4));
    *Sop = 1.28e-22 * (O_PI2 * f / O_CLIGHT) * (O_PI2 * f / O_CLIGHT) * (1.0 + pow(2.0e-3 / f, 4));
}

static double noise_transfer(double x)
{
    double s = sin(x);
    return s * s;
}

double orc_XYZnoise_FF(double fstar, double f, double Spm, double Sop)
{
    double x = f / fstar, cx = cos(x);
    return 16. * noise_transfer(x) * (2. * (1.0 + cx * cx) * Spm + Sop);
}

double orc_XYZcross_FF(double fstar, double f, double Spm, double Sop)
{
    double x = f / fstar, cx = cos(x);
    return -8. * noise_transfer(x) * cx * (4. * Spm + Sop);
}

void orc_xyz_noise_covariance(const orc_orbit *o, double T, double fmin, int N, double *C)
{
    for (int n = 0; n < N; n++)
    {
        double f = fmin + (double)(n) / T, Spm, Sop;
        orc_noise_levels_sangria(f, &Spm, &Sop);
        double dg = orc_XYZnoise_FF(o->fstar, f, Spm, Sop);
        double cr = orc_XYZcross_FF(o->fstar, f, Spm, Sop);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) C[(a * 3 + b) * N + n] = ((a == b) ? dg : cr) / 4.;
    }
}

static double signed_log_div(double num, double logdetC)
{
    return copysign(exp(log(fabs(num)) - logdetC), num);
}

void orc_invert_noise_covariance(const double *C, double *invC, double *logdetC, int N)
{
#define CC(a, b) C[((a) * 3 + (b)) * N + n]
#define IC(a, b) invC[((a) * 3 + (b)) * N + n]
    for (int n = 0; n < N; n++)
    {
        double cxx = CC(0, 0), cyy = CC(1, 1), czz = CC(2, 2);
        double cxy = CC(0, 1), cxz = CC(0, 2), cyz = CC(1, 2);
        double d1 = cxx;
        double d2 = cyy - cxy * cxy / d1;
        double t = cyz - cxy * cxz / d1;
        double d3 = (czz - cxz * cxz / d1) - (t * t) / d2;
        double ld = log(d1) + log(d2) + log(d3);
        if (logdetC) logdetC[n] = ld;
        IC(0, 0) = signed_log_div(cyy * czz - cyz * cyz, ld);
        IC(1, 1) = signed_log_div(czz * cxx - cxz * cxz, ld);
        IC(2, 2) = signed_log_div(cxx * cyy - cxy * cxy, ld);
        IC(0, 1) = signed_log_div(cxz * cyz - czz * cxy, ld);
        IC(0, 2) = signed_log_div(cxy * cyz - cxz * cyy, ld);
        IC(1, 2) = signed_log_div(cxy * cxz - cxx * cyz, ld);
        IC(1, 0) = IC(0, 1);
        IC(2, 0) = IC(0, 2);
        IC(2, 1) = IC(1, 2);
    }
#undef CC
#undef IC
}

/* ---------------------------------------------------- bandwidth ------- */
static double analytic_snr(double A, double Sn, double Sf, double sqT)
{
    return A * sqT * Sf / sqrt(Sn);
}

int orc_ucb_bandwidth(double L, double fstar, double f, double fdot, double costheta, double A,
                      double T, int N)
{
    int Nmin = 16;
    int Nmax = (int)pow(2, (int)log2((double)(N / 2)));
    double sqT = sqrt(T);
    double sf = sin(f / fstar);
    double sn = orc_AEnoise(L, fstar, f);
    double sintheta = sin(acos(costheta));
    double bw = 2 * T * ((4. + O_PI2 * f * (O_AU / O_CLIGHT) * sintheta) / O_YEAR + fabs(fdot) * T);
    int DS = (int)pow(2, (int)log2(bw - 1) + 1);
    if (DS > Nmax) DS = Nmax;
    if (DS < Nmin) DS = Nmin;
    double SNRm = analytic_snr(A, sn, sf, sqT);
    int SS = (int)pow(2, (int)log2(SNRm - 1) + 1);
    if (SS > Nmax) SS = Nmax;
    if (SS < Nmin) SS = Nmin;
    return (DS > SS) ? DS : SS;
}

void orc_ucb_alignment(const orc_orbit *o, double T, int NFFT, int data_qmin, const double *params,
                       int *BW, int *qmin, int *imin)
{
    double f0 = params[0] / T;
    double costheta = params[1];
    double amp = exp(params[3]);
    double dfdt = params[7] / (T * T);
    *BW = 2 * orc_ucb_bandwidth(o->L, o->fstar, f0, dfdt, costheta, amp, T, NFFT);
    *qmin = (int)(f0 * T) - *BW / 2;
    *imin = *qmin - data_qmin;
}

/* ------------------------------------------------- inner products ----- */
double orc_fourier_nwip(const double *a, const double *b, const double *invC, int N)
{
    double arg = 0.0;
    for (int n = 0; n < N; n++)
    {
        int j = n * 2, k = j + 1;
        double product = a[j] * b[j] + a[k] * b[k];
        arg += product * invC[n];
    }
    return (2.0 * arg);
}

double orc_wavelet_nwip(const double *a, const double *b, const double *invC, const int *list, int N)
{
    double arg = 0.0;
    for (int n = 0; n < N; n++)
    {
        int k = list[n];
        arg += a[k] * b[k] * invC[k];
    }
    return arg;
}

double orc_snr(const double *X, const double *Y, const double *Z, const double *invC, int Ninv, int N)
{
    const double *ic00 = invC, *ic11 = invC + 4 * (size_t)Ninv, *ic22 = invC + 8 * (size_t)Ninv;
    const double *ic01 = invC + 1 * (size_t)Ninv, *ic02 = invC + 2 * (size_t)Ninv,
                 *ic12 = invC + 5 * (size_t)Ninv;
    double s = 0.0;
    s += orc_fourier_nwip(X, X, ic00, N);
    s += orc_fourier_nwip(Y, Y, ic11, N);
    s += orc_fourier_nwip(Z, Z, ic22, N);
    s += orc_fourier_nwip(X, Y, ic01, N) * 2.;
    s += orc_fourier_nwip(X, Z, ic02, N) * 2.;
    s += orc_fourier_nwip(Y, Z, ic12, N) * 2.;
    return sqrt(s);
}

double orc_gaussian_log_likelihood(const double *dX, const double *dY, const double *dZ,
                                   const double *dA, const double *dE, const double *mX,
                                   const double *mY, const double *mZ, const double *mA,
                                   const double *mE, const double *invC, int NFFT, int Nchannel)
{
    size_t N2 = 2 * (size_t)NFFT;
    double *r = (double *)malloc(sizeof(double) * 3 * N2);
    double chi2 = 0.0;
#define PL(a, b) (invC + ((a) * 3 + (b)) * (size_t)NFFT)
    if (Nchannel == 3)
    {
        double *rX = r, *rY = r + N2, *rZ = r + 2 * N2;
        for (size_t i = 0; i < N2; i++)
        {
            rX[i] = dX[i] - mX[i];
            rY[i] = dY[i] - mY[i];
            rZ[i] = dZ[i] - mZ[i];
        }
        chi2 += orc_fourier_nwip(rX, rX, PL(0, 0), NFFT);
        chi2 += orc_fourier_nwip(rY, rY, PL(1, 1), NFFT);
        chi2 += orc_fourier_nwip(rZ, rZ, PL(2, 2), NFFT);
        chi2 += 2.0 * orc_fourier_nwip(rX, rY, PL(0, 1), NFFT);
        chi2 += 2.0 * orc_fourier_nwip(rX, rZ, PL(0, 2), NFFT);
        chi2 += 2.0 * orc_fourier_nwip(rY, rZ, PL(1, 2), NFFT);
    }
    else if (Nchannel == 2)
    {
        double *rA = r, *rE = r + N2;
        for (size_t i = 0; i < N2; i++)
        {
            rA[i] = dA[i] - mA[i];
            rE[i] = dE[i] - mE[i];
        }
        chi2 += orc_fourier_nwip(rA, rA, PL(0, 0), NFFT);
        chi2 += orc_fourier_nwip(rE, rE, PL(1, 1), NFFT);
    }
    else
    {
        double *rX = r;
        for (size_t i = 0; i < N2; i++) rX[i] = dX[i] - mX[i];
        chi2 += orc_fourier_nwip(rX, rX, PL(0, 0), NFFT);
    }
#undef PL
    free(r);
    return -0.5 * chi2;
}

void orc_add_ucb_model(double *mX, double *mY, double *mZ, const double *X, const double *Y,
                       const double *Z, int BW, int imin, int NFFT)
{
    for (int i = 0; i < BW; i++)
    {
        int j = i + imin;
        if (j > -1 && j < NFFT)
        {
            mX[2 * j] += X[2 * i];  mX[2 * j + 1] += X[2 * i + 1];
            mY[2 * j] += Y[2 * i];  mY[2 * j + 1] += Y[2 * i + 1];
            mZ[2 * j] += Z[2 * i];  mZ[2 * j + 1] += Z[2 * i + 1];
        }
    }
}

/* ------------------------------------------------------------ Fisher -- */
void orc_ucb_fisher_fd(const orc_orbit *o, const char *format, double T, double t0, int NFFT,
                       int data_qmin, const double *params, const double *invC, double *fisher)
{
    const double eps = 1.0e-6, inv = 1. / eps;
    size_t N2 = 2 * (size_t)NFFT;
    int BW0, qmin0, imin0, BWp = 0, qp, ip;
    /* the unperturbed template as the sampler stores it (sized 2*NFFT, zero
     * beyond its BW) */
    double *h = (double *)calloc(5 * N2, sizeof(double));
    double *hp = (double *)calloc(5 * N2, sizeof(double));
    double *dh = (double *)calloc((size_t)O_NP * 3 * N2, sizeof(double));
    orc_ucb_alignment(o, T, NFFT, data_qmin, params, &BW0, &qmin0, &imin0);
    orc_ucb_waveform(o, format, T, t0, params, O_NP, h, h + N2, h + 2 * N2, h + 3 * N2, h + 4 * N2, BW0, 3);
    for (int i = 0; i < O_NP; i++)
    {
        double pp[O_NP];
        for (int j = 0; j < O_NP; j++) pp[j] = params[j];
        pp[i] += eps;
        if (i == 1 || i == 4)
            if (pp[i] > 1.0) pp[i] = 1.0;
        memset(hp, 0, sizeof(double) * 5 * N2);
        orc_ucb_alignment(o, T, NFFT, data_qmin, pp, &BWp, &qp, &ip);
        orc_ucb_waveform(o, format, T, t0, pp, O_NP, hp, hp + N2, hp + 2 * N2, hp + 3 * N2, hp + 4 * N2, BWp, 3);
        for (int c = 0; c < 3; c++)
            for (int n = 0; n < 2 * BWp; n++)
                dh[((size_t)i * 3 + c) * N2 + n] = (hp[c * N2 + n] - h[c * N2 + n]) * inv;
    }
    /* Gram over the LAST perturbed BW with invC from window bin 0 */
    for (int i = 0; i < O_NP; i++)
        for (int j = i; j < O_NP; j++)
        {
            const double *di[3], *dj[3];
            for (int c = 0; c < 3; c++)
            {
                di[c] = dh + ((size_t)i * 3 + c) * N2;
                dj[c] = dh + ((size_t)j * 3 + c) * N2;
            }
#define IP(a, b) (invC + ((a) * 3 + (b)) * (size_t)NFFT)
            double F = orc_fourier_nwip(di[0], dj[0], IP(0, 0), BWp);
            F += orc_fourier_nwip(di[1], dj[1], IP(1, 1), BWp);
            F += orc_fourier_nwip(di[2], dj[2], IP(2, 2), BWp);
            F += orc_fourier_nwip(di[0], dj[1], IP(0, 1), BWp);
            F += orc_fourier_nwip(di[0], dj[2], IP(0, 2), BWp);
            F += orc_fourier_nwip(di[1], dj[2], IP(1, 2), BWp);
            F += orc_fourier_nwip(di[1], dj[0], IP(1, 0), BWp);
            F += orc_fourier_nwip(di[2], dj[0], IP(2, 0), BWp);
            F += orc_fourier_nwip(di[2], dj[1], IP(2, 1), BWp);
#undef IP
            if (F != F) F = 10.0;
            fisher[i * O_NP + j] = F;
            fisher[j * O_NP + i] = F;
        }
    free(h);
    free(hp);
    free(dh);
}
