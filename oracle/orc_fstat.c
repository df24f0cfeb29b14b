/*
 * orc_fstat.c -- CPU restatement of the F-statistic of the UCB
 * F-statistic proposal (test infrastructure only).
 *
 * Reference functions restated (paths relative to the reference root):
 *   get_Fstat_logL                ucb/src/glass_ucb_fstatistic.c:646-703
 *   get_filters                   ucb/src/glass_ucb_fstatistic.c:231-348
 *   get_N, init_M_matrix, get_M   ucb/src/glass_ucb_fstatistic.c:172-517
 *   calc_Fstat_logL, calc_a_i     ucb/src/glass_ucb_fstatistic.c:599-644
 *   get_F_params                  ucb/src/glass_ucb_fstatistic.c:525-597
 *   ucb_fdot                      ucb/src/glass_ucb_waveform.c:307-314
 *   setup_ / build_fstatistic_proposal  ucb/src/glass_ucb_proposal.c:916-1180
 * The 4x4 inverse (invert_matrix, LAPACKE dgetrf/dgetri in the reference,
 * utils/src/glass_math.c:501-523) is Gauss-Jordan elimination with partial
 * pivoting here and on the device.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "glass_oracle.h"

static void invert4(double m[4][4])
{
    double a[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
        {
            a[i][j] = m[i][j];
            a[i][4 + j] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++)
    {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
        if (p != c)
            for (int j = 0; j < 8; j++)
            {
                double t = a[c][j];
                a[c][j] = a[p][j];
                a[p][j] = t;
            }
        const double piv = a[c][c];
        for (int j = 0; j < 8; j++) a[c][j] = a[c][j] / piv;
        for (int r = 0; r < 4; r++)
        {
            if (r == c) continue;
            const double f = a[r][c];
            for (int j = 0; j < 8; j++) a[r][j] = a[r][j] - f * a[c][j];
        }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) m[i][j] = a[i][4 + j];
}

void orc_invert4(const double *in, double *out)
{
    double m[4][4];
    memcpy(m, in, sizeof(m));
    invert4(m);
    memcpy(out, m, sizeof(m));
}

static int sgn(double v) { return v < 0 ? -1 : (v > 0 ? 1 : 0); }

static double fstat_value(double M[4][4], const double *N)
{
    return 0.5 * (M[0][0] * N[0] * N[0] + 2 * M[0][1] * N[0] * N[1] + 2 * M[0][2] * N[0] * N[2] +
                  2 * M[0][3] * N[0] * N[3] + M[1][1] * N[1] * N[1] + 2 * M[1][2] * N[1] * N[2] +
                  2 * M[1][3] * N[1] * N[3] + M[2][2] * N[2] * N[2] + 2 * M[2][3] * N[2] * N[3] +
                  M[3][3] * N[3] * N[3]);
}

static void fstat_params(double M[4][4], const double *N, double *out)
{
    double a[4];
    for (int i = 0; i < 4; i++) a[i] = M[i][0] * N[0] + M[i][1] * N[1] + M[i][2] * N[2] + M[i][3] * N[3];
    double a1p4 = a[0] + a[3], a2m3 = a[1] - a[2], a1m4 = a[0] - a[3], a2p3 = a[1] + a[2];
    double s1 = sqrt(a1p4 * a1p4 + a2m3 * a2m3), s2 = sqrt(a1m4 * a1m4 + a2p3 * a2p3);
    double Ap = s1 + s2, Ac = s1 - s2;
    double psi = atan2(Ap * a[3] - Ac * a[0], -(Ac * a[1] + Ap * a[2]));
    psi *= 0.5;
    double c = (double)sgn(sin(2.0 * psi));
    double r = sqrt(Ap * Ap - Ac * Ac);
    out[0] = 0.5 * (Ap + r);
    out[1] = acos(-Ac / (Ap + r));
    out[2] = psi;
    out[3] = atan2(c * (Ap * a[3] - Ac * a[0]), -c * (Ac * a[2] + Ap * a[1]));
}

int orc_get_Fstat_logL(const orc_orbit *orbit, const char *format, double T, double t0, int NFFT, int qmin0,
                       const double *X, const double *A, const double *E, const double *C00, double f0, double fdot,
                       double theta, double phi, double *logL_X, double *logL_AE, double *Fparams)
{
    long q = (long)(f0 * T);
    int M = orc_ucb_bandwidth(orbit->L, orbit->fstar, f0, fdot, cos(theta), 1.e-22, T, NFFT);
    double *h = calloc((size_t)4 * 5 * 2 * M, sizeof(double));
    const double psi[4] = {0.0, M_PI_4, 0.0, M_PI_4}, ph[4] = {0.0, M_PI, 3.0 * M_PI_2, M_PI_2};
    for (int fi = 0; fi < 4; fi++)
    {
        double p[9] = {f0 * T, cos(theta), phi, log(2.0), cos(M_PI_2), psi[fi], ph[fi], fdot * (T * T), 0.0};
        double *o = h + (size_t)fi * 10 * M;
        if (orc_ucb_waveform(orbit, format, T, t0, p, 8, o, o + 2 * M, o + 4 * M, o + 6 * M, o + 8 * M, M, 2))
        {
            free(h);
            return -1;
        }
    }
#define FX(fi) (h + (size_t)(fi) * 10 * M)
#define FA(fi) (h + (size_t)(fi) * 10 * M + 6 * M)
#define FE(fi) (h + (size_t)(fi) * 10 * M + 8 * M)
    double nx[4] = {0, 0, 0, 0}, nae[4] = {0, 0, 0, 0};
    double MX[4][4], MA[4][4];
    memset(MX, 0, sizeof(MX));
    memset(MA, 0, sizeof(MA));
    long qo = q - qmin0;
    for (int i = 0; i < M; i++)
    {
        long k = qo + i - M / 2;
        if (!(k > 0 && k < NFFT)) continue;
        double c = C00[k];
        for (int a = 0; a < 4; a++)
        {
            nx[a] += (X[2 * k] * FX(a)[2 * i] + X[2 * k + 1] * FX(a)[2 * i + 1]) / c;
            nae[a] += (A[2 * k] * FA(a)[2 * i] + A[2 * k + 1] * FA(a)[2 * i + 1] + E[2 * k] * FE(a)[2 * i] +
                       E[2 * k + 1] * FE(a)[2 * i + 1]) /
                      c;
        }
        for (int a = 0; a < 4; a++)
            for (int b = a; b < 4; b++)
            {
                MX[a][b] += 4.0 * (FX(a)[2 * i] * FX(b)[2 * i] + FX(a)[2 * i + 1] * FX(b)[2 * i + 1]) / c;
                MA[a][b] += 4.0 * (FA(a)[2 * i] * FA(b)[2 * i] + FA(a)[2 * i + 1] * FA(b)[2 * i + 1]) / c +
                            4.0 * (FE(a)[2 * i] * FE(b)[2 * i] + FE(a)[2 * i + 1] * FE(b)[2 * i + 1]) / c;
            }
    }
#undef FX
#undef FA
#undef FE
    free(h);
    for (int a = 0; a < 4; a++)
    {
        nx[a] *= 4.0;
        nae[a] *= 4.0;
    }
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++)
        {
            MX[b][a] = MX[a][b];
            MA[b][a] = MA[a][b];
        }
    invert4(MX);
    invert4(MA);
    *logL_X = fstat_value(MX, nx);
    *logL_AE = fstat_value(MA, nae);
    fstat_params(MA, nae, Fparams);
    return M;
}

double orc_ucb_fdot(double Mc, double f0)
{
    double M = Mc * 4.92569043916e-6;
    double Q = 19.2;
    return Q * pow(pow(M_PI, 8) * pow(M, 5) * pow(f0, 11), 1. / 3.);
}
