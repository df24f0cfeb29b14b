/*
 * orc_noise.c -- CPU restatement of the instrument-noise model, the noise
 * likelihoods and the instrument-noise MCMC step.  Test infrastructure
 * only: the checker for the HIP noise kernels and the CPU baseline of the
 * noise leg.
 *
 * Reference functions restated (paths relative to the reference root):
 *   generate_instrument_noise_model           noise/src/glass_noise_model.c:389-545
 *   generate_instrument_noise_model_wavelet   noise/src/glass_noise_model.c:561-610
 *   generate_full_covariance_matrix           noise/src/glass_noise_model.c:894-907
 *   noise_log_likelihood                      noise/src/glass_noise_model.c:1104-1134
 *   my_noise_log_likelihood_wavelet           noise/src/glass_noise_model.c:1182-1206
 *   my_noise_log_likelihood                   noise/src/glass_noise_model.c:1275-1306
 *   noise_instrument_model_mcmc               noise/src/glass_noise_sampler.c:395-592
 *   noise_ptmcmc                              noise/src/glass_noise_sampler.c:70-117
 *   rand_r_U_0_1 / rand_r_N_0_1               utils/src/glass_utils.c:19-35
 * The random numbers come from the C library's rand_r, as in the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "glass_oracle.h"

#define O_PI2 6.283185307179586476925286766559
#define O_CLIGHT 299792458.

static double noise_transfer(double x)
{
    double s = sin(x);
    return s * s;
}

/* XYZnoise_FF / AEnoise_FF, utils/src/glass_lisa.c:1121-1145 */
static double xyz_noise_ff(double fstar, double f, double Spm, double Sop)
{
    double x = f / fstar, cx = cos(x);
    return 16. * noise_transfer(x) * (2. * (1.0 + cx * cx) * Spm + Sop);
}

static double ae_noise_ff(double fstar, double f, double Spm, double Sop)
{
    double x = f / fstar, cx = cos(x), c2x = cos(2. * x);
    return 8. * noise_transfer(x) * (2. * Spm * (3. + 2. * cx + c2x) + Sop * (2. + cx));
}

/* Parameter order of sacc[6] / soms[6]: links 12, 21, 23, 32, 13, 31
 * (map_array_to_noise_params, noise/src/glass_noise_model.c:24-44). */
void orc_instrument_noise_model(double fstar, const double *f, int N, int nch, const double *sacc,
                                const double *soms, double *C)
{
    const double acc_units = 2.81837551648e-19;
    const double oms_units = 4.39256635604e-16;
    const double acc_fmax = 0.01;
    const double a12 = sacc[0], a21 = sacc[1], a23 = sacc[2], a32 = sacc[3], a13 = sacc[4], a31 = sacc[5];
    const double o12 = soms[0], o21 = soms[1], o23 = soms[2], o32 = soms[3], o13 = soms[4], o31 = soms[5];
#define CP(i, j) C[((size_t)(i) * nch + (j)) * N + n]
    for (int n = 0; n < N; n++)
    {
        double fn = f[n];
        if (fn == 0.0)
        {
            /* the reference zeroes these five entries and skips the rest */
            if (nch > 0) CP(0, 0) = 0.0;
            if (nch > 1) CP(0, 1) = 0.0;
            if (nch > 2)
            {
                CP(0, 2) = 0.0;
                CP(1, 2) = 0.0;
                CP(2, 2) = 0.0;
            }
            continue;
        }
        double x = fn / fstar, f2 = fn * fn;
        double cosx = cos(x);
        double tdi = noise_transfer(x);
        double acc = (fn < acc_fmax)
                         ? acc_units / f2 * (1.0 + pow(0.4e-3 / fn, 2)) * (1.0 + pow(fn / 8.0e-3, 4))
                         : 0.0;
        double oms = oms_units * f2 * (1.0 + pow(2.0e-3 / fn, 4));
        tdi *= -8.0;
        if (nch == 1)
        {
            CP(0, 0) = xyz_noise_ff(fstar, fn, a12 * acc, o12 * oms);
        }
        else if (nch == 2)
        {
            CP(0, 0) = ae_noise_ff(fstar, fn, a12 * acc, o12 * oms);
            CP(1, 1) = ae_noise_ff(fstar, fn, a12 * acc, o12 * oms);
            CP(0, 1) = 0.0;
        }
        else
        {
            double S12 = 0, S21 = 0, S23 = 0, S32 = 0, S13 = 0, S31 = 0;
            if (fn < acc_fmax)
            {
                S12 = a12 * acc;
                S21 = a21 * acc;
                S23 = a23 * acc;
                S32 = a32 * acc;
                S13 = a13 * acc;
                S31 = a31 * acc;
            }
            double O12 = o12 * oms, O21 = o21 * oms, O23 = o23 * oms;
            double O32 = o32 * oms, O13 = o13 * oms, O31 = o31 * oms;
            double cxx = 0., cyy = 0., czz = 0., cxy = 0., cxz = 0., cyz = 0.;
            cxx += (O12 + O21 + O13 + O31) * .25;
            cyy += (O23 + O32 + O21 + O12) * .25;
            czz += (O31 + O13 + O32 + O23) * .25;
            cxy += (O12 + O21) * .5;
            cxz += (O13 + O31) * .5;
            cyz += (O32 + O23) * .5;
            if (fn < acc_fmax)
            {
                cxx += ((S12 + S13) + ((S21 + S31)) * cosx * cosx);
                cyy += ((S23 + S21) + ((S32 + S12)) * cosx * cosx);
                czz += ((S31 + S32) + ((S13 + S23)) * cosx * cosx);
                cxy += 2. * (S12 + S21);
                cxz += 2. * (S13 + S31);
                cyz += 2. * (S32 + S23);
            }
            cxx *= -2. * tdi;
            cyy *= -2. * tdi;
            czz *= -2. * tdi;
            cxy *= tdi * cosx;
            cxz *= tdi * cosx;
            cyz *= tdi * cosx;
            CP(0, 0) = cxx;
            CP(1, 1) = cyy;
            CP(2, 2) = czz;
            CP(0, 1) = cxy;
            CP(0, 2) = cxz;
            CP(1, 2) = cyz;
            CP(1, 0) = cxy;
            CP(2, 0) = cxz;
            CP(2, 1) = cyz;
        }
        /* every entry is normalised, including C[1][0] of the two-channel
         * case, which the reference does not write */
        for (int i = 0; i < nch; i++)
            for (int j = 0; j < nch; j++) CP(i, j) /= 4.0;
    }
#undef CP
}

int orc_instrument_noise_model_wavelet(double fstar, const double *f, int N, double df, int ND,
                                       const double *sacc, const double *soms, double *C)
{
    int imin = (int)round(f[0] / df);
    int imax = (int)round(f[N - 1] / df) + 1;
    int Na = imax - imin;
    if (Na <= 0 || Na > N) return -1;
    double *fg = malloc(sizeof(double) * Na);
    double *G = calloc(9 * (size_t)Na, sizeof(double));
    for (int n = 0; n < Na; n++) fg[n] = (double)(imin + n) * df;
    orc_instrument_noise_model(fstar, fg, Na, 3, sacc, soms, G);
    double inv_ND = 1.0 / (double)ND;
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++)
            for (int k = 0; k < Na; k++)
            {
                double v = G[(a * 3 + b) * Na + k] * inv_ND;
                C[(size_t)(a * 3 + b) * N + k] = v;
                C[(size_t)(b * 3 + a) * N + k] = v;
            }
    free(fg);
    free(G);
    return Na;
}

double orc_my_noise_log_likelihood(const double *X, const double *Y, const double *Z, const double *invC,
                                   const double *logdetC, int N)
{
    double logL = 0.0;
    const double log2pi = log(2 * M_PI);
#define IC(a, b) invC[((a) * 3 + (b)) * (size_t)N + n]
    for (int n = 0; n < N; n++)
    {
        double rex = X[2 * n], imx = X[2 * n + 1];
        double rey = Y[2 * n], imy = Y[2 * n + 1];
        double rez = Z[2 * n], imz = Z[2 * n + 1];
        double xx = rex * rex + imx * imx;
        double yy = rey * rey + imy * imy;
        double zz = rez * rez + imz * imz;
        double xy = rex * rey + imx * imy;
        double xz = rex * rez + imx * imz;
        double yz = rey * rez + imy * imz;
        logL -= 0.5 * xx * IC(0, 0) * 2;
        logL -= 0.5 * yy * IC(1, 1) * 2;
        logL -= 0.5 * zz * IC(2, 2) * 2;
        logL -= xy * IC(0, 1) * 2;
        logL -= xz * IC(0, 2) * 2;
        logL -= yz * IC(1, 2) * 2;
        logL -= logdetC[n];
    }
#undef IC
    logL -= 3 * N * log2pi;
    return logL;
}

double orc_noise_log_likelihood(const double *X, const double *Y, const double *Z, const double *invC,
                                const double *logdetC, int N)
{
    double logL = 0.0;
    const double *iC = invC;
    logL += -0.5 * orc_fourier_nwip(X, X, iC + 0 * (size_t)N, N);
    logL += -0.5 * orc_fourier_nwip(Y, Y, iC + 4 * (size_t)N, N);
    logL += -0.5 * orc_fourier_nwip(Z, Z, iC + 8 * (size_t)N, N);
    logL += -orc_fourier_nwip(X, Y, iC + 1 * (size_t)N, N);
    logL += -orc_fourier_nwip(X, Z, iC + 2 * (size_t)N, N);
    logL += -orc_fourier_nwip(Y, Z, iC + 5 * (size_t)N, N);
    for (int n = 0; n < N; n++) logL -= logdetC[n];
    return logL;
}

double orc_my_noise_log_likelihood_wavelet(const double *X, const double *Y, const double *Z,
                                           const double *invC, const double *logdetC, int NT, int Nlayer,
                                           int Ntot, int ldc)
{
    const double log2pi = log(2 * M_PI);
    double logL = 0.0;
#define IC(a, b) invC[((a) * 3 + (b)) * (size_t)ldc + k]
    for (int j = 0; j < Nlayer; j++)
        for (int i = 0; i < NT; i++)
        {
            int k = i + j * NT;
            double wx = X[k], wy = Y[k], wz = Z[k];
            double s = 0.0;
            s -= 0.5 * 1.0 * IC(0, 0) * (wx * wx);
            s -= 0.5 * 1.0 * IC(1, 1) * (wy * wy);
            s -= 0.5 * 1.0 * IC(2, 2) * (wz * wz);
            s -= 1.0 * IC(0, 1) * (wx * wy);
            s -= 1.0 * IC(0, 2) * (wx * wz);
            s -= 1.0 * IC(1, 2) * (wy * wz);
            s -= 0.5 * 1.0 * logdetC[k];
            logL += s;
        }
#undef IC
    logL -= 0.5 * 3 * Ntot * log2pi;
    return logL;
}

/* ------------------------------------------------------------ sampler -- */

double orc_rand_U_0_1(unsigned int *seed) { return (double)rand_r(seed) / (double)RAND_MAX; }

double orc_rand_N_0_1(unsigned int *seed)
{
    double u, v, s;
    do
    {
        u = 2 * orc_rand_U_0_1(seed) - 1;
        v = 2 * orc_rand_U_0_1(seed) - 1;
        s = u * u + v * v;
    } while (s >= 1 || s == 0);
    return u * sqrt(-2 * log(s) / s);
}

/* logL of one parameter set: instrument model, + fixed components, inverse,
 * my_noise_log_likelihood (the evaluation inside the MCMC step) */
double orc_noise_logl_params(double fstar, const double *f, int N, const double *X, const double *Y,
                             const double *Z, const double *Cadd, const double *p12, double *work)
{
    double *C = work, *iC = work + 9 * (size_t)N, *ld = work + 18 * (size_t)N;
    orc_instrument_noise_model(fstar, f, N, 3, p12, p12 + 6, C);
    if (Cadd)
        for (size_t i = 0; i < 9 * (size_t)N; i++) C[i] += Cadd[i];
    orc_invert_noise_covariance(C, iC, ld, N);
    return orc_my_noise_log_likelihood(X, Y, Z, iC, ld, N);
}

static int link_draw(unsigned int *seed)
{
    int i = (int)(orc_rand_U_0_1(seed) * 6.0);
    return i > 5 ? 5 : i;
}

/* correlation_matrix of noise/src/glass_noise_sampler.c:436-461 */
static const double kCorr[6][6] = {
    {+1, -1, +1, -1, -1, +1}, {-1, +1, -1, +1, +1, -1}, {+1, -1, +1, -1, -1, +1},
    {-1, +1, -1, +1, +1, -1}, {-1, +1, -1, +1, +1, -1}, {+1, -1, +1, -1, -1, +1},
};

int orc_noise_mcmc_call(double fstar, const double *f, int N, const double *X, const double *Y,
                        const double *Z, const double *Cadd, double temperature, double *x, double *logLx,
                        unsigned int *seed, double *work)
{
    const double Sacc = 9.00e-30, Soms = 2.25e-22;
    const double Sacc_min = 9.00e-30 / 10, Sacc_max = 9.00e-30 * 10;
    const double Soms_min = 2.25e-22 / 10, Soms_max = 2.25e-22 * 10;
    double y[12], logLy = *logLx;
    double logH = 0.0, logPx = 0.0, logPy = 0.0;
    int nacc = 0;
    memcpy(y, x, sizeof(y));
    double *ya = y, *yo = y + 6;
    const double *xa = x, *xo = x + 6;
    for (int mc = 0; mc < 10; mc++)
    {
        double scale;
        if (orc_rand_U_0_1(seed) > 0.75)
            scale = 1;
        else if (orc_rand_U_0_1(seed) > 0.5)
            scale = 0.1;
        else if (orc_rand_U_0_1(seed) > 0.25)
            scale = 0.01;
        else
            scale = 0.001;
        /* rand_r can return RAND_MAX, so U = 1 exactly (probability 2^-31):
         * type 3 then matches no case and changes nothing, as in the
         * reference; a link index of 6 would index past the six links, so
         * link draws are clamped to 5 (the reference reads out of bounds) */
        int type = (int)(orc_rand_U_0_1(seed) * 3.);
        int i, j;
        switch (type)
        {
        case 0:
            i = link_draw(seed);
            ya[i] = xa[i] + scale * Sacc * orc_rand_N_0_1(seed);
            yo[i] = xo[i] + scale * Soms * orc_rand_N_0_1(seed);
            if (i % 2 == 0)
                yo[i + 1] = yo[i];
            else
                yo[i - 1] = yo[i];
            if (ya[i] < Sacc_min || ya[i] > Sacc_max) logPy = -INFINITY;
            if (yo[i] < Soms_min || yo[i] > Soms_max) logPy = -INFINITY;
            break;
        case 1:
        {
            i = link_draw(seed);
            do
                j = link_draw(seed);
            while (i == j);
            double aj = scale * Sacc * orc_rand_N_0_1(seed);
            double oj = scale * Soms * orc_rand_N_0_1(seed);
            if (abs(i - j) == 1)
            {
                ya[i] = xa[i] + aj;
                ya[j] = xa[j] - aj;
                yo[i] = xo[i] + oj;
            }
            else
            {
                ya[i] = xa[i] + aj;
                ya[j] = xa[j] + aj;
                yo[i] = xo[i] + oj;
                yo[j] = xo[j] - oj;
            }
            if (i % 2 == 0)
                yo[i + 1] = yo[i];
            else
                yo[i - 1] = yo[i];
            if (j % 2 == 0)
                yo[j + 1] = yo[j];
            else
                yo[j - 1] = yo[j];
            if (ya[i] < Sacc_min || ya[i] > Sacc_max) logPy = -INFINITY;
            if (ya[j] < Sacc_min || ya[j] > Sacc_max) logPy = -INFINITY;
            if (yo[i] < Soms_min || yo[i] > Soms_max) logPy = -INFINITY;
            if (yo[j] < Soms_min || yo[j] > Soms_max) logPy = -INFINITY;
            break;
        }
        case 2:
        {
            double jv[6];
            for (i = 0; i < 6; i++)
            {
                jv[i] = scale * Sacc * orc_rand_N_0_1(seed);
                ya[i] = xa[i];
            }
            for (i = 0; i < 6; i++)
            {
                for (j = 0; j < 6; j++) ya[i] += kCorr[i][j] * jv[j];
                if (ya[i] < Sacc_min || ya[i] > Sacc_max) logPy = -INFINITY;
            }
            break;
        }
        }
        if (logPy > -INFINITY)
        {
            logLy = orc_noise_logl_params(fstar, f, N, X, Y, Z, Cadd, y, work);
            logH += (logLy - *logLx) / temperature;
        }
        logH += logPy - logPx;
        double loga = log(orc_rand_U_0_1(seed));
        if (logH > loga)
        {
            memcpy(x, y, sizeof(y));
            *logLx = logLy;
            nacc++;
        }
    }
    return nacc;
}

void orc_noise_ptmcmc(int NC, int *index, const double *temperature, const double *logL, unsigned int *seeds,
                      int *acceptance)
{
    for (int b = NC - 1; b > 0; b--)
    {
        int a = b - 1;
        acceptance[a] = 0;
        int olda = index[a], oldb = index[b];
        double heat1 = temperature[a], heat2 = temperature[b];
        double logL1 = logL[olda], logL2 = logL[oldb];
        if (orc_rand_U_0_1(&seeds[a]) < 1.0)
        {
            double dlogL = logL2 - logL1;
            double H = (heat2 - heat1) / (heat2 * heat1);
            double alpha = exp(dlogL * H);
            double beta = orc_rand_U_0_1(&seeds[a]);
            if (alpha >= beta)
            {
                index[a] = oldb;
                index[b] = olda;
                acceptance[a] = 1;
            }
        }
    }
}
