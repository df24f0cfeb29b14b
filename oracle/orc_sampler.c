/*
 * orc_sampler.c -- CPU restatement of the UCB sampler's per-step logic and
 * of the Fisher eigen-decomposition (test infrastructure only).
 *
 * Reference functions restated (paths relative to the reference root):
 *   ucb_mcmc (propose / accept)     ucb/src/glass_ucb_sampler.c:75-173
 *   step loop, RJ coin              apps/src/ucb_mcmc.c:239-258
 *   draw_from_uniform_prior         utils/src/glass_sampler.c:96-107
 *   draw_from_fisher                ucb/src/glass_ucb_proposal.c:348-391
 *   fm_shift, psi_phi_jump          ucb/src/glass_ucb_proposal.c:577-613
 *   draw_from_fstatistic, jump_from_fstatistic,
 *   evaluate_fstatistic_proposal    ucb/src/glass_ucb_proposal.c:1526-1673
 *   check_range, evaluate_ucb_prior, evalaute_sky_location_prior,
 *   evaluate_uniform_priors         ucb/src/glass_ucb_prior.c:396-646
 *   ptmcmc, adapt_temperature_ladder  utils/src/glass_sampler.c:19-94
 * Random numbers: the C library's rand_r (orc_rand_U_0_1 / orc_rand_N_0_1).
 *
 * The eigen-decomposition replaces matrix_eigenstuff's LAPACKE_dgeev
 * (utils/src/glass_math.c:475-499) by cyclic Jacobi (Golub & Van Loan,
 * Matrix Computations, section 8.5) with eigenvalues ascending and each
 * eigenvector's largest-magnitude component positive; the same method and
 * operation order as the device kernel.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "glass_oracle.h"

#define S_PI 3.14159265358979323846
#define S_PI2 6.283185307179586
#define S_YEAR 3.15581497632e7
#define S_WRAPMAX 64

/* ------------------------------------------------------------ eigen ---- */

static int tri(int p, int q) { return p * 8 - p * (p - 1) / 2 + (q - p); }

int orc_jacobi_eigen8(const double *F, double *evec, double *eval)
{
    double a[36], v[64];
#define A(p, q) a[(p) <= (q) ? tri(p, q) : tri(q, p)]
#define V(p, q) v[(p) * 8 + (q)]
    for (int p = 0; p < 8; p++)
        for (int q = p; q < 8; q++) A(p, q) = F[p * 8 + q];
    for (int p = 0; p < 8; p++)
        for (int q = 0; q < 8; q++) V(p, q) = p == q ? 1.0 : 0.0;
    double nrm = 0.0;
    for (int p = 0; p < 8; p++)
        for (int q = 0; q < 8; q++) nrm += A(p, q) * A(p, q);
    const double tol = 1.2e-32 * nrm;
    int sw = 0;
    for (; sw < 40; sw++)
    {
        double off = 0.0;
        for (int p = 0; p < 7; p++)
            for (int q = p + 1; q < 8; q++) off += A(p, q) * A(p, q);
        if (!(off > tol)) break;
        for (int p = 0; p < 7; p++)
            for (int q = p + 1; q < 8; q++)
            {
                const double apq = A(p, q);
                if (apq == 0.0) continue;
                const double app = A(p, p), aqq = A(q, q);
                const double theta = (aqq - app) / (2.0 * apq);
                const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
                const double c = 1.0 / sqrt(1.0 + tt * tt), s = tt * c;
                A(p, p) = app - tt * apq;
                A(q, q) = aqq + tt * apq;
                A(p, q) = 0.0;
                for (int r = 0; r < 8; r++)
                {
                    if (r == p || r == q) continue;
                    const double arp = A(r, p), arq = A(r, q);
                    A(r, p) = c * arp - s * arq;
                    A(r, q) = s * arp + c * arq;
                }
                for (int r = 0; r < 8; r++)
                {
                    const double vrp = V(r, p), vrq = V(r, q);
                    V(r, p) = c * vrp - s * vrq;
                    V(r, q) = s * vrp + c * vrq;
                }
            }
    }
    int ord[8];
    double lam[8];
    for (int j = 0; j < 8; j++)
    {
        ord[j] = j;
        lam[j] = A(j, j);
    }
    for (int i = 1; i < 8; i++)
    {
        int k = ord[i];
        double x = lam[i];
        int j = i - 1;
        while (j >= 0 && lam[j] > x)
        {
            lam[j + 1] = lam[j];
            ord[j + 1] = ord[j];
            j--;
        }
        lam[j + 1] = x;
        ord[j + 1] = k;
    }
    for (int j = 0; j < 8; j++)
    {
        const int c = ord[j];
        int imax = 0;
        double amax = -1.0;
        for (int i = 0; i < 8; i++)
            if (fabs(V(i, c)) > amax)
            {
                amax = fabs(V(i, c));
                imax = i;
            }
        const double sg = V(imax, c) < 0.0 ? -1.0 : 1.0;
        for (int i = 0; i < 8; i++) evec[i * 8 + j] = sg * V(i, c);
        eval[j] = lam[j];
    }
#undef A
#undef V
    return sw;
}

/* ------------------------------------------------------------ priors --- */

static double wrap_up_down(double x, double lo, double hi, double period)
{
    int k = 0;
    while (x < lo && k < S_WRAPMAX)
    {
        x += period;
        k++;
    }
    while (x > hi && k < S_WRAPMAX)
    {
        x -= period;
        k++;
    }
    if (k == S_WRAPMAX && (x < lo || x > hi))
    {
        x = lo + fmod(x - lo, period);
        if (x < lo) x += period;
    }
    return x;
}

static int check_range(double *p, const double *pr)
{
    for (int n = 0; n < 8; n++)
        if (p[n] != p[n]) return 1;
    if (p[0] < pr[0] || p[0] > pr[1]) return 1;
    if (p[1] < pr[2] || p[1] > pr[3]) return 1;
    if (p[2] < pr[4] || p[2] >= pr[5])
    {
        p[2] = atan2(sin(p[2]), cos(p[2]));
        if (p[2] < 0.0) p[2] += S_PI2;
    }
    if (p[3] < pr[6] || p[3] > pr[7]) return 1;
    if (p[4] < pr[8] || p[4] > pr[9]) return 1;
    p[5] = wrap_up_down(p[5], pr[10], pr[11], S_PI);
    p[6] = wrap_up_down(p[6], pr[12], pr[13], S_PI2);
    if (p[7] < pr[14] || p[7] > pr[15]) return 1;
    return 0;
}

static double sky_prior(double *p, const double *pr, const double *lpv)
{
    double logP = 0.0;
    if (p[1] < pr[2] || p[1] > pr[3])
        return -INFINITY;
    else
        logP -= lpv[1];
    if (pr[4] > 0.0 || pr[5] < S_PI2)
    {
        if (p[2] < pr[4] || p[2] > pr[5]) return -INFINITY;
    }
    else if (p[2] < pr[4] || p[2] >= pr[5])
    {
        p[2] = atan2(sin(p[2]), cos(p[2]));
        if (p[2] < 0.0) p[2] += S_PI2;
    }
    logP -= lpv[2];
    return logP;
}

static double uniform_priors(double *p, const double *pr, const double *lpv)
{
    if (check_range(p, pr)) return -INFINITY;
    double logP = 0.0;
    logP -= lpv[0];
    logP -= lpv[3];
    logP -= lpv[4];
    logP -= lpv[5];
    logP -= lpv[6];
    logP -= lpv[7];
    return logP;
}

double orc_ucb_prior(double *p, const double *pr, const double *lpv)
{
    if (check_range(p, pr)) return -INFINITY;
    double logP = 0.0;
    logP += sky_prior(p, pr, lpv);
    logP += uniform_priors(p, pr, lpv);
    return logP;
}

void orc_ucb_log_prior_volume(const double *pr, double *lpv)
{
    for (int k = 0; k < 8; k++) lpv[k] = log(pr[2 * k + 1] - pr[2 * k]);
}

/* ---------------------------------------------------------- one step --- */

/* evaluate_fstatistic_proposal (ucb/src/glass_ucb_proposal.c:1640-1673) */
static double fstat_density(double *p, const double *pr, const double *lpv, const orc_fstat_prop *f, double T)
{
    const double d_theta = 2. / (double)f->nt, d_phi = S_PI2 / (double)f->nph;
    double logP = 0.0;
    logP += sky_prior(p, pr, lpv);
    logP += uniform_priors(p, pr, lpv);
    int i = (int)floor((p[0] - f->qmin) / (f->d_f * T));
    int j = (int)floor((p[1] - -1) / d_theta);
    int k = (int)floor((p[2]) / d_phi);
    if (i < 0 || i >= f->nf) return -INFINITY;
    if (j < 0 || j >= f->nt) return -INFINITY;
    if (k < 0 || k >= f->nph) return -INFINITY;
    logP += log(f->ten[((size_t)i * f->nt + j) * f->nph + k]);
    return logP;
}

/* rejection sampling of draw_from_fstatistic / jump_from_fstatistic
 * (:1526-1638); the jump's range check (failed only by U = 1) ends the
 * proposal with y[0..2] unchanged, the draw's missing check is a clamp */
static void fstat_sample(unsigned int *seed, double *y, const orc_fstat_prop *f, double T, int jump)
{
    const double d_theta = 2. / (double)f->nt, d_phi = S_PI2 / (double)f->nph;
    double i, j, k, q, costheta, phi, p, alpha;
    do
    {
        i = orc_rand_U_0_1(seed) * f->nf;
        j = orc_rand_U_0_1(seed) * f->nt;
        k = orc_rand_U_0_1(seed) * f->nph;
        q = f->qmin + i * f->d_f * T;
        costheta = -1. + j * d_theta;
        phi = k * d_phi;
        int ii = (int)i, jj = (int)j, kk = (int)k;
        if (jump && (ii < 0 || ii > f->nf - 1 || jj < 0 || jj > f->nt - 1 || kk < 0 || kk > f->nph - 1)) return;
        if (ii > f->nf - 1) ii = f->nf - 1;
        if (jj > f->nt - 1) jj = f->nt - 1;
        if (kk > f->nph - 1) kk = f->nph - 1;
        p = f->ten[((size_t)ii * f->nt + jj) * f->nph + kk];
        alpha = orc_rand_U_0_1(seed) * f->maxp;
    } while (p < alpha);
    y[0] = q;
    y[1] = costheta;
    y[2] = phi;
}

void orc_ucb_propose(unsigned int *seed, int nlive, double *sparams, const double *evec, const double *eval,
                     const double *pr, const double *lpv, const double *w, double T, int fix_sky, int fix_freq,
                     int fix_fdot, const orc_fstat_prop *fs, int *n_out, int *np_out, double *y, double *lpx_out,
                     double *lpy_out, double *lq_out)
{
    (void)orc_rand_U_0_1(seed); /* the step loop's RJ coin, flags->rj off */
    int n = (int)(orc_rand_U_0_1(seed) * (double)nlive);
    if (n >= nlive) n = nlive - 1;
    int np;
    double draw;
    do
    {
        np = (int)floor(9.0 * orc_rand_U_0_1(seed));
        if (np > 8) np = 8;
        draw = orc_rand_U_0_1(seed);
    } while (w[np] <= draw);
    double *xs = sparams + (size_t)n * 8;
    double x[8];
    for (int k = 0; k < 8; k++) x[k] = y[k] = xs[k];
    if (np == 0)
    {
        for (int k = 0; k < 8; k++) y[k] = pr[2 * k] + orc_rand_U_0_1(seed) * (pr[2 * k + 1] - pr[2 * k]);
    }
    else if (np == 1 || np == 2)
    {
        if (np == 1)
            for (int k = 0; k < 8; k++) y[k] = pr[2 * k] + orc_rand_U_0_1(seed) * (pr[2 * k + 1] - pr[2 * k]);
        else
            (void)(orc_rand_U_0_1(seed) < -0.5);
        fstat_sample(seed, y, fs, T, np == 2);
    }
    else if (np == 4)
    {
        const double *ev = evec + (size_t)n * 64, *la = eval + (size_t)n * 8;
        double amp[8], jump[8];
        for (int i = 0; i < 8; i++)
        {
            amp[i] = orc_rand_N_0_1(seed) / sqrt(la[i]);
            jump[i] = 0.0;
        }
        int i = (int)(orc_rand_U_0_1(seed) * 8.0);
        if (i > 7) i = 7;
        for (int j = 0; j < 8; j++) jump[j] += amp[i] * ev[j * 8 + i];
        for (int k = 0; k < 8; k++)
            if (jump[k] != jump[k]) jump[k] = 0.01 * x[k];
        for (int k = 0; k < 8; k++) y[k] = x[k] + jump[k];
        if (y[1] >= 1.) y[1] = x[1] - jump[1];
        if (y[1] <= -1.) y[1] = x[1] - jump[1];
        if (y[4] >= 1.) y[4] = x[4] - jump[4];
        if (y[4] <= -1.) y[4] = x[4] - jump[4];
    }
    else if (np == 5)
    {
        const double fm = T / S_YEAR;
        int scale;
        do
            scale = (int)floor(-3. + 7. * orc_rand_U_0_1(seed));
        while (scale == 0);
        y[0] += scale * fm;
        y[2] = y[2] + -0.2 + orc_rand_U_0_1(seed) * 0.4;
    }
    else if (np == 6)
    {
        const double scale = orc_rand_U_0_1(seed) * S_PI2;
        double d_phi = scale, d_psi = 0.5 * scale;
        if (orc_rand_U_0_1(seed) < 0.5) d_phi *= -1.;
        if (orc_rand_U_0_1(seed) < 0.5) d_psi *= -1.;
        y[5] += d_psi;
        y[6] += d_phi;
    }
    if (fix_sky)
    {
        y[1] = x[1];
        y[2] = x[2];
    }
    if (fix_freq) y[0] = x[0];
    if (fix_fdot) y[7] = x[7];
    double lqyx = 0.0, lqxy = 0.0;
    if (np == 0)
    {
        for (int k = 0; k < 8; k++) lqyx -= lpv[k];
        for (int k = 0; k < 8; k++) lqxy -= lpv[k];
    }
    else if (np == 1 || np == 2)
    {
        lqyx = fstat_density(y, pr, lpv, fs, T);
        lqxy = fstat_density(x, pr, lpv, fs, T);
    }
    *lq_out = lqxy - lqyx;
    *lpx_out = orc_ucb_prior(x, pr, lpv);
    *lpy_out = orc_ucb_prior(y, pr, lpv);
    for (int k = 0; k < 8; k++) xs[k] = x[k];
    *n_out = n;
    *np_out = np;
}

int orc_ucb_accept(unsigned int *seed, double logLy, double logLx, double snr, double snr_min, double temperature,
                   double lpx, double lpy, double lq)
{
    if (snr < snr_min) lpy = -INFINITY;
    double logH = 0.0;
    logH += (logLy - logLx) / temperature;
    logH += lpy - lpx;
    logH += lq;
    const double loga = log(orc_rand_U_0_1(seed));
    return isfinite(logH) && logH > loga;
}

void orc_ucb_ptmcmc(int NC, int *index, const double *temperature, const double *logL, const double *lognorm,
                    unsigned int *seeds, int *acceptance)
{
    for (int b = NC - 1; b > 0; b--)
    {
        int a = b - 1;
        acceptance[a] = 0;
        int olda = index[a], oldb = index[b];
        double heat1 = temperature[a], heat2 = temperature[b];
        double logL1 = logL[olda] + lognorm[olda];
        double logL2 = logL[oldb] + lognorm[oldb];
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

void orc_adapt_temperature_ladder(int NC, double *temperature, const int *acceptance, int mcmc)
{
    double S[64], A[64][2];
    double nu = 10, t0 = 10000.;
    if (NC > 64) return;
    for (int ic = 1; ic < NC - 1; ic++)
    {
        S[ic] = log(temperature[ic] - temperature[ic - 1]);
        A[ic][0] = acceptance[ic - 1];
        A[ic][1] = acceptance[ic];
    }
    for (int ic = 1; ic < NC - 1; ic++)
    {
        S[ic] += (A[ic][0] - A[ic][1]) * (t0 / ((double)mcmc + t0)) / nu;
        temperature[ic] = temperature[ic - 1] + exp(S[ic]);
        if (temperature[ic] / temperature[ic - 1] < 1.1) temperature[ic] = temperature[ic - 1] * 1.1;
    }
}
