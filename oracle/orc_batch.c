/*
 * orc_batch.c -- the reference's per-step likelihood work, timed as the CPU
 * baseline (test infrastructure only).
 *
 * One evaluation restates what one MCMC proposal costs the reference:
 *   generate_ucb_model(source_id)  ucb/src/glass_ucb_model.c:55-103
 *     - zero the 5 window model channels (X,Y,Z,A,E) and the regenerated
 *       source's 5 template channels
 *     - ucb_alignment + ucb_waveform for the proposed source
 *     - add_ucb_model of every live source (stored templates for the others)
 *   gaussian_log_likelihood        utils/src/glass_model.c:726-774
 *     - residual of all 5 channels, 6 fourier_nwip terms over the window
 * Proposals are independent and spread over OpenMP threads with dynamic
 * scheduling, as the reference spreads chains (apps/src/ucb_mcmc.c:230).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "glass_oracle.h"

double orc_eval_logl(const orc_orbit *orbit, double T, double t0, const orc_chain *ch, int sid,
                     const double *params)
{
    int W = ch->W;
    size_t N2 = 2 * (size_t)W;
    double *m = (double *)malloc(sizeof(double) * 10 * N2); /* model X..E + residual X..E */
    double *h = (double *)malloc(sizeof(double) * 5 * N2);  /* regenerated template X..E  */
    memset(m, 0, sizeof(double) * 5 * N2);
    memset(h, 0, sizeof(double) * 5 * N2);
    double *mX = m, *mY = m + N2, *mZ = m + 2 * N2;
    for (int s = 0; s < ch->nsrc; s++)
    {
        if (s == sid)
        {
            int BW, qmin, imin;
            orc_ucb_alignment(orbit, T, W, ch->qmin, params, &BW, &qmin, &imin);
            orc_ucb_waveform(orbit, "sangria", T, t0, params, 8, h, h + N2, h + 2 * N2, h + 3 * N2,
                             h + 4 * N2, BW, 3);
            orc_add_ucb_model(mX, mY, mZ, h, h + N2, h + 2 * N2, BW, imin, W);
        }
        else
        {
            const double *t = ch->tmpl + (size_t)s * 3 * 2 * ch->bwmax;
            size_t st = 2 * (size_t)ch->bwmax;
            orc_add_ucb_model(mX, mY, mZ, t, t + st, t + 2 * st, ch->src_bw[s], ch->src_imin[s], W);
        }
    }
    const double *dX = ch->data, *dY = ch->data + N2, *dZ = ch->data + 2 * N2;
    double *r = m + 5 * N2;
    /* residual of all five channels as the reference forms it; the data's
     * A,E channels are not used by the 3-channel likelihood and are formed
     * from zero here */
    for (size_t i = 0; i < N2; i++)
    {
        r[i] = dX[i] - mX[i];
        r[N2 + i] = dY[i] - mY[i];
        r[2 * N2 + i] = dZ[i] - mZ[i];
        r[3 * N2 + i] = 0.0 - m[3 * N2 + i];
        r[4 * N2 + i] = 0.0 - m[4 * N2 + i];
    }
    const double *ic = ch->invC;
    size_t P = (size_t)W;
    double chi2 = 0.0;
    chi2 += orc_fourier_nwip(r, r, ic + 0 * P, W);
    chi2 += orc_fourier_nwip(r + N2, r + N2, ic + 4 * P, W);
    chi2 += orc_fourier_nwip(r + 2 * N2, r + 2 * N2, ic + 8 * P, W);
    chi2 += 2.0 * orc_fourier_nwip(r, r + N2, ic + 1 * P, W);
    chi2 += 2.0 * orc_fourier_nwip(r, r + 2 * N2, ic + 2 * P, W);
    chi2 += 2.0 * orc_fourier_nwip(r + N2, r + 2 * N2, ic + 5 * P, W);
    free(m);
    free(h);
    return -0.5 * chi2;
}

double orc_eval_batch(const orc_orbit *orbit, double T, double t0, const orc_chain *chains, int n,
                      const int *chain_of, const int *sid_of, const double *params, double *out,
                      int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double t_start = omp_get_wtime();
#pragma omp parallel for schedule(dynamic)
    for (int e = 0; e < n; e++)
        out[e] = orc_eval_logl(orbit, T, t0, &chains[chain_of[e]], sid_of[e], params + (size_t)8 * e);
    return omp_get_wtime() - t_start;
}
