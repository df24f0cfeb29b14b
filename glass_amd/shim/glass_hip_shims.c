/*
 * glass_hip_shims.c -- the reference-side shims: drop into the reference
 * tree as utils/src/glass_hip_shims.c and compile it with -DGLASS_HIP (the
 * build and rename steps are in the integration notes next to the headers).
 *
 * Each shim has the name and signature of the reference function it
 * replaces, unpacks the reference structs into the plain pointers of the C
 * ABI (include/glass_hip*.h) and turns a non-zero status into the
 * reference's exit(1), as the replaced functions exit on unsupported input
 * (ucb/src/glass_ucb_waveform.c:971-975).  The originals stay available as
 * *_cpu for the cases the library does not take (fewer than 3 channels).
 *
 * GLASS_HIP_SHIM_SUBSET: compile against a caller-provided declaration of
 * the struct fields used here instead of the reference headers (the
 * library's own compile check does this).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef GLASS_HIP_SHIM_SUBSET
#include "glass_utils.h"
#include "glass_ucb.h"
#endif

#include "glass_hip.h"
#include "glass_hip_wdm.h"

static void hip_fail(const char *what)
{
    fprintf(stderr, "ERROR: %s: %s\n", what, glass_hip_last_error());
    exit(1);
}

static glass_hip_orbit hip_orbit(const struct Orbit *o)
{
    glass_hip_orbit g = {o->L, o->fstar, o->R};
    return g;
}

static glass_hip_wdm hip_wdm(const struct Wavelets *w)
{
    glass_hip_wdm g = {w->NF, w->NT, w->dt, w->df, w->A, w->kmin, w->kmax, w->imin, w->imax};
    return g;
}

/* the orbit's knots as 3 x Norb blocks (row i = spacecraft i): x | y | z */
static double *hip_knots(const struct Orbit *o)
{
    int N = o->Norb;
    double *x = malloc(9 * (size_t)N * sizeof(double));
    for (int i = 0; i < 3; i++)
    {
        memcpy(x + i * (size_t)N, o->x[i], N * sizeof(double));
        memcpy(x + (3 + i) * (size_t)N, o->y[i], N * sizeof(double));
        memcpy(x + (6 + i) * (size_t)N, o->z[i], N * sizeof(double));
    }
    return x;
}

/* 3x3 planes of a struct Noise as one contiguous row-major block */
static double *hip_invC9(struct Noise *noise, int N)
{
    double *c = malloc(9 * (size_t)N * sizeof(double));
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            memcpy(c + (3 * a + b) * (size_t)N, noise->invC[a][b], N * sizeof(double));
    return c;
}

/* ---------------------------------------------------- Fourier domain --- */

void ucb_waveform(struct Orbit *orbit, char *format, double T, double t0, double *params, int NParams,
                  double *X, double *Y, double *Z, double *A, double *E, int BW, int NI)
{
    int rc;
    if (orbit->orbit_function == &analytic_orbits)
    {
        glass_hip_orbit o = hip_orbit(orbit);
        rc = glass_hip_ucb_waveform(&o, format, T, t0, params, NParams, X, Y, Z, A, E, BW, NI);
    }
    else
    {
        /* numeric ephemeris (interpolate_orbits): the knots go to the device */
        double *x = hip_knots(orbit);
        glass_hip_numeric_orbit o = {orbit->Norb, orbit->t, x, x + 3 * (size_t)orbit->Norb,
                                     x + 6 * (size_t)orbit->Norb, orbit->L, orbit->fstar};
        rc = glass_hip_ucb_waveform_numeric(&o, format, T, t0, params, NParams, X, Y, Z, A, E, BW, NI);
        free(x);
    }
    if (rc) hip_fail("ucb_waveform");
}

double fourier_nwip(const double *a, const double *b, const double *invC, int N)
{
    return glass_hip_fourier_nwip(a, b, invC, N);
}

double snr(struct Source *source, struct Noise *noise)
{
    if (source->tdi->Nchannel != 3) return snr_cpu(source, noise);
    double *ic = hip_invC9(noise, noise->N);
    double s = glass_hip_snr(source->tdi->X, source->tdi->Y, source->tdi->Z, ic, noise->N, source->tdi->N / 2);
    free(ic);
    return s;
}

double gaussian_log_likelihood(struct Data *data, struct Model *model)
{
    struct TDI *d = data->tdi, *m = model->tdi;
    double *ic = hip_invC9(model->noise, data->NFFT);
    double L = glass_hip_gaussian_log_likelihood(d->X, d->Y, d->Z, d->A, d->E, m->X, m->Y, m->Z, m->A, m->E, ic,
                                                 data->NFFT, data->Nchannel);
    free(ic);
    return L;
}

void invert_noise_covariance_matrix(struct Noise *noise)
{
    if (noise->Nchannel != 3)
    {
        invert_noise_covariance_matrix_cpu(noise);
        return;
    }
    int N = noise->N;
    double *C = malloc(9 * (size_t)N * sizeof(double)), *iC = malloc(9 * (size_t)N * sizeof(double));
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) memcpy(C + (3 * a + b) * (size_t)N, noise->C[a][b], N * sizeof(double));
    if (glass_hip_invert_noise_covariance(C, iC, noise->logdetC, N)) hip_fail("invert_noise_covariance");
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) memcpy(noise->invC[a][b], iC + (3 * a + b) * (size_t)N, N * sizeof(double));
    free(C);
    free(iC);
}

void ucb_fisher(struct Orbit *orbit, struct Data *data, struct Source *source, struct Noise *noise)
{
    if (orbit->orbit_function != &analytic_orbits)
    {
        ucb_fisher_cpu(orbit, data, source, noise);
        return;
    }
    glass_hip_orbit o = hip_orbit(orbit);
    double F[64], *ic = hip_invC9(noise, data->NFFT);
    if (glass_hip_ucb_fisher(&o, data->format, data->T, data->t0, data->NFFT, data->qmin, source->params, ic, F))
        hip_fail("ucb_fisher");
    free(ic);
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) source->fisher_matrix[i][j] = F[8 * i + j];
    /* the eigensystem stays with the reference (LAPACKE, glass_math.c:475-499) */
    matrix_eigenstuff(source->fisher_matrix, source->fisher_evectr, source->fisher_evalue, 8);
}

/* ---------------------------------------------------- wavelet domain --- */

void ucb_waveform_wavelet(struct Orbit *orbit, struct Wavelets *wdm, double Tobs, double t0, double *params,
                          int *wavelet_list, int *Nwavelet, double *X, double *Y, double *Z)
{
    int N = orbit->Norb;
    double *x = hip_knots(orbit);
    glass_hip_wdm w = hip_wdm(wdm);
    int rc = glass_hip_ucb_waveform_wavelet(&w, N, orbit->t, x, x + 3 * (size_t)N, x + 6 * (size_t)N, Tobs, t0,
                                            params, wavelet_list, Nwavelet, X, Y, Z);
    free(x);
    if (rc) hip_fail("ucb_waveform_wavelet");
}

void ucb_fisher_wavelet(struct Orbit *orbit, struct Data *data, struct Source *source, struct Noise *noise)
{
    int N = orbit->Norb;
    double *x = hip_knots(orbit), F[64];
    glass_hip_wdm w = hip_wdm(data->wdm);
    double *ic = hip_invC9(noise, data->N);
    int rc = glass_hip_ucb_fisher_wavelet(&w, N, orbit->t, x, x + 3 * (size_t)N, x + 6 * (size_t)N, data->T,
                                          data->t0, source->params, ic, data->N, F);
    free(x);
    free(ic);
    if (rc) hip_fail("ucb_fisher_wavelet");
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) source->fisher_matrix[i][j] = F[8 * i + j];
    /* eigensystem and the one regularisation retry stay on the host
       (glass_ucb_waveform.c:620-686) */
    matrix_eigenstuff(source->fisher_matrix, source->fisher_evectr, source->fisher_evalue, 8);
    int bad = 0;
    for (int i = 0; i < 8; i++) bad |= !(source->fisher_evalue[i] > 0);
    if (bad)
    {
        for (int i = 0; i < 8; i++) source->fisher_matrix[i][i] += 1.0e-1;
        matrix_eigenstuff(source->fisher_matrix, source->fisher_evectr, source->fisher_evalue, 8);
        for (int i = 0; i < 8; i++)
            if (!(source->fisher_evalue[i] > 0))
            {
                fprintf(stderr, "ERROR: bad eigenvalue %d\n", i);
                exit(1);
            }
    }
}

double wavelet_nwip(const double *a, const double *b, const double *invC, const int *list, int N)
{
    int nmax = 0;
    for (int n = 0; n < N; n++)
        if (list[n] + 1 > nmax) nmax = list[n] + 1;
    return glass_hip_wavelet_nwip(a, b, invC, list, N, nmax);
}

double snr_wavelet(struct Source *source, struct Noise *noise)
{
    double *ic = hip_invC9(noise, noise->N);
    double s = glass_hip_snr_wavelet(source->tdi->X, source->tdi->Y, source->tdi->Z, ic, noise->N, source->list,
                                     source->Nlist);
    free(ic);
    return s;
}

double gaussian_log_likelhood_wavelet(struct Data *data, struct Model *model)
{
    glass_hip_wdm w = hip_wdm(data->wdm);
    double *ic = hip_invC9(model->noise, data->N);
    double L = glass_hip_gaussian_log_likelihood_wavelet(&w, data->tdi->X, data->tdi->Y, data->tdi->Z,
                                                         model->tdi->X, model->tdi->Y, model->tdi->Z, model->list,
                                                         model->Nlist, ic, data->N);
    free(ic);
    return L;
}

void wavelet_transform_timefreq(struct Wavelets *wdm, double *timedata)
{
    glass_hip_wdm w = hip_wdm(wdm);
    if (glass_hip_wavelet_transform_timefreq(&w, timedata)) hip_fail("wavelet_transform_timefreq");
}

void wavelet_transform_freq(struct Wavelets *wdm, double *freqdata, double *wdmdata)
{
    glass_hip_wdm w = hip_wdm(wdm);
    if (glass_hip_wavelet_transform_freq(&w, freqdata, wdmdata)) hip_fail("wavelet_transform_freq");
}

void wavelet_transform_inverse_freq(struct Wavelets *wdm, double *wdmdata, double *freqdata)
{
    glass_hip_wdm w = hip_wdm(wdm);
    if (glass_hip_wavelet_transform_inverse_freq(&w, wdmdata, freqdata)) hip_fail("wavelet_transform_inverse_freq");
}

void stationary_dft_psd_to_wdm_psd(struct Wavelets *wdm, double *dft_psd, double *wdm_psd)
{
    glass_hip_wdm w = hip_wdm(wdm);
    if (glass_hip_stationary_dft_psd_to_wdm_psd(&w, dft_psd, wdm_psd)) hip_fail("stationary_dft_psd_to_wdm_psd");
}

void stationary_dft_psd_to_wdm_layer_var(struct Wavelets *wdm, double *dft_psd, double *layer_var)
{
    glass_hip_wdm w = hip_wdm(wdm);
    if (glass_hip_stationary_dft_psd_to_wdm_layer_var(&w, dft_psd, layer_var))
        hip_fail("stationary_dft_psd_to_wdm_layer_var");
}
