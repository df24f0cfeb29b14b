/*
 * glass_hip_fstat.h -- the F-statistic of the UCB F-statistic proposal on
 * the device (libglass_hip, gfx950).
 *
 * Reference (paths relative to the reference root):
 *   get_Fstat_logL                  ucb/src/glass_ucb_fstatistic.c:646-703
 *   setup_fstatistic_proposal,
 *   build_fstatistic_proposal       ucb/src/glass_ucb_proposal.c:916-1180
 * Data layout: X, A, E the data's DFT (data->dft) as interleaved complex
 * double[2 NFFT] from bin qmin0 = (int)(data->fmin * data->T); C00[NFFT]
 * the noise's C[0][0], which the reference divides every channel by.
 * Return codes are GLASS_HIP_* (0 = ok); glass_hip_last_error() has the
 * message.  The 4x4 Gram inverses use Gauss-Jordan elimination with
 * partial pivoting in place of LAPACKE dgetrf/dgetri.
 */
#ifndef GLASS_HIP_FSTAT_H
#define GLASS_HIP_FSTAT_H

#include "glass_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* get_Fstat_logL(struct Orbit*, struct Data*, f0, fdot, theta, phi,
 * &logL_X, &logL_AE, Fparams): Fparams = A, iota, psi, phase of the A,E
 * maximum-likelihood fit.  Host arrays; uploads the data on every call. */
int glass_hip_get_Fstat_logL(const glass_hip_orbit *orbit, const char *format, double T, double t0, int NFFT,
                             int qmin0, const double *X, const double *A, const double *E, const double *C00,
                             double f0, double fdot, double theta, double phi, double *logL_X, double *logL_AE,
                             double *Fparams);

/* Many cells on one data segment: the data stays resident; max_cells the
 * most cells per glass_hip_fstat_eval. */
typedef struct glass_hip_fstat glass_hip_fstat;
int glass_hip_fstat_create(glass_hip_fstat **fstat, const glass_hip_orbit *orbit, const char *format, double T,
                           double t0, int NFFT, int qmin0, const double *X, const double *A, const double *E,
                           const double *C00, int max_cells, void *stream);
int glass_hip_fstat_destroy(glass_hip_fstat *fstat);
/* n <= max_cells cells, device arrays: cells[n][4] = f0, fdot, theta, phi;
 * logL_X[n], logL_AE[n], Fparams[n][4] out (NaN for a cell whose filter
 * length ucb_bandwidth(f0, fdot, cos theta, 1e-22) falls outside 16..2048).
 * Synchronises the stream once (after bucketing the cells by filter length,
 * so that only the non-empty lengths launch); not for graph capture. */
int glass_hip_fstat_eval(glass_hip_fstat *fstat, int n, const double *d_cells, double *d_logL_X, double *d_logL_AE,
                         double *d_Fparams, void *stream);
/* build_fstatistic_proposal over the band [fmin, fmax): n_f = (int)((fmax -
 * fmin) T) frequency cells of width d_f, n_theta x n_phi sky cells; tensor
 * (host, n_f * n_theta * n_phi, frequency-major; NULL to only get n_f and
 * d_f) receives the normalised proposal density, norm and maxp the
 * reference's proposal->norm and proposal->maxp. */
int glass_hip_fstat_proposal(glass_hip_fstat *fstat, double fmin, double fmax, int n_theta, int n_phi, double *tensor,
                             int *n_f, double *d_f, double *norm, double *maxp, void *stream);

#ifdef __cplusplus
}
#endif
#endif
