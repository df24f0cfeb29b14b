/*
 * glass_hip_noise.h -- instrument-noise model, noise likelihoods and the
 * batched instrument-noise MCMC of libglass_hip (gfx950).
 *
 * Drop-ins keep the reference's arithmetic and argument meaning with plain
 * arrays in place of its structs; return codes are GLASS_HIP_* (0 = ok),
 * likelihoods return NaN on error; glass_hip_last_error() has the message.
 *
 * Array layouts: C / invC are nine planes [3][3][N] (plane (i,j) at
 * (3 i + j) N, nch x nch planes for nch channels); X, Y, Z in the Fourier
 * domain are interleaved complex double[2N]; link parameters are sacc[6],
 * soms[6] in the link order 12, 21, 23, 32, 13, 31 (the reference's
 * map_array_to_noise_params, noise/src/glass_noise_model.c:24-44).
 */
#ifndef GLASS_HIP_NOISE_H
#define GLASS_HIP_NOISE_H

#include <stddef.h>

#include "glass_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* generate_instrument_noise_model(struct Orbit*, struct InstrumentModel*)
 * (noise/src/glass_noise_model.c:389-545): model->psd->C[i][j][n] for the
 * bins model->psd->f[n], nch = model->psd->Nchannel (1, 2, 3).  C is in/out:
 * entries the reference leaves alone (f = 0 bins; C[1][0] of two channels,
 * which it only divides by 4) keep the caller's values. */
int glass_hip_instrument_noise_model(double fstar, const double *f, int N, int nch, const double *sacc,
                                     const double *soms, double *C);

/* generate_instrument_noise_model_wavelet(struct Wavelets*, struct Orbit*,
 * struct InstrumentModel*) (:561-610): the model at the layer centres
 * (imin + k) df of the active layers imin = round(f[0]/df) ..
 * round(f[N-1]/df), divided by ND = NF NT; planes of stride N; *nactive
 * receives the number of layers written. */
int glass_hip_instrument_noise_model_wavelet(double fstar, const double *f, int N, double df, int ND,
                                             const double *sacc, const double *soms, double *C, int *nactive);

/* my_noise_log_likelihood(struct Data*, struct Noise*) (:1275-1306) */
double glass_hip_my_noise_log_likelihood(const double *X, const double *Y, const double *Z, const double *invC,
                                         const double *logdetC, int N);
/* noise_log_likelihood(struct Data*, struct Noise*) (:1104-1134), 3 channels */
double glass_hip_noise_log_likelihood(const double *X, const double *Y, const double *Z, const double *invC,
                                      const double *logdetC, int N);
/* my_noise_log_likelihood_wavelet(struct Data*, struct Noise*) (:1182-1206):
 * real pixels k = i + j NT (i < NT, j < Nlayer), invC planes of stride
 * ldc >= NT Nlayer, Ntot = data->N for the normalisation constant. */
double glass_hip_my_noise_log_likelihood_wavelet(const double *X, const double *Y, const double *Z,
                                                 const double *invC, const double *logdetC, int NT, int Nlayer,
                                                 int Ntot, int ldc);

/* ---- batched instrument-noise sampler (no single reference function:
 * noise_instrument_model_mcmc, :395-592 of noise/src/glass_noise_sampler.c,
 * for every chain at once, and noise_ptmcmc, :70-117) ---- */
typedef struct glass_hip_noise_plan glass_hip_noise_plan;

/* One data segment of N bins (every f[n] > 0) shared by nladder ladders of
 * ntemp temperature slots.  X, Y, Z host interleaved complex [2N]; Cadd
 * (NULL for none) the fixed extra covariance [3][3][N] added to every
 * chain's instrument model (generate_full_covariance_matrix,
 * noise/src/glass_noise_model.c:894-907).  Device memory: 10 N doubles
 * (+ 6 N with Cadd) plus about 40 doubles per slot. */
int glass_hip_noise_plan_create(glass_hip_noise_plan **plan, int N, int nladder, int ntemp, double fstar,
                                const double *f, const double *X, const double *Y, const double *Z,
                                const double *Cadd, void *stream);
int glass_hip_noise_plan_destroy(glass_hip_noise_plan *plan);
/* my_noise_log_likelihood of the model of n <= nladder*ntemp parameter sets
 * (device [n][12]: sacc[6], soms[6]) plus Cadd -> device logL[n] */
int glass_hip_noise_logl(glass_hip_noise_plan *plan, int n, const double *d_params, double *d_logL,
                         void *stream);
/* Host arrays, slot-major (slot s = ladder * ntemp + t): params[S][12] of
 * the states (state s starts in slot s), rand_r seeds[S] and
 * temperature[S] of the slots.  Computes every state's logL. */
int glass_hip_noise_set_state(glass_hip_noise_plan *plan, const double *params, const unsigned *seeds,
                              const double *temperature, void *stream);
/* Any pointer may be NULL.  params/logL by state, seeds/index/nacc/swaps by
 * slot (index: slot -> state; nacc: accepted moves; swaps: accepted PT swaps
 * with the slot above). */
int glass_hip_noise_get_state(glass_hip_noise_plan *plan, double *params, double *logL, unsigned *seeds,
                              int *index, int *nacc, int *swaps, void *stream);
/* ncalls calls of noise_instrument_model_mcmc (10 moves each) for every
 * slot, then, if swap, one noise_ptmcmc pass per ladder; graph != 0 runs it
 * as one hipGraph (captured once per (ncalls, swap)). */
int glass_hip_noise_mcmc(glass_hip_noise_plan *plan, int ncalls, int swap, int graph, void *stream);
int glass_hip_noise_ptswap(glass_hip_noise_plan *plan, void *stream);
int glass_hip_noise_plan_info(glass_hip_noise_plan *plan, int *nchunk, int *chunk, size_t *bytes);

/* n uniforms rand_r_U_0_1 then n normals rand_r_N_0_1 of the device's
 * restatement of the C library's rand_r, from one seed (checking aid) */
int glass_hip_rand_stream(unsigned seed, int n, double *u, double *g);

#ifdef __cplusplus
}
#endif
#endif
