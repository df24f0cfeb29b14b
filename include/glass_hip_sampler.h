/*
 * glass_hip_sampler.h -- the UCB sampler's per-step work on the device:
 * proposals, priors, Metropolis-Hastings, parallel-tempering swaps, and the
 * Fisher eigen-decomposition its Fisher proposal reads (libglass_hip, gfx950).
 *
 * Reference (paths relative to the reference root):
 *   ucb_mcmc                     ucb/src/glass_ucb_sampler.c:75-173
 *   step loop of the sampler     apps/src/ucb_mcmc.c:230-259
 *   draw_from_uniform_prior      utils/src/glass_sampler.c:96-107
 *   draw_from_fisher             ucb/src/glass_ucb_proposal.c:348-391
 *   fm_shift, psi_phi_jump       ucb/src/glass_ucb_proposal.c:577-613
 *   draw_from_fstatistic, jump_from_fstatistic,
 *   evaluate_fstatistic_proposal ucb/src/glass_ucb_proposal.c:1526-1673
 *   proposal table and weights   ucb/src/glass_ucb_proposal.c:615-760
 *   evaluate_ucb_prior et al.    ucb/src/glass_ucb_prior.c:396-646
 *   ptmcmc, adapt_temperature_ladder  utils/src/glass_sampler.c:19-94
 *   matrix_eigenstuff            utils/src/glass_math.c:475-499
 *   rand_r_U_0_1 / rand_r_N_0_1  utils/src/glass_utils.c:19-35
 * Return codes are GLASS_HIP_* (0 = ok); glass_hip_last_error() has the message.
 */
#ifndef GLASS_HIP_SAMPLER_H
#define GLASS_HIP_SAMPLER_H

#include "glass_hip.h"
#include "glass_hip_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Eigen-decomposition of n symmetric 8x8 matrices, device arrays F[n][8][8]
 * -> evec[n][8][8] (evec[i][j] = component i of eigenvector j, the layout
 * draw_from_fisher reads) and eval[n][8], eigenvalues ascending, each
 * eigenvector's largest-magnitude component positive (cyclic Jacobi);
 * sweeps[n] (may be NULL) receives the Jacobi sweeps used.  Replaces the
 * LAPACKE_dgeev of matrix_eigenstuff, whose order and signs are not fixed. */
int glass_hip_fisher_eigen(int n, const double *d_F, double *d_evec, double *d_eval, int *d_sweeps, void *stream);

/* matrix_eigenstuff(double **matrix, double **evectors, double *evalues, int N)
 * (utils/src/glass_math.c:475-499) on host row-major arrays; N = 8. */
int glass_hip_matrix_eigenstuff(const double *matrix, double *evectors, double *evalues, int N);

typedef struct glass_hip_ucb_sampler glass_hip_ucb_sampler;

typedef struct
{
    /* proposal weights in the reference's table order: 0 uniform prior,
     * 1 F-statistic draw, 2 F-statistic jump, 3 extrinsic prior, 4 Fisher,
     * 5 fm shift, 6 psi-phi jump, 7 GMM prior, 8 covariance.  3, 7 and 8
     * are not built and must be 0; 1 and 2 need every window's proposal
     * density (glass_hip_ucb_sampler_set_fstat). */
    double weight[9];
    int fix_sky, fix_freq, fix_fdot; /* flags->fixSky, fixFreq, fixFdot */
    double snr_min;                  /* 5 in the reference */
} glass_hip_ucb_sampler_opts;

/* A sampler over the chains of a batched plan: nladder ladders of ntemp
 * temperature slots, nladder * ntemp = the plan's chains; slot s starts on
 * chain (state) s.  prior[G][8][2] per data window (model->prior,
 * ucb/src/glass_ucb_prior.c:221-398, set by the caller); every chain needs
 * at least one live source.  The reference's galaxy prior, GMM prior,
 * calibration and RJ moves are not built (their flags off). */
int glass_hip_ucb_sampler_create(glass_hip_ucb_sampler **sampler, glass_hip_fd_batch *plan, int nladder, int ntemp,
                                 const double *prior, const glass_hip_ucb_sampler_opts *opts, void *stream);
int glass_hip_ucb_sampler_destroy(glass_hip_ucb_sampler *sampler);

/* rand_r seeds and temperatures of the slots, logLnorm per chain (NULL: 0),
 * adapt_iter0 the first iteration count adapt_temperature_ladder sees;
 * resets the counters and the slot -> chain map to the identity. */
int glass_hip_ucb_sampler_set_state(glass_hip_ucb_sampler *sampler, const unsigned *seeds, const double *temperature,
                                    const double *logLnorm, int adapt_iter0, void *stream);

/* The F-statistic proposal density of window g (setup_fstatistic_proposal,
 * ucb/src/glass_ucb_proposal.c:916-1084): host tensor [n_f][n_theta][n_phi]
 * (normalised, e.g. from glass_hip_fstat_proposal), cell width d_f in Hz,
 * its maximum maxp, and the window's first bin data->qmin.  Refused
 * (GLASS_HIP_EINVAL, checked on the host) unless every cell is finite and
 * >= 0, the maximum is > 0 and maxp is at most that maximum: the device's
 * rejection sampler could otherwise never accept.  A draw that still finds
 * no cell in 2^24 tries leaves the proposal's f0 and sky unchanged and is
 * reported by glass_hip_ucb_sampler_get_state (GLASS_HIP_ERUNTIME). */
int glass_hip_ucb_sampler_set_fstat(glass_hip_ucb_sampler *sampler, int g, const double *tensor, int n_f, double d_f,
                                    int n_theta, int n_phi, double maxp, int data_qmin, void *stream);

/* Fisher matrices of every live source at its current parameters (the
 * plan's exact derivatives, replacing ucb_fisher) and their eigen-
 * decompositions, stored per (chain, source) for the Fisher proposal. */
int glass_hip_ucb_sampler_fisher(glass_hip_ucb_sampler *sampler, void *stream);
/* host arrays [C][maxsrc][8][8] and [C][maxsrc][8] of the stored eigen data */
int glass_hip_ucb_sampler_set_eigen(glass_hip_ucb_sampler *sampler, const double *evec, const double *eval,
                                    void *stream);
int glass_hip_ucb_sampler_get_eigen(glass_hip_ucb_sampler *sampler, double *evec, double *eval, void *stream);

/* nsteps calls of ucb_mcmc on every slot (the RJ coin of the step loop is
 * drawn and ignored); accepted moves update the plan's chains in place. */
int glass_hip_ucb_sampler_steps(glass_hip_ucb_sampler *sampler, int nsteps, void *stream);
/* ptmcmc on every ladder; adapt != 0 then runs adapt_temperature_ladder
 * with the iteration counter, which it advances */
int glass_hip_ucb_sampler_ptswap(glass_hip_ucb_sampler *sampler, int adapt, void *stream);
/* one iteration of the reference's outer loop: nsteps steps, the Fisher
 * update, ptmcmc (+ adapt); graph != 0 runs it as one hipGraph (captured
 * once per (nsteps, adapt)) */
int glass_hip_ucb_sampler_sweep(glass_hip_ucb_sampler *sampler, int nsteps, int adapt, int graph, void *stream);

/* Any pointer may be NULL.  Per slot: seeds, index (slot -> chain),
 * temperature, trials[S][9] / accepts[S][9] per proposal type, swaps
 * (accepted ptmcmc swaps with the slot above).  Returns GLASS_HIP_ERUNTIME
 * (after copying) if an F-statistic draw gave up since set_state. */
int glass_hip_ucb_sampler_get_state(glass_hip_ucb_sampler *sampler, unsigned *seeds, int *index, double *temperature,
                                    int *trials, int *accepts, int *swaps, void *stream);

/* save_chain_state (ucb/src/glass_ucb_io.c:418-446) of every ladder, each
 * ladder being one run of the reference: dirs[l]/chain_state_<t>.dat holds
 * the chain now at temperature t of ladder l (slot l*ntemp + t through the
 * slot -> chain map that ptmcmc swaps permute), with logLmax[l] (NULL: 0),
 * that chain's logLnorm (NULL: 0), t0 and eta[neta].  The reference's
 * restore reads file t back into temperature t with the map reset to the
 * identity, which glass_hip_ucb_sampler_set_state does.  Synchronises. */
int glass_hip_ucb_sampler_save_chain_states(glass_hip_ucb_sampler *sampler, const char *const *dirs, int step,
                                            const double *logLmax, const double *logLnorm, double t0,
                                            const double *eta, int neta, void *stream);

#ifdef __cplusplus
}
#endif
#endif
