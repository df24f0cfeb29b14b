/*
 * glass_hip_batch.h -- batched, device-resident UCB waveform + likelihood
 * evaluation on gfx950.  No counterpart in the reference: the reference
 * evaluates one proposal per OpenMP thread (apps/src/ucb_mcmc.c:230,
 * ucb/src/glass_ucb_sampler.c:139-155); this API evaluates a whole sweep of
 * proposals (chains x temperatures x sources) in one stream-ordered,
 * graph-capturable launch sequence.
 *
 * Model.  A plan holds G data windows ("groups": the frequency segment a
 * set of parallel-tempered chains analyses) and C chains.  Chain c belongs
 * to group g(c) and carries nsrc(c) <= 8 galactic binaries; the plan keeps,
 * on the device, every source's current template (X,Y,Z, "sangria" unless
 * another format is given), every chain's residual r_c = d_g - sum_s h_cs
 * and its logL_c = -1/2 (r_c|r_c).
 *
 * A proposal is (chain c, source s, params[8]).  Its evaluation returns
 *   logL  = log-likelihood of chain c with source s moved to params:
 *           logL_c + (r_c|dh) - 1/2 (dh|dh),  dh = h(params) - h_cs,
 *           equal to generate_ucb_model(source_id=s) followed by
 *           gaussian_log_likelihood (ucb/src/glass_ucb_model.c:55-103,
 *           utils/src/glass_model.c:726-774) over the whole window;
 *   snr   = snr() of the proposed template as the sampler's rejection test
 *           computes it (utils/src/glass_math.c:403-426: template-local bins
 *           against window-indexed invC).
 * ucb_alignment (BW, qmin; ucb/src/glass_ucb_waveform.c:698-736) is
 * evaluated on the device for every proposal.
 *
 * Layouts (all FP64, complex = interleaved re,im):
 *   data:  per group, 3 channels x W_g complex, groups concatenated
 *   invC:  per group, 6 planes x W_g: XX, YY, ZZ, XY, XZ, YZ
 *   src_params: C x 8 x 8 (C x max_src x 8 with glass_hip_fd_batch_create_ex;
 *               unused source slots ignored)
 * Device memory beyond the data: per chain its residual (3 W_g complex)
 * and, for each live source plus one accept slot, 3 cap_g complex with cap_g
 * the largest power of two <= W_g (the largest BW ucb_bandwidth allows in
 * the window); e.g. 4000 windows x 256 chains of ~780 bins with ~5 sources
 * each take ~38 GB of residuals and ~150 GB of templates.
 * Entry points taking `stream` accept a hipStream_t (NULL = default
 * stream) and only enqueue work: they neither allocate nor synchronise, so
 * a caller may capture them in a hipGraph.  The plan's own graph path is
 * glass_hip_fd_batch_eval_graph.
 */
#ifndef GLASS_HIP_BATCH_H
#define GLASS_HIP_BATCH_H

#include "glass_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glass_hip_fd_batch glass_hip_fd_batch;

/* Creates a plan and uploads the data (host arrays, copied).  format:
 * "sangria" | "phase" | "frequency".  max_prop: largest proposal count a
 * single eval will be given.  Windows hold 32 <= W_g <= 8191 bins (the
 * largest template ucb_bandwidth allows there is 4096 bins); wider windows
 * are refused with GLASS_HIP_EINVAL.  Computes every source's template, the
 * residuals and logL_c on the device before returning.  An evaluated item
 * whose template fits no class gets logL -inf and snr 0 (and is counted in
 * the plan's error count), so a sampler never accepts it. */
int glass_hip_fd_batch_create(glass_hip_fd_batch **plan, const glass_hip_orbit *orbit,
                              const char *format, double T, double t0, int ngroup,
                              const int *group_qmin, const int *group_W, const double *data,
                              const double *invC, int nchain, const int *chain_group,
                              const int *chain_nsrc, const double *src_params, int max_prop);

/* Same with max_src source slots per chain (1..64; src_params is nchain x
 * max_src x 8), for data windows holding more than eight binaries. */
int glass_hip_fd_batch_create_ex(glass_hip_fd_batch **plan, const glass_hip_orbit *orbit,
                                 const char *format, double T, double t0, int ngroup,
                                 const int *group_qmin, const int *group_W, const double *data,
                                 const double *invC, int nchain, const int *chain_group,
                                 const int *chain_nsrc, int max_src, const double *src_params,
                                 int max_prop);

int glass_hip_fd_batch_destroy(glass_hip_fd_batch *plan);

/* Device memory the plan holds: its template slots and in total (bytes). */
int glass_hip_fd_batch_memory(glass_hip_fd_batch *plan, long long *template_bytes,
                              long long *total_bytes);

/* Device buffers owned by the plan, for callers that fill proposals on the
 * device: chain [max_prop] int, src [max_prop] int, params [max_prop*8],
 * logL [max_prop], snr [max_prop]. */
int glass_hip_fd_batch_buffers(glass_hip_fd_batch *plan, int **d_chain, int **d_src,
                               double **d_params, double **d_logL, double **d_snr);

/* Evaluates nprop proposals held in device memory (any device pointers;
 * the plan's own buffers work).  Results to d_logL, d_snr (device). */
int glass_hip_fd_batch_eval(glass_hip_fd_batch *plan, int nprop, const int *d_chain,
                            const int *d_src, const double *d_params, double *d_logL,
                            double *d_snr, void *stream);

/* Same, replaying a hipGraph of the launch sequence built on first use for
 * this (nprop, pointer) combination. */
int glass_hip_fd_batch_eval_graph(glass_hip_fd_batch *plan, int nprop, const int *d_chain,
                                  const int *d_src, const double *d_params, double *d_logL,
                                  double *d_snr, void *stream);

/* Accepts proposals: for each index i in d_idx[0..nacc) (indices into the
 * last evaluated proposal arrays, at most one per chain), moves the source
 * to its proposed parameters, updates the chain residual and sets logL_c to
 * the proposal's logL. */
int glass_hip_fd_batch_accept(glass_hip_fd_batch *plan, int nacc, const int *d_idx,
                              const int *d_chain, const int *d_src, const double *d_params,
                              const double *d_logL, void *stream);

/* Recomputes every template, residual and logL_c from the current
 * parameters (removes drift accumulated by accepts). */
int glass_hip_fd_batch_refresh(glass_hip_fd_batch *plan, void *stream);

/* Profiling of the eager glass_hip_fd_batch_eval (not the graph path):
 * with on != 0 every eval records HIP events around each BW-class kernel
 * on the launch stream, waits for them and accumulates the elapsed time
 * per class (evals become synchronous).  Turning it on or off clears the
 * totals.  BW class k covers BW = 32 << k, k = 0..7. */
int glass_hip_fd_batch_profile(glass_hip_fd_batch *plan, int on);

/* ms[8]: summed kernel time per class; launches[8]: evals timed;
 * count[8]: proposals per class in the last timed eval. */
int glass_hip_fd_batch_profile_read(glass_hip_fd_batch *plan, double *ms, int *launches,
                                    int *count);

/* Concurrent BW classes (no reference counterpart; on by default): with
 * on != 0 the class kernels of an eval, an accept, a template rebuild or
 * the exact derivatives run on three side streams of the plan forked from and joined back into
 * the launch stream (parallel branches under graph capture), so small and
 * latency-bound classes overlap; with on = 0 they run in turn on the
 * launch stream.  Results are identical either way.  Drops the captured
 * graph. */
int glass_hip_fd_batch_set_concurrent(glass_hip_fd_batch *plan, int on);

/* Launch shape ("tile") of one BW class's fused eval kernel (no reference
 * counterpart): nthr threads per workgroup (one template per workgroup;
 * fewer threads than template bins take several bins each), the grid
 * (gmul 0 = one workgroup per proposal slot, gmul > 0 = gmul x (workgroups
 * resident per CU) x CUs workgroups striding over the class's proposals),
 * and waves: the waves per SIMD the kernel's register allocation must
 * allow (1 = the compiler's own choice; 5 or 6 spill registers to keep more
 * workgroups resident).  Class k covers BW = 32 << k, k = 0..7.  Every
 * tile gives the same results up to the order of the final sums (~1e-16
 * relative). */
typedef struct glass_hip_tile
{
    int nthr;
    int gmul;
    int waves;
} glass_hip_tile;

/* The compiled tiles of class k: writes up to max of them to out and their
 * number to *n. */
int glass_hip_fd_batch_tile_candidates(int k, glass_hip_tile *out, int max, int *n);

/* The plan's current tiles (tiles[8]). */
int glass_hip_fd_batch_get_tiles(glass_hip_fd_batch *plan, glass_hip_tile *tiles);

/* Sets all eight tiles (each a compiled candidate of its class, e.g. from
 * an earlier autotune); drops the captured graph. */
int glass_hip_fd_batch_set_tiles(glass_hip_fd_batch *plan, const glass_hip_tile *tiles);

/* Autotuner: buckets the nprop proposals (device arrays, as for an eval)
 * into BW classes, times every compiled tile of each non-empty class over
 * them (one untimed launch, then reps timed launches, HIP events on
 * stream) and keeps the fastest per class in the plan.  Returns the chosen
 * tiles (best[8], may be NULL) and their mean launch times in ms
 * (best_ms[8], 0 for empty classes; may be NULL).  Synchronous; overwrites
 * the plan's own logL/snr buffers; drops the captured graph. */
int glass_hip_fd_batch_autotune(glass_hip_fd_batch *plan, int nprop, const int *d_chain,
                                const int *d_src, const double *d_params, int reps,
                                glass_hip_tile *best, double *best_ms, void *stream);

/* Copies the chains' current logL (host array of nchain) and, if not
 * NULL, their current source parameters (nchain*max_src*8, max_src = 8
 * for glass_hip_fd_batch_create). Synchronous. */
int glass_hip_fd_batch_state(glass_hip_fd_batch *plan, double *logL, double *src_params);

/* Exact derivatives (no reference counterpart; replaces the forward
 * differences of ucb_fisher, ucb/src/glass_ucb_waveform.c:364-522).
 * For n sources with parameters d_params[n*8] and templates of BW bins
 * (q and BW held fixed) placed at window offsets d_imin[n] of windows of W
 * bins: with R the residual WITHOUT the source (3 x W complex at
 * d_resid + e*resid_stride doubles; NULL = zero) and invC (6 planes x W
 * at d_invC + e*invC_stride doubles; stride 0 = one window for all),
 *   d_logL[e]   = L = (R|h) - (h|h)/2   (logL minus a theta-free constant)
 *   d_grad[e*8] = dL/dtheta = (R - h | dh/dtheta)
 *   d_fisher[e*64] = (dh/dtheta_k | dh/dtheta_l)
 * in the parameter-array units of map_ucb_params_to_array
 * (ucb/src/glass_ucb_model.c:40-53).  mode 0 weights bin m of the
 * template with invC at window bin imin+m; mode 1 reproduces the
 * reference's Fisher alignment (invC at window bin m).  BW: a power of two,
 * 32..4096.  BW 2048 and 4096 run decimated kernels over at most one
 * workgroup per CU with a scratch slice each (0.6 / 1.6 MB) that the
 * calling host thread keeps (grown with hipMalloc on first use; calls of
 * one host thread on different streams must not overlap). */
int glass_hip_fd_source_derivs(const glass_hip_orbit *orbit, const char *format, double T,
                               double t0, int n, const double *d_params, int BW,
                               const int *d_imin, int W, const double *d_resid,
                               long long resid_stride, const double *d_invC,
                               long long invC_stride, int mode, double *d_logL, double *d_grad,
                               double *d_fisher, void *stream);

/* Reverse-mode (adjoint) gradient: the same L and dL/dtheta as
 * glass_hip_fd_source_derivs (same arguments and conventions; BW 32..1024)
 * without the Fisher matrix, by transposing the TDI and FFT steps: one
 * waveform, six more FFTs and one pass of per-sample tangents per source
 * (no reference counterpart; for gradient-only consumers such as autograd
 * and gradient-based proposals).  d_grad may be NULL: L only. */
int glass_hip_fd_source_grad(const glass_hip_orbit *orbit, const char *format, double T,
                             double t0, int n, const double *d_params, int BW,
                             const int *d_imin, int W, const double *d_resid,
                             long long resid_stride, const double *d_invC,
                             long long invC_stride, int mode, double *d_logL, double *d_grad,
                             void *stream);

/* Exact derivatives of chain sources at their current parameters, from
 * the plan's state (no reference counterpart; for Fisher-matrix proposals,
 * draw_from_fisher, ucb/src/glass_ucb_proposal.c:348-391): item e = (chain
 * d_chain[e], source d_src[e]) (device arrays, n <= the plan's capacity,
 * max(max_prop, live sources, nchain)).  With R the chain's residual plus
 * the source's stored template (the residual without the source) and the
 * window's invC, as glass_hip_fd_source_derivs mode 0:
 *   d_logL[e] = (R|h) - (h|h)/2,  d_grad[8e..] = (R - h | dh/dtheta),
 *   d_fisher[64e..] = (dh/dtheta_k | dh/dtheta_l).
 * Sources with a BW outside 32..4096, or invalid indices, get NaN outputs.  Graph
 * safe (no allocation); shares the plan's class lists with evals, so it
 * must not overlap an eval of the same plan on another stream. */
int glass_hip_fd_batch_derivs(glass_hip_fd_batch *plan, int n, const int *d_chain, const int *d_src,
                              double *d_logL, double *d_grad, double *d_fisher, void *stream);

/* One sampler sweep as one hipGraph (built on first use for this set of
 * counts and pointers, replayed after): glass_hip_fd_batch_eval of nprop
 * proposals followed by glass_hip_fd_batch_derivs of nder (chain, source)
 * items for Fisher-matrix proposals. */
int glass_hip_fd_batch_sweep_graph(glass_hip_fd_batch *plan, int nprop, const int *d_chain,
                                   const int *d_src, const double *d_params, double *d_logL,
                                   double *d_snr, int nder, const int *d_dchain, const int *d_dsrc,
                                   double *d_dlogL, double *d_grad, double *d_fisher,
                                   void *stream);

/* Batched full-window likelihood: n independent calls of
 *   double gaussian_log_likelihood(struct Data *data, struct Model *model)
 * (utils/src/glass_model.h:219, utils/src/glass_model.c:726-774), Nchannel
 * 3, in one launch (no reference counterpart as a batch).  Item e has a
 * window of W[e] bins; its data occupy 3 channels (X, Y, Z) x 2 W[e]
 * doubles at d_data + 6*data_off[e], its model the same at d_model +
 * 6*model_off[e], its inverse covariance the 6 planes XX, YY, ZZ, XY, XZ,
 * YZ of W[e] doubles at d_invC + 6*invC_off[e] (offsets in bins; items
 * may share data or invC by sharing offsets).  The plan holds the item
 * table (host arrays at create, copied) on the current device; an eval
 * writes d_logL[e] = -1/2 chi^2 of item e.  Evals on one plan must not
 * overlap (they share the plan's partial sums). */
typedef struct glass_hip_logl_batch glass_hip_logl_batch;

int glass_hip_logl_batch_create(glass_hip_logl_batch **plan, int n, const int *W,
                                const long long *data_off, const long long *model_off,
                                const long long *invC_off);
int glass_hip_logl_batch_destroy(glass_hip_logl_batch *plan);
int glass_hip_logl_batch_eval(glass_hip_logl_batch *plan, const double *d_data,
                              const double *d_model, const double *d_invC, double *d_logL,
                              void *stream);

/* glass_hip_logl_batch_eval without a plan and without allocating: no
 * counterpart in the reference; for callers that cannot call hipMalloc
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  The item table is passed as device arrays (d_W[n] and the
 * three offset arrays, in bins, as glass_hip_logl_batch_create takes them on
 * the host) with max_W >= every W[e]; the chunk partial sums go to a static
 * device array of 2^20 doubles (8 MiB per GPU, held until the process ends):
 * n * ceil(max_W / 2048) <= 2^20, else GLASS_HIP_EINVAL.  Results equal
 * glass_hip_logl_batch_eval's bit for bit (same chunks, same order); an item
 * with W[e] <= 0 or W[e] > max_W gets NaN.  Enqueues two kernels on stream,
 * no allocation, copy or synchronisation.  Every call on one GPU shares its
 * scratch array: calls that overlap on one GPU (other streams, other host
 * threads) race and must be ordered by the caller. */
int glass_hip_logl_batch_eval_noalloc(int n, const int *d_W, const long long *d_data_off,
                                      const long long *d_model_off, const long long *d_invC_off,
                                      int max_W, const double *d_data, const double *d_model,
                                      const double *d_invC, double *d_logL, void *stream);

#ifdef __cplusplus
}
#endif
#endif
