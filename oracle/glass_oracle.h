/*
 * glass_oracle.h -- CPU restatement of the GLASS galactic-binary (UCB)
 * likelihood path, used ONLY as test infrastructure: the parity checker for
 * the HIP library and the timed CPU baseline.  Nothing in the product links
 * or loads this library.
 *
 * Every function restates one reference function; the citation on each
 * declaration names the reference file:line it follows (paths relative to the
 * reference repository root).  Arithmetic is FP64 with the reference's
 * operation order; the library is compiled with -ffp-contract=off so no
 * multiply-add is fused (the reference's parity build, see oracle/Makefile).
 *
 * Pinning (what checks each function independently of this restatement):
 *   orc_fft_*            -> bitwise vs kissfft built from the reference sources
 *                           (oracle/_ref), brute-force O(N^2) DFT, impulse
 *                           known answers (apps/src/fft_wdm_unit_tests.c:454-545)
 *   orc_fourier_nwip,
 *   orc_snr, orc_gaussian_log_likelihood,
 *   orc_wavelet_nwip     -> closed-form sums evaluated by numpy on small inputs
 *   orc_ucb_waveform     -> invariants the formulas fix: linearity in A, the
 *                           phi0 global phase, psi -> psi+pi/2 sign flip, A/E as
 *                           linear combinations of X/Y/Z, carrier-bin
 *                           placement, and a brute-force time-domain evaluation
 *                           of the same one-arm response (tests/)
 *   orc_ucb_bandwidth    -> closed form by hand at chosen inputs
 *   orc_invert_noise_covariance -> C * invC = I
 *   orc_instrument_noise_model -> equal-link parameters reproduce the
 *                           reference's closed forms XYZnoise_FF/XYZcross_FF
 *                           with get_noise_levels("sangria"); the cyclic
 *                           relabelling of the spacecraft permutes X,Y,Z;
 *                           linear in every link parameter
 *   orc_my_noise_log_likelihood, orc_noise_log_likelihood,
 *   orc_my_noise_log_likelihood_wavelet -> numpy per-bin inverse and
 *                           log-determinant on small inputs
 *   orc_rand_*           -> the C library's rand_r itself
 *   orc_jacobi_eigen8    -> numpy.linalg.eigh (LAPACK) on random and graded
 *                           matrices, A v = lambda v, orthonormality
 *   orc_ucb_propose, orc_ucb_prior, orc_ucb_accept, orc_ucb_ptmcmc,
 *   orc_adapt_temperature_ladder -> closed-form prior values and wraps, the
 *                           Fisher jump covariance F^-1/8, the fm-shift and
 *                           psi-phi jump laws, Metropolis rates, a hand-worked
 *                           ladder update
 *   orc_get_Fstat_logL   -> the F-statistic identity: on noise-free data of
 *                           the cell's own intrinsic parameters F = (d|d)/2
 *                           for any extrinsic parameters, quadratic in the
 *                           data; the fit recovers cos(iota); the inverse
 *                           against numpy
 *   orc_noise_mcmc_call, orc_noise_ptmcmc -> detailed balance of the swap
 *                           (exact swap probabilities), and the cold chain of
 *                           a long run settling on the injected noise levels
 *   orc_ucb_fisher_fd    -> symmetric, and equal to the Gram matrix of the
 *                           finite-difference waveform derivatives
 *   libm (glibc sin/cos/exp/log/pow): unpinned, ulp level only.
 */
#ifndef GLASS_ORACLE_H
#define GLASS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Analytic (rigid, equal-arm) LISA orbit constants.
 * utils/src/glass_lisa.c:125-136 (initialize_analytic_orbit). */
typedef struct
{
    double L;      /* arm length [m]                 */
    double fstar;  /* transfer frequency c/(2 pi L)  */
    double R;      /* AU * eccentricity              */
} orc_orbit;

void orc_init_analytic_orbit(orc_orbit *orbit);

/* utils/src/glass_lisa.c:77-124; x,y,z are 1-indexed arrays of length 4 */
void orc_analytic_orbits(const orc_orbit *orbit, double t, double *x, double *y, double *z);

/* ---------------------------------------------------------------- FFT ---- */
/* Mixed-radix decimation-in-time FFT with kissfft's factorisation and
 * butterflies (external/kissfft/kiss_fft.c:15-330, twiddles :356-363).
 * Complex data interleaved re,im.  Unnormalised in both directions. */
int  orc_fft(const double *in, double *out, int n, int inverse);
/* glass_forward_complex_fft, utils/src/glass_math.c:703-730 (in place) */
void orc_forward_complex_fft(double *data, int n);
/* glass_inverse_complex_fft, utils/src/glass_math.c:740-766 (in place, /N) */
void orc_inverse_complex_fft(double *data, int n);
/* kiss_fftr forward real FFT, external/kissfft/kiss_fftr.c:20-117:
 * n real samples in, n/2+1 complex bins out (interleaved). */
int  orc_fftr(const double *timedata, double *freqdata, int n);
/* kiss_fftri inverse real FFT (unnormalised), kiss_fftr.c:119-153 */
int  orc_fftri(const double *freqdata, double *timedata, int n);

/* ------------------------------------------------------- FD waveform ---- */
/* ucb/src/glass_ucb_waveform.c:738-1001.  format: "sangria" (default
 * production TDI, utils/src/glass_lisa.c:645-710), "phase" (:458-549) or
 * "frequency" (:551-643).  X..E receive 2*BW doubles each (template-local
 * bins; bin m <-> absolute bin q - BW/2 + m).  Returns 0, or -1 for an
 * unsupported format (the reference prints and exits, :971-975). */
int orc_ucb_waveform(const orc_orbit *orbit, const char *format, double T, double t0,
                     const double *params, int NParams,
                     double *X, double *Y, double *Z, double *A, double *E,
                     int BW, int NI);

/* The 6 slow link series before TDI, for tests: d[link][2*BW] after the
 * FFT, fftshift and 1/(2BW) normalisation (glass_ucb_waveform.c:940-962).
 * link order: 12, 21, 31, 13, 23, 32. */
void orc_ucb_slow_series(const orc_orbit *orbit, double T, double t0, const double *params,
                         int NParams, int BW, double *d);

/* ucb/src/glass_ucb_waveform.c:698-725 */
int orc_ucb_bandwidth(double L, double fstar, double f, double fdot, double costheta,
                      double A, double T, int N);
/* ucb/src/glass_ucb_waveform.c:727-736 with map_array_to_ucb_params
 * (ucb/src/glass_ucb_model.c:25-38).  out: BW, qmin, imin */
void orc_ucb_alignment(const orc_orbit *orbit, double T, int NFFT, int data_qmin,
                       const double *params, int *BW, int *qmin, int *imin);

/* utils/src/glass_lisa.c:1188-1198 */
double orc_AEnoise(double L, double fstar, double f);
/* utils/src/glass_lisa.c:1082-1119 ("sangria" levels), :1121-1135 */
void   orc_noise_levels_sangria(double f, double *Spm, double *Sop);
double orc_XYZnoise_FF(double fstar, double f, double Spm, double Sop);
double orc_XYZcross_FF(double fstar, double f, double Spm, double Sop);
/* 3-channel sangria covariance of utils/src/glass_data.c:1156-1242 (no
 * confusion noise): C[9][N] row-major 3x3 planes for bins fmin + n/T. */
void orc_xyz_noise_covariance(const orc_orbit *orbit, double T, double fmin, int N, double *C);
/* utils/src/glass_math.c:240-289, Nchannel = 3.  C, invC: 9 planes of N. */
void orc_invert_noise_covariance(const double *C, double *invC, double *logdetC, int N);

/* ------------------------------------------------- inner products ------- */
/* utils/src/glass_math.c:362-380 (factor 2, not 4) */
double orc_fourier_nwip(const double *a, const double *b, const double *invC, int N);
/* utils/src/glass_math.c:382-391 */
double orc_wavelet_nwip(const double *a, const double *b, const double *invC, const int *list, int N);
/* utils/src/glass_math.c:403-426, Nchannel = 3: template-local X,Y,Z
 * against window-indexed invC (the reference's alignment), N bins. */
double orc_snr(const double *X, const double *Y, const double *Z, const double *invC, int Ninv, int N);
/* utils/src/glass_model.c:726-774.  data/model: X,Y,Z,(A,E) arrays of 2*NFFT;
 * invC 9 planes of NFFT (row-major 3x3).  Nchannel 1 (X), 2 (A,E), 3 (XYZ). */
double orc_gaussian_log_likelihood(const double *dX, const double *dY, const double *dZ,
                                   const double *dA, const double *dE,
                                   const double *mX, const double *mY, const double *mZ,
                                   const double *mA, const double *mE,
                                   const double *invC, int NFFT, int Nchannel);

/* ucb/src/glass_ucb_model.c:192-232 (Nchannel 3): model[j] += template[i],
 * j = i + imin, 0 <= j < NFFT */
void orc_add_ucb_model(double *mX, double *mY, double *mZ, const double *X, const double *Y,
                       const double *Z, int BW, int imin, int NFFT);

/* ----------------------------------------------------------- Fisher ----- */
/* ucb/src/glass_ucb_waveform.c:364-522 up to (not including) the
 * eigensolve: forward difference eps=1e-6 on the parameter array, cosine
 * clipping, per-step ucb_alignment, Gram over wave_p->BW bins of 9
 * fourier_nwip channel-pair terms with invC taken from window bin 0,
 * NaN -> 10.  Nchannel 3, "sangria" unless format given. fisher: 8x8. */
void orc_ucb_fisher_fd(const orc_orbit *orbit, const char *format, double T, double t0,
                       int NFFT, int data_qmin, const double *params,
                       const double *invC, double *fisher);

/* polarization tensors and propagation vector, 1-indexed
 * (utils/src/glass_lisa.c:1566-1596) */
void orc_polarization_tensor(double costh, double phi, double ep[4][4], double ec[4][4], double k[4]);

/* ------------------------------------------------------------- WDM ----- */
/* The fields of struct Wavelets (utils/src/glass_wavelet.h:25-67) that the
 * wavelet path reads, filled by hand for an NF x NT grid of pixel duration
 * dt: df = 1/(2 dt), A = 0 (setup_wdm_basis, utils/src/glass_wavelet.c:39-49),
 * imin/imax = 100 / NT-100 (WAVELET_EDGE_BUFFER), kmin = NT (pixel (0,1)),
 * kmax = NT*(NF-1) (pixel (0,NF-1)); window = build_wdm_filter_freq(.., NF,
 * NT, A, forward) in caller memory of NT/2+1 doubles. */
typedef struct
{
    int NF, NT;
    double dt, df, A;
    int kmin, kmax, imin, imax;
    double *window;
} orc_wdm;

void   orc_wdm_init(orc_wdm *w, int NF, int NT, double dt, double *window);
/* utils/src/glass_math.c:1214-1242 */
double orc_incomplete_beta(double a, double b, double x);
/* utils/src/glass_wavelet.c:327-349 */
void   orc_wdm_filter(double *phif, int Nf, int Nt, double A, int forward);
/* utils/src/glass_wavelet.c:410-460: freq ND/2+1 complex bins -> ND pixels
 * (pixel (n, m) at n + NT m) */
void   orc_wavelet_transform_freq(const orc_wdm *w, const double *freq, double *wdm);
/* utils/src/glass_wavelet.c:293-303 (in place, ND = NF*NT real samples) */
void   orc_wavelet_transform_timefreq(const orc_wdm *w, double *timedata);
/* utils/src/glass_wavelet.c:351-405: ND pixels -> ND/2+1 complex bins */
void   orc_wavelet_transform_inverse_freq(const orc_wdm *w, const double *wdm, double *freq);
/* utils/src/glass_wavelet.c:544-611 (in place, NT*(Nlayers+1) samples) */
void   orc_wavelet_transform_timefreq_by_layers(const orc_wdm *w, int jmin, int Nlayers,
                                                const double *phif, double *timedata);
/* utils/src/glass_wavelet.c:865-900: NF+1 layer variances from a one-sided
 * DFT PSD of ND/2+1 bins */
void   orc_wdm_layer_var(const orc_wdm *w, const double *dft_psd, double *layer_var);

/* natural cubic spline (utils/src/glass_math.c:98-190): c = y0|y1|y2|y3,
 * N doubles each; lookup even_sampled_search (even != 0) or binary_search
 * (utils/src/glass_math.c:393-467) */
void   orc_spline_coefficients(int N, const double *x, const double *y, double *c);
double orc_spline_eval(int N, const double *x, const double *c, int even, double xv);
int    orc_binary_search(const double *a, int lo, int hi, double x);
int    orc_even_search(const double *a, int lo, int hi, double x);

/* initialize_interpolated_analytic_orbits (utils/src/glass_lisa.c:271-380):
 * N = orc_iorbit_size(Tobs) knots t[N], positions pos[9][N] (spacecraft i,
 * coordinate d at row 3i+d) and their spline coefficients coef[9][4][N]. */
int    orc_iorbit_size(double Tobs);
void   orc_iorbit_fill(const orc_orbit *o, double Tobs, double t0, int N, double *t, double *pos,
                       double *coef);

/* utils/src/glass_math.c:1244-1308 and :1034-1052 */
void   orc_extract_amplitude_and_phase(int Ns, double *As, double *Dphi, const double *M,
                                       const double *Mf, const double *phiR);
void   orc_unwrap_phase(int N, double *phase);

/* 0 (default): the template path calls the platform libm's sin/cos, as the
 * reference does; 1: correctly rounded sin/cos (see orc_wdm.c). */
void   orc_set_cr_trig(int on);
void   orc_set_wavelet_layers(int L);

/* ucb/src/glass_ucb_waveform.c:1003-1037 */
void   orc_ucb_wavelet_layers(const orc_wdm *w, double Tobs, const double *params, int *jstart,
                              int *jwidth);
/* The interpolated TDI series A(t) cos(Phi(t) - 2 pi fhet t) of
 * ucb_waveform_wavelet at N samples t0 + i dt_s (out: 3 x N); see orc_wdm.c */
void   orc_set_numeric_orbit(int Norb, const double *t, const double *coef);
double orc_numeric_orbit_armlength(int N, const double *pos);
void   orc_ucb_barycenter_waveform(const double *params, int N, const double *times, double *phase,
                                   double *amp, double T);
void   orc_ucb_wavelet_series(int Norb, const double *ot, const double *oc, double Tobs, double t0,
                              const double *params, int N, double dt_s, double fhet, double *out);
/* The response stage of ucb_waveform_wavelet alone: barycentre phase at the
 * knots, phase at spacecraft 1, signed TDI amplitude and unwrapped phase
 * (3 x Norb each) at the response times. */
void   orc_ucb_wavelet_response(int Norb, const double *ot, const double *oc, double Tobs, double t0,
                                const double *params, double *phase_ssb, double *phase_sc,
                                double *amp3, double *ph3);
/* ucb_waveform_wavelet (ucb/src/glass_ucb_waveform.c:1061-1240) with the
 * orbit given as its knots/coefficients.  The compact form writes the
 * pixel values in list order (wX[n] belongs to pixel list[n]); the drop-in
 * form scatters them into X,Y,Z at list[n] as the reference does.  Both
 * return the number of layers. */
int    orc_ucb_waveform_wavelet_compact(const orc_wdm *w, int Norb, const double *ot,
                                        const double *oc, double Tobs, double t0,
                                        const double *params, int *list, int *Nwavelet,
                                        double *wX, double *wY, double *wZ);
int    orc_ucb_waveform_wavelet(const orc_wdm *w, int Norb, const double *ot, const double *oc,
                                double Tobs, double t0, const double *params, int *list,
                                int *Nwavelet, double *X, double *Y, double *Z);
/* utils/src/glass_math.c:428-438; invC 9 planes of Npix */
double orc_snr_wavelet(const double *X, const double *Y, const double *Z, const double *invC,
                       int Npix, const int *list, int Nlist);
/* utils/src/glass_model.c:805-855; data/model 3 planes of Npix, invC 9 */
double orc_gaussian_log_likelihood_wavelet(const orc_wdm *w, const double *dX, const double *dY,
                                           const double *dZ, const double *mX, const double *mY,
                                           const double *mZ, const int *mlist, int Nlist,
                                           const double *invC, int Npix);
/* ucb_fisher_wavelet (ucb/src/glass_ucb_waveform.c:524-696) before the
 * eigensolve: central differences (epsilon 1e-6) of the wavelet template on
 * the plus template's pixel list, nine-term wavelet_nwip Fisher over all
 * pixels, NaN -> 10; invC 9 planes of Npix; fisher[64] row-major. */
int    orc_ucb_fisher_wavelet(const orc_wdm *w, int Norb, const double *ot, const double *oc,
                              double Tobs, double t0, const double *params, const double *invC,
                              int Npix, double *fisher);
/* CPU baseline of the wavelet workload: per template the compact waveform,
 * snr_wavelet and the logL change the template makes (see orc_wdm.c);
 * data 3 planes and invC 9 planes of Npix.  Returns wall seconds. */
double orc_wdm_eval_batch(const orc_wdm *w, int Norb, const double *ot, const double *oc,
                          double Tobs, double t0, int n, const double *params, const double *data,
                          const double *invC, int Npix, double *dlogL, double *snr, int *nwave,
                          int nthreads);

/* ------------------------------------------------------ CPU baseline ---- */
/* One MCMC-step likelihood evaluation the way the reference's sampler
 * performs it (ucb/src/glass_ucb_sampler.c:139-155 -> generate_ucb_model,
 * ucb/src/glass_ucb_model.c:55-103, then gaussian_log_likelihood):
 * zero the window model, regenerate source `sid` from `params`, add every
 * source's stored template, return the full-window logL.  Templates of the
 * other sources are taken from `tmpl` (3 channels x 2*BW each, laid out as
 * described in orc_batch.c).  Evaluations are independent and run under
 * OpenMP with nthreads threads (0 = runtime default). */
typedef struct
{
    int     nsrc;         /* live sources of the chain                       */
    int     W;            /* window bins (NFFT)                              */
    int     qmin;         /* absolute bin of window bin 0                    */
    const double *data;   /* 3 x 2W: X,Y,Z data                              */
    const double *invC;   /* 9 x W                                           */
    const double *src_params; /* nsrc x 8 current parameters                 */
    const double *tmpl;   /* nsrc x 3 x 2*BWmax current templates (bins 0..BW) */
    const int    *src_bw;   /* nsrc current BW                              */
    const int    *src_imin; /* nsrc current imin                            */
    int     bwmax;        /* template stride                                 */
} orc_chain;

double orc_eval_logl(const orc_orbit *orbit, double T, double t0, const orc_chain *chain,
                     int sid, const double *params);
/* Evaluate n proposals (chain index, source index, params[8] each);
 * returns logL in out.  nthreads 0 = default. Returns wall seconds. */
double orc_eval_batch(const orc_orbit *orbit, double T, double t0, const orc_chain *chains,
                      int n, const int *chain_of, const int *sid_of, const double *params,
                      double *out, int nthreads);

/* ------------------------------------------------------- noise model ---- */
/* generate_instrument_noise_model, noise/src/glass_noise_model.c:389-545.
 * f[N] bin frequencies, nch 1/2/3 channels, sacc[6]/soms[6] in the link
 * order 12,21,23,32,13,31 (:24-44); C[nch][nch][N] (entries the reference
 * leaves alone -- f = 0 bins, C[1][0] of two channels -- are not written
 * apart from the final /4 of the latter). */
void orc_instrument_noise_model(double fstar, const double *f, int N, int nch, const double *sacc,
                                const double *soms, double *C);
/* generate_instrument_noise_model_wavelet, :561-610; C planes of stride N;
 * returns the number of active layers written, -1 if it exceeds N */
int orc_instrument_noise_model_wavelet(double fstar, const double *f, int N, double df, int ND,
                                       const double *sacc, const double *soms, double *C);
/* my_noise_log_likelihood, :1275-1306 (X,Y,Z interleaved complex [2N]) */
double orc_my_noise_log_likelihood(const double *X, const double *Y, const double *Z, const double *invC,
                                   const double *logdetC, int N);
/* noise_log_likelihood, :1104-1134, three channels */
double orc_noise_log_likelihood(const double *X, const double *Y, const double *Z, const double *invC,
                                const double *logdetC, int N);
/* my_noise_log_likelihood_wavelet, :1182-1206 (pixel k = i + j NT, planes
 * of stride ldc, Ntot = data->N) */
double orc_my_noise_log_likelihood_wavelet(const double *X, const double *Y, const double *Z,
                                           const double *invC, const double *logdetC, int NT, int Nlayer,
                                           int Ntot, int ldc);
/* rand_r_U_0_1 / rand_r_N_0_1, utils/src/glass_utils.c:19-35 (libc rand_r) */
double orc_rand_U_0_1(unsigned int *seed);
double orc_rand_N_0_1(unsigned int *seed);
/* instrument model of p12 = sacc[6],soms[6] (+ Cadd[9][N] if not NULL),
 * inverse, my_noise_log_likelihood; work holds 19 N doubles */
double orc_noise_logl_params(double fstar, const double *f, int N, const double *X, const double *Y,
                             const double *Z, const double *Cadd, const double *p12, double *work);
/* one call of noise_instrument_model_mcmc (noise/src/glass_noise_sampler.c:
 * 395-592) for one chain: x[12] and *logLx updated in place; returns the
 * number of accepted moves of its 10 */
int orc_noise_mcmc_call(double fstar, const double *f, int N, const double *X, const double *Y,
                        const double *Z, const double *Cadd, double temperature, double *x, double *logLx,
                        unsigned int *seed, double *work);
/* noise_ptmcmc, noise/src/glass_noise_sampler.c:70-117: one ladder of NC
 * slots; index[] slot -> state, logL[] per state, seeds[] per slot */
void orc_noise_ptmcmc(int NC, int *index, const double *temperature, const double *logL, unsigned int *seeds,
                      int *acceptance);

/* ------------------------------------------------------- UCB sampler --- */
/* Eigen-decomposition of a symmetric 8x8 matrix replacing matrix_eigenstuff
 * (utils/src/glass_math.c:475-499): cyclic Jacobi; evec[i][j] = component i
 * of eigenvector j, eigenvalues ascending, largest component of each vector
 * positive.  Returns the sweeps used. */
int orc_jacobi_eigen8(const double *F, double *evec, double *eval);
/* evaluate_ucb_prior, galaxy/update flags off (ucb/src/glass_ucb_prior.c:
 * 536-646); pr[8][2] = model->prior, lpv[8] its log volumes; wraps p */
double orc_ucb_prior(double *p, const double *pr, const double *lpv);
/* model->logPriorVolume (ucb/src/glass_ucb_prior.c:396) */
void orc_ucb_log_prior_volume(const double *pr, double *lpv);
/* the draws and proposal of one ucb_mcmc step (ucb/src/glass_ucb_sampler.c:
 * 75-136, with the RJ coin of apps/src/ucb_mcmc.c:242): chain with nlive
 * sources sparams[nlive][8] (the chosen source's params wrapped in place by
 * the prior), their eigen data evec[nlive][64], eval[nlive][8]; returns the
 * source n, proposal type np, proposed y[8], log priors of x and y */
typedef struct
{
    const double *ten; /* [nf][nt][nph] normalised F-statistic proposal density */
    int nf, nt, nph;
    double d_f, maxp, qmin; /* cell width (Hz), maximum, data->qmin of the window */
} orc_fstat_prop;
void orc_ucb_propose(unsigned int *seed, int nlive, double *sparams, const double *evec, const double *eval,
                     const double *pr, const double *lpv, const double *w, double T, int fix_sky, int fix_freq,
                     int fix_fdot, const orc_fstat_prop *fs, int *n_out, int *np_out, double *y, double *lpx_out,
                     double *lpy_out, double *lq_out);
/* the rest of the step when the prior of y is finite (:138-172): snr cut,
 * Hastings ratio with lq = log q(x|y) - log q(y|x), accept draw; returns 1
 * to accept */
int orc_ucb_accept(unsigned int *seed, double logLy, double logLx, double snr, double snr_min, double temperature,
                   double lpx, double lpy, double lq);
/* ptmcmc (utils/src/glass_sampler.c:19-59) with model->logLnorm */
void orc_ucb_ptmcmc(int NC, int *index, const double *temperature, const double *logL, const double *lognorm,
                    unsigned int *seeds, int *acceptance);
/* adapt_temperature_ladder (utils/src/glass_sampler.c:61-94) */
void orc_adapt_temperature_ladder(int NC, double *temperature, const int *acceptance, int mcmc);

/* ----------------------------------------------------- F-statistic ---- */
/* get_Fstat_logL (ucb/src/glass_ucb_fstatistic.c:646-703): X, A, E the
 * data DFT [2 NFFT] from bin qmin0, C00 [NFFT]; returns the filter length */
int orc_get_Fstat_logL(const orc_orbit *orbit, const char *format, double T, double t0, int NFFT, int qmin0,
                       const double *X, const double *A, const double *E, const double *C00, double f0, double fdot,
                       double theta, double phi, double *logL_X, double *logL_AE, double *Fparams);
/* the 4x4 inverse used for the F-statistic Grams (row-major in/out) */
void orc_invert4(const double *in, double *out);
/* ucb_fdot (ucb/src/glass_ucb_waveform.c:307-314) */
double orc_ucb_fdot(double Mc, double f0);

#ifdef __cplusplus
}
#endif
#endif
