/*
 * glass_hip_wdm.h -- C ABI of the wavelet-domain (WDM) part of
 * libglass_hip: the WDM transforms, galactic-binary templates in the
 * wavelet domain, and the wavelet inner products / likelihood.
 *
 * Drop-ins keep the reference function's arguments with struct Wavelets and
 * struct Orbit unpacked into the fields the path reads; they work on host
 * memory and are synchronous.  The batched entry points take device
 * pointers and only enqueue work on the given stream.
 */
#ifndef GLASS_HIP_WDM_H
#define GLASS_HIP_WDM_H

#include "glass_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The fields of struct Wavelets (utils/src/glass_wavelet.h:25-67) that the
 * wavelet path reads.  NF frequency layers x NT time pixels of duration dt
 * [s], layer bandwidth df [Hz] (1/(2 dt)), Meyer window parameter A (0 in
 * the reference, setup_wdm_basis utils/src/glass_wavelet.c:39-49), the
 * pixel range [kmin, kmax) that data arrays hold (array index k - kmin),
 * and the time pixels [imin, imax] the likelihood keeps.  The forward
 * window (window_freq_forward) is rebuilt from NF, NT and A exactly as
 * initialize_wavelet does (build_wdm_filter_freq, :327-349). */
typedef struct glass_hip_wdm
{
    int NF, NT;
    double dt, df, A;
    int kmin, kmax;
    int imin, imax;
} glass_hip_wdm;

/* Replaces void initialize_interpolated_analytic_orbits(struct Orbit *orbit,
 *   double Tobs, double t0) (utils/src/glass_lisa.c:271-380) for the fields
 * the wavelet path reads: *Norb = max(20, (int)(200 Tobs / YEAR)) knots
 * t[Norb] and the analytic spacecraft positions x, y, z (each 3 x Norb,
 * row i = spacecraft i).  With t == NULL only *Norb is returned. */
int glass_hip_interpolated_analytic_orbit(const glass_hip_orbit *orbit, double Tobs, double t0,
                                          int *Norb, double *t, double *x, double *y, double *z);

/* ------------------------------------------------------------ drop-ins */

/* Replaces void wavelet_transform_timefreq(struct Wavelets *wdm,
 *   double *timedata) (utils/src/glass_wavelet.h, glass_wavelet.c:293-303):
 * in place, ND = NF*NT samples in, ND pixels out (pixel (n, m) at
 * n + NT m).  NT a power of two in 32..4096, NF of the form 2^a or 3*2^a. */
int glass_hip_wavelet_transform_timefreq(const glass_hip_wdm *wdm, double *timedata);

/* Replaces void wavelet_transform_freq(struct Wavelets *wdm,
 *   double *freqdata, double *wdmdata) (glass_wavelet.c:410-460):
 * ND/2+1 complex bins (ND+2 doubles) in, ND pixels out. */
int glass_hip_wavelet_transform_freq(const glass_hip_wdm *wdm, const double *freqdata, double *wdmdata);

/* Replaces void wavelet_transform_inverse_freq(struct Wavelets *wdm,
 *   double *wdmdata, double *freqdata) (glass_wavelet.c:351-405):
 * ND pixels in, ND/2+1 complex bins (ND+2 doubles) out. */
int glass_hip_wavelet_transform_inverse_freq(const glass_hip_wdm *wdm, const double *wdmdata,
                                             double *freqdata);

/* Replaces void wavelet_transform_timefreq_by_layers(struct Wavelets *wdm,
 *   int jmin, int Nlayers, double *window, double *timedata)
 * (utils/src/glass_wavelet.h:119, glass_wavelet.c:544-611): in place on
 * NT (Nlayers + 1) heterodyned samples; window NT/2 + 1 doubles (the
 * (Nlayers+1)-layer filter of build_wdm_filter_freq).  Pixel (n, m) lands
 * at n Nlayers + m - jmin, the remaining samples are zeroed.  NT 256 or 512,
 * Nlayers 1..3 (ucb_wavelet_layers' range). */
int glass_hip_wavelet_transform_timefreq_by_layers(const glass_hip_wdm *wdm, int jmin, int Nlayers,
                                                   const double *window, double *timedata);

/* Replaces void stationary_dft_psd_to_wdm_psd(struct Wavelets *wdm,
 *   double *dft_psd, double *wdm_psd) (utils/src/glass_wavelet.c:901-911,
 * with stationary_dft_psd_to_wdm_layer_var :865-899): dft_psd the DFT
 * variance <|FFT[k]|^2> at k = 0..ND/2 (ND/2 + 1 doubles), wdm_psd the
 * ND = NF*NT pixel variances out (pixel (n, m) at n + NT m; row 0 holds
 * layer 0 at even n and layer NF at odd n).  NT even; the window is
 * build_wdm_filter_freq's of (NF, NT, A). */
int glass_hip_stationary_dft_psd_to_wdm_psd(const glass_hip_wdm *wdm, const double *dft_psd, double *wdm_psd);

/* Replaces void stationary_dft_psd_to_wdm_layer_var(struct Wavelets *wdm,
 *   double *dft_psd, double *layer_var) (utils/src/glass_wavelet.c:865-899):
 * the NF + 1 layer variances. */
int glass_hip_stationary_dft_psd_to_wdm_layer_var(const glass_hip_wdm *wdm, const double *dft_psd,
                                                  double *layer_var);

/* Replaces void ucb_waveform_wavelet(struct Orbit *orbit,
 *   struct Wavelets *wdm, double Tobs, double t0, double *params,
 *   int *wavelet_list, int *Nwavelet, double *X, double *Y, double *Z)
 * (ucb/src/glass_ucb_waveform.h:231, glass_ucb_waveform.c:1061-1240).
 * The orbit is given by its Norb knots orbit_t and the spacecraft
 * positions orbit_x/y/z, each 3 x Norb (row i = spacecraft i; struct Orbit
 * t, x[i], y[i], z[i], utils/src/glass_lisa.h:60-109), as
 * initialize_interpolated_analytic_orbits or an orbit file fill them.
 * Writes Nwavelet (<= 3 NT) list entries (pixel index - kmin) and X,Y,Z at
 * those entries only.  NT 256 or 512, Norb <= 512; the sample times must
 * lie inside the knots and inside [t0, t0 + Tobs]. */
int glass_hip_ucb_waveform_wavelet(const glass_hip_wdm *wdm, int Norb, const double *orbit_t,
                                   const double *orbit_x, const double *orbit_y,
                                   const double *orbit_z, double Tobs, double t0,
                                   const double *params, int *wavelet_list, int *Nwavelet, double *X,
                                   double *Y, double *Z);

/* Replaces double snr_wavelet(struct Source *source, struct Noise *noise)
 * (utils/src/glass_math.h:261, glass_math.c:428-438): X,Y,Z and invC (9
 * row-major 3x3 planes) of N pixels, list of Nlist indices. */
double glass_hip_snr_wavelet(const double *X, const double *Y, const double *Z, const double *invC,
                             int N, const int *list, int Nlist);

/* Replaces double gaussian_log_likelhood_wavelet(struct Data *data,
 *   struct Model *model) (utils/src/glass_model.h:220,
 * glass_model.c:805-855): data and model X,Y,Z of N = data->N pixels, the
 * model's list (Nlist entries), invC 9 planes of N; residual pixels on the
 * model list whose time index lies outside [imin, imax] are zeroed, as in
 * the reference. */
double glass_hip_gaussian_log_likelihood_wavelet(const glass_hip_wdm *wdm, const double *dX,
                                                 const double *dY, const double *dZ,
                                                 const double *mX, const double *mY,
                                                 const double *mZ, const int *mlist, int Nlist,
                                                 const double *invC, int N);

/* Replaces void ucb_fisher_wavelet(struct Orbit *orbit, struct Data *data,
 *   struct Source *source, struct Noise *noise)
 * (ucb/src/glass_ucb_waveform.h, glass_ucb_waveform.c:524-696) up to the
 * eigensolve: central differences (epsilon = 1e-6 on the parameter array,
 * cosines clipped to +-1 on the outer side) of ucb_waveform_wavelet, taken
 * on the plus template's pixel list at pixels whose time index lies
 * strictly inside (imin, imax); fisher[8*i+j] = the nine wavelet_nwip terms
 * XX, YY, ZZ, XY, XZ, YZ, YX, ZX, ZY over all pixels, NaN -> 10.  Orbit and
 * grid as glass_hip_ucb_waveform_wavelet; invC 9 row-major 3x3 planes of
 * N = kmax - kmin pixels (only the layers the templates touch are read).
 * The reference then calls matrix_eigenstuff (LAPACKE_dgeev) and, if an
 * eigenvalue is not positive, adds 0.1 to the diagonal once more
 * (:620-686); that stays on the host. */
int glass_hip_ucb_fisher_wavelet(const glass_hip_wdm *wdm, int Norb, const double *orbit_t,
                                 const double *orbit_x, const double *orbit_y,
                                 const double *orbit_z, double Tobs, double t0,
                                 const double *params, const double *invC, int N, double *fisher);

/* ------------------------------------------------ device-pointer API --- */

/* stationary_dft_psd_to_wdm_psd on device pointers (no reference
 * counterpart as a device call): d_dft_psd ND/2 + 1 in, d_layer_var NF + 1
 * out and, if not NULL, d_wdm_psd ND out; enqueues on stream, no
 * allocation after the window's first use on a grid. */
int glass_hip_wdm_psd(const glass_hip_wdm *wdm, const double *d_dft_psd, double *d_layer_var,
                      double *d_wdm_psd, void *stream);

/* Full forward transform on device memory: d_time (ND samples) ->
 * d_pixels (ND).  Same grid limits as glass_hip_wavelet_transform_timefreq.
 * Uses the calling thread's scratch (no reference counterpart). */
int glass_hip_wdm_transform(const glass_hip_wdm *wdm, const double *d_time, double *d_pixels,
                            void *stream);

/* Batched wavelet-domain templates (no reference counterpart: the
 * reference generates one template per call).  A plan holds the grid, the
 * orbit tables, the data (3 planes X,Y,Z of npix = kmax - kmin pixels,
 * host, copied) and invC (6 planes XX,YY,ZZ,XY,XZ,YZ of npix).  For n <=
 * max_templates parameter vectors d_params[n*8] (device) an eval writes
 *   d_snr[e]   = snr_wavelet of template e over its pixel list;
 *   d_dlogL[e] = gaussian_log_likelhood_wavelet(data, model = template e)
 *                - gaussian_log_likelhood_wavelet(data, empty model),
 * and keeps every template's Nwavelet, pixel list (vcap ints per template)
 * and values (3 channels x vcap doubles per template, in list order) in
 * plan buffers (glass_hip_wdm_batch_buffers); vcap = 3*NT
 * (glass_hip_wdm_batch_capacity). */
typedef struct glass_hip_wdm_batch glass_hip_wdm_batch;

int glass_hip_wdm_batch_create(glass_hip_wdm_batch **plan, const glass_hip_wdm *wdm, int Norb,
                               const double *orbit_t, const double *orbit_x, const double *orbit_y,
                               const double *orbit_z, double Tobs, double t0, const double *data,
                               const double *invC, int npix, int max_templates);
/* A plan whose templates all take `layers` frequency layers instead of the
 * reference's ucb_wavelet_layers count (1-3, ucb/src/glass_ucb_waveform.c:
 * 1003-1042): the same first layer and heterodyne frequency (jmin - 1) df,
 * then NT (layers + 1) heterodyned samples, the window of layers + 1 layers
 * and layers inverse transforms (glass_ucb_waveform.c:1110-1180 with
 * Nlayers = layers).  No reference counterpart: a workload knob for timing
 * templates that touch 8-64 layers.  layers + 1 a power of two from 8 to
 * 64, NT = 256 or 512, NT (layers + 1) <= 32768; vcap = NT * layers.
 * glass_hip_wdm_batch_fisher refuses such a plan. */
int glass_hip_wdm_batch_create_layers(glass_hip_wdm_batch **plan, const glass_hip_wdm *wdm, int Norb,
                                      const double *orbit_t, const double *orbit_x, const double *orbit_y,
                                      const double *orbit_z, double Tobs, double t0, const double *data,
                                      const double *invC, int npix, int max_templates, int layers);
/* Pixels per template and channel in the plan's list and value buffers. */
int glass_hip_wdm_batch_capacity(const glass_hip_wdm_batch *plan, int *vcap);
int glass_hip_wdm_batch_destroy(glass_hip_wdm_batch *plan);
int glass_hip_wdm_batch_buffers(glass_hip_wdm_batch *plan, int **d_nwave, int **d_list,
                                double **d_values);
int glass_hip_wdm_batch_eval(glass_hip_wdm_batch *plan, int n, const double *d_params,
                             double *d_dlogL, double *d_snr, void *stream);
/* Host copies of the last eval's first n templates (synchronous): nwave[n],
 * list[n*vcap], values[n*3*vcap] (channel-major per template). */
int glass_hip_wdm_batch_templates(glass_hip_wdm_batch *plan, int n, int *nwave, int *list,
                                  double *values);
/* ucb_fisher_wavelet (as glass_hip_ucb_fisher_wavelet) for n sources in
 * one launch sequence (no reference counterpart as a batch): d_params[n*8]
 * (device) -> d_fisher[n*64] (device), with the plan's invC (6 symmetric
 * planes).  Uses 16 template slots per source (16 n <= max_templates) and
 * overwrites the plan's template buffers.  A source whose 16 templates span
 * more than 6 layers gets NaN entries. */
int glass_hip_wdm_batch_fisher(glass_hip_wdm_batch *plan, int n, const double *d_params,
                               double *d_fisher, void *stream);
/* Diagnostics: the response stage of template e of the last eval (host copy,
 * synchronous): 64 header doubles, the barycentre phase spline (4 x Norb:
 * values, then the three coefficient rows) and the six TDI splines
 * (amp X, phase X, amp Y, phase Y, amp Z, phase Z; 4 x Norb each);
 * nout >= 64 + 28 Norb. */
int glass_hip_wdm_batch_scratch(glass_hip_wdm_batch *plan, int e, double *out, int nout);

#ifdef __cplusplus
}
#endif
#endif
