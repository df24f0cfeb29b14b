/*
 * glass_hip.h -- C ABI of the MI355X (gfx950) UCB likelihood library
 * (glass_amd/libglass_hip.so).
 *
 * Plain pointers and sizes only.  Two groups of entry points:
 *
 *  1. Drop-ins for the reference's C functions on the galactic-binary
 *     likelihood path.  Each takes the reference function's arguments with
 *     the reference structs unpacked into pointers (a shim per function in
 *     the reference tree is shown in INTEGRATION.md), works on host memory,
 *     is synchronous, and is safe to call from several host threads at once
 *     (each host thread owns a HIP stream and staging buffers).
 *  2. Device-pointer entry points (no reference counterpart): batched
 *     waveform + delta-logL evaluation, exact Fisher/gradient, WDM.
 *
 * Error convention: int-returning functions return GLASS_HIP_OK (0) or a
 * negative code; double-returning drop-ins return NaN on error.  The message
 * of the last error on the calling thread is glass_hip_last_error().
 * The reference exits on these conditions instead (e.g. unsupported TDI
 * format, ucb/src/glass_ucb_waveform.c:971-975); the INTEGRATION.md shims
 * turn a non-zero status back into that exit.
 */
#ifndef GLASS_HIP_H
#define GLASS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define GLASS_HIP_OK 0
#define GLASS_HIP_EINVAL -1      /* bad argument (size, format, pointer)   */
#define GLASS_HIP_ERUNTIME -2    /* HIP runtime failure                     */
#define GLASS_HIP_ENODEV -3      /* no usable gfx950 device                 */

/* Analytic (rigid) LISA orbit: the three constants of struct Orbit that
 * analytic_orbits() reads (utils/src/glass_lisa.h:60-109,
 * utils/src/glass_lisa.c:77-136).  L [m], fstar [Hz], R [m]. */
typedef struct glass_hip_orbit
{
    double L;
    double fstar;
    double R;
} glass_hip_orbit;

/* Fills the analytic orbit of initialize_analytic_orbit()
 * (utils/src/glass_lisa.c:125-136). */
void glass_hip_analytic_orbit(glass_hip_orbit *orbit);

/* An orbit read from a file: struct Orbit after initialize_numeric_orbit
 * (utils/src/glass_lisa.c:138-269), whose orbit_function interpolate_orbits
 * (:62-72) evaluates a natural cubic spline of every coordinate over the
 * file's Norb knots t.  x, y, z: 3 x Norb each (row i = spacecraft i, as
 * orbit->x[i]); L, fstar: the average arm length and c / (2 pi L) the
 * reference stores (glass_hip_numeric_orbit_armlength computes them). */
typedef struct glass_hip_numeric_orbit
{
    int Norb;
    const double *t;
    const double *x, *y, *z;
    double L;
    double fstar;
} glass_hip_numeric_orbit;

/* initialize_numeric_orbit's average arm length (glass_lisa.c:207-259) and
 * transfer frequency fstar = CLIGHT / (2 pi L) of an orbit file's samples. */
int glass_hip_numeric_orbit_armlength(int Norb, const double *x, const double *y, const double *z,
                                      double *L, double *fstar);

/* Selects the HIP device for the calling thread (default 0). */
int glass_hip_set_device(int device);

const char *glass_hip_last_error(void);

/* ------------------------------------------------------------ drop-ins */

/* Replaces void ucb_waveform(struct Orbit *orbit, char *format, double T,
 *   double t0, double *params, int NParams, double *X, double *Y, double *Z,
 *   double *A, double *E, int BW, int NI)
 * (ucb/src/glass_ucb_waveform.h:210, ucb/src/glass_ucb_waveform.c:738-1001).
 * format "sangria" | "phase" | "frequency"; writes 2*BW doubles to each of
 * X,Y,Z,A,E ("phase"/"frequency" with NI<=1 write X only, as the reference).
 * BW: a power of two, 16 <= BW <= 4096. */
int glass_hip_ucb_waveform(const glass_hip_orbit *orbit, const char *format, double T, double t0,
                           const double *params, int NParams, double *X, double *Y, double *Z,
                           double *A, double *E, int BW, int NI);

/* ucb_waveform with a numeric orbit (orbit->orbit_function ==
 * interpolate_orbits): the same template with the spacecraft positions at
 * the BW slow-sample times t0 + T n / BW taken from the orbit's splines
 * (coefficients and positions computed once per orbit, T, t0 and BW on the
 * host in the reference's operation order, then cached on the device).
 * The sample times must lie within the knots [t[0], t[Norb-1]]. */
int glass_hip_ucb_waveform_numeric(const glass_hip_numeric_orbit *orbit, const char *format, double T,
                                   double t0, const double *params, int NParams, double *X, double *Y,
                                   double *Z, double *A, double *E, int BW, int NI);

/* Replaces int ucb_bandwidth(double L, double fstar, double f, double fdot,
 *   double costheta, double A, double T, int N)
 * (ucb/src/glass_ucb_waveform.h, ucb/src/glass_ucb_waveform.c:698-725).
 * Host computation (integer result that sizes the template). */
int glass_hip_ucb_bandwidth(double L, double fstar, double f, double fdot, double costheta,
                            double A, double T, int N);

/* Replaces void ucb_alignment(struct Orbit *orbit, struct Data *data,
 *   struct Source *source)  (ucb/src/glass_ucb_waveform.c:727-736):
 * from params[8] (the reference's parameter array, NFFT = data->NFFT,
 * data_qmin = data->qmin) returns BW, qmin and imin (source fields). */
int glass_hip_ucb_alignment(const glass_hip_orbit *orbit, double T, int NFFT, int data_qmin,
                            const double *params, int *BW, int *qmin, int *imin);

/* Replaces double fourier_nwip(const double *a, const double *b,
 *   const double *invC, int N)  (utils/src/glass_math.h:243,
 *   utils/src/glass_math.c:362-380): 2 * sum_n Re(a_n conj b_n) invC_n. */
double glass_hip_fourier_nwip(const double *a, const double *b, const double *invC, int N);

/* Replaces double wavelet_nwip(const double *a, const double *b,
 *   const double *invC, const int *list, int N)  (utils/src/glass_math.h:244,
 *   utils/src/glass_math.c:382-391).  a, b, invC hold nmax doubles; list N
 *   indices into them. */
double glass_hip_wavelet_nwip(const double *a, const double *b, const double *invC, const int *list,
                              int N, int nmax);

/* Replaces double snr(struct Source *source, struct Noise *noise)
 * (utils/src/glass_math.h:255, utils/src/glass_math.c:403-426), 3-channel
 * case: X,Y,Z template arrays (2*N doubles), invC 9 planes of Ninv
 * (row-major 3x3, plane (a,b) at invC + (3a+b)*Ninv) read from bin 0 as the
 * reference does; N = source->tdi->N/2. */
double glass_hip_snr(const double *X, const double *Y, const double *Z, const double *invC,
                     int Ninv, int N);

/* Replaces double gaussian_log_likelihood(struct Data *data,
 *   struct Model *model)  (utils/src/glass_model.h:219,
 *   utils/src/glass_model.c:726-774).  d? and m? (data, model): channel arrays of
 *   2*NFFT doubles (X,Y,Z for Nchannel 3 or 1; A,E for Nchannel 2; unused
 *   ones may be NULL); invC 9 planes of NFFT. */
double glass_hip_gaussian_log_likelihood(const double *dX, const double *dY, const double *dZ,
                                         const double *dA, const double *dE, const double *mX,
                                         const double *mY, const double *mZ, const double *mA,
                                         const double *mE, const double *invC, int NFFT,
                                         int Nchannel);

/* Replaces void ucb_fisher(struct Orbit *orbit, struct Data *data,
 *   struct Source *source, struct Noise *noise)
 * (ucb/src/glass_ucb_waveform.h:150, ucb/src/glass_ucb_waveform.c:364-522)
 * up to the eigensolve: the forward-difference Fisher matrix with the
 * reference's conventions (eps = 1e-6 on the parameter array, cosine
 * clipping, per-step ucb_alignment, Gram over the last perturbed BW with
 * invC read from window bin 0, NaN -> 10), 3 channels.  params[8];
 * invC 9 planes of NFFT; fisher[64] row-major.  The reference then calls
 * matrix_eigenstuff (LAPACKE_dgeev) on the host. */
int glass_hip_ucb_fisher(const glass_hip_orbit *orbit, const char *format, double T, double t0,
                         int NFFT, int data_qmin, const double *params, const double *invC,
                         double *fisher);

/* Replaces void invert_noise_covariance_matrix(struct Noise *noise)
 * (utils/src/glass_math.c:240-289), Nchannel 3: C, invC 9 planes of N,
 * logdetC N (may be NULL). */
int glass_hip_invert_noise_covariance(const double *C, double *invC, double *logdetC, int N);

#ifdef __cplusplus
}
#endif
#endif
