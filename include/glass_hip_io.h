/* glass_hip_io.h -- the reference's UCB file formats on the host: the
 * verification-binary and injection files in, the chain, parameter,
 * likelihood, temperature and checkpoint files out and back, and a
 * checkpoint of a batched plan's device state.
 *
 * Text is produced with the reference's printf formats and read with its
 * scanf sequences, so files are byte-identical to the reference's for the
 * same values and interchangeable with them.  Parameters are the
 * reference's params[8] layout (f0 T, cos theta, phi, ln A, cos i, psi,
 * phi0, fdot T^2).  Return codes as in glass_hip.h; glass_hip_last_error()
 * describes a failure. */
#ifndef GLASS_HIP_IO_H
#define GLASS_HIP_IO_H

#include "glass_hip_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* print_source_params (ucb/src/glass_ucb_io.c:598-613): the source's line
 * without its newline, "%.16g " f0 then "%.12g " fdot, A, phi, cos theta,
 * cos i, psi, phi0, into buf (cap bytes, NUL-terminated).  Returns the
 * length, or -1 when cap is too small. */
int glass_hip_format_source_params(const double *params, double T, char *buf, int cap);

/* print_chain_state (utils/src/glass_sampler.c:153-161):
 * "%i %i %lg %lg %.12g \n" of step, Nlive, logL, logLnorm, t0. */
int glass_hip_format_chain_state(int step, int nlive, double logL, double logLnorm, double t0, char *buf, int cap);

/* print_psd_state (ucb/src/glass_ucb_io.c:580-589): "%i " step,
 * "%lg %lg " logL logLnorm, "%lg " per eta, "\n". */
int glass_hip_format_psd_state(int step, double logL, double logLnorm, const double *eta, int neta, char *buf,
                               int cap);

/* One line of print_chain_files' likelihood or temperature file
 * (ucb/src/glass_ucb_io.c:494-509): "%i " step, "%lg " per value, "\n";
 * values = logL + logLnorm per chain, or 1/T per chain. */
int glass_hip_format_value_line(int step, const double *values, int n, char *buf, int cap);

/* scan_source_params (ucb/src/glass_ucb_io.c:615-636): eight "%lg" from
 * text (f0, fdot, A, phi, cos theta, cos i, psi, phi0), mapped by
 * map_ucb_params_to_array (ucb/src/glass_ucb_model.c:40-53).  *consumed:
 * characters read. */
int glass_hip_scan_source_params(const char *text, double T, double *params, int *consumed);

/* GetVerificationBinary (ucb/src/glass_ucb_data.c:627-673) after the
 * header line UCBInjectVerificationSource strips (:88-93): f0, fdot, cos
 * theta, phi, m1, m2, cos i, d (kpc) -> A from the chirp mass and
 * distance, phi0 = 2 pi U and psi = pi/4 U from rand_r seeded with iseed,
 * and the data segment centred on f0: qmin = (int)((f0 - (NFFT/2)/T) T),
 * fmin = qmin/T, fmax = (qmin + NFFT)/T. */
int glass_hip_read_verification_binary(const char *path, double T, int NFFT, unsigned iseed, double *params,
                                       int *qmin, double *fmin, double *fmax);

/* The source file of UCBInjectSimulatedSource (ucb/src/glass_ucb_data.c:
 * 270-384): lines of f0, fdot, theta, phi, A, iota, psi, phi0, counted as
 * the reference counts them (one fewer than its read loop's passes),
 * psi > pi less pi, phi0 > 2 pi less 2 pi, cos theta = cos(pi/2 - theta),
 * cos i = cos(iota).  Writes up to max sources to params[max][8] and the
 * file's count to *n. */
int glass_hip_read_injection_file(const char *path, double T, int max, double *params, int *n);

/* save_chain_state (ucb/src/glass_ucb_io.c:418-446) of one chain to path:
 * "%.12g\n" logLmax, the chain-state line, the psd-state line (eta[neta]),
 * then one print_source_params line per live source (params[nlive][8]). */
int glass_hip_write_chain_state_file(const char *path, double logLmax, int step, int nlive, double logL,
                                     double logLnorm, double t0, const double *eta, int neta, const double *params,
                                     double T);

/* restore_chain_state's reads (ucb/src/glass_ucb_io.c:448-490) of one
 * chain-state file: logLmax, scan_chain_state (step, Nlive, logL, logLnorm,
 * t0), scan_psd_state (step, logL, logLnorm, eta[neta]), then Nlive
 * sources (at most maxsrc) into params[maxsrc][8].  The psd state's step
 * and likelihoods overwrite the chain state's, as in the reference.  Every
 * field must convert (the reference accepts a file as long as some field of
 * each read does, so it takes truncated files). */
int glass_hip_read_chain_state_file(const char *path, int maxsrc, int neta, double T, double *logLmax, int *step,
                                    int *nlive, double *logL, double *logLnorm, double *t0, double *eta,
                                    double *params);

/* save_chain_state for every chain of a batched plan (no reference
 * counterpart as one call): copies the plan's chain logL and live source
 * parameters from the device and writes dir/chain_state_<c>.dat for chain
 * c = 0..nchain-1, with logLnorm[c] (NULL: 0), the common t0 and eta
 * (neta values, NULL: none).  Synchronises the device.  Files are named by
 * plan chain: once a sampler's temperature swaps have permuted which chain
 * sits at which temperature, checkpoint with
 * glass_hip_ucb_sampler_save_chain_states instead (files by temperature). */
int glass_hip_fd_batch_save_chain_states(glass_hip_fd_batch *plan, const char *dir, int step, double logLmax,
                                         const double *logLnorm, double t0, const double *eta, int neta);

#ifdef __cplusplus
}
#endif

#endif
