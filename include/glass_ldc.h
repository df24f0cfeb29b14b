/*
 * glass_ldc.h -- C ABI of the LDC HDF5 readers (glass_amd/libglass_ldc.so,
 * host code linked against libhdf5).
 *
 * Both readers are two calls: the sample count, then a fill of caller
 * arrays of that many doubles (no memory crosses the boundary).  Every
 * function returns GLASS_LDC_OK (0) or a negative code, with the message in
 * glass_ldc_last_error() on the calling thread.  HDF5's own error printing
 * is off during each call.  Callable from several host threads when libhdf5
 * is a thread-safe build (H5_HAVE_THREADSAFE, as the image's 1.10.6 is),
 * which serialises the HDF5 calls themselves.
 */
#ifndef GLASS_LDC_H
#define GLASS_LDC_H

#ifdef __cplusplus
extern "C" {
#endif

#define GLASS_LDC_OK 0
#define GLASS_LDC_EINVAL -1   /* bad argument (null pointer, wrong Nsamples)      */
#define GLASS_LDC_EIO -4      /* file or dataset missing or unreadable           */
#define GLASS_LDC_EFORMAT -5  /* dataset present but not of the expected layout  */

const char *glass_ldc_last_error(void);

/* Replaces void LISA_Read_HDF5_LDC_TDI(struct TDI *tdi, char *fileName,
 *   const char *dataName)  (utils/src/glass_lisa.h:327,
 *   utils/src/glass_lisa.c:1394-1454), e.g. dataName "/obs/tdi" (Sangria
 *   data) or "/sky/mbhb/tdi", "/sky/dgb/tdi", "/sky/igb/tdi", "/sky/vgb/tdi"
 *   (glass_data.c:660-751).  The dataset: one column of compound rows with
 *   double members t, X, Y, Z (matched by name; more members are ignored).
 *   glass_ldc_tdi_samples gives tdi->N (the row count, >= 2);
 *   glass_ldc_read_tdi fills X, Y, Z, A, E, T (Nsamples each, A/E/T by
 *   XYZ2AET, glass_lisa.c:429-455) and *delta = t[1] - t[0]. */
int glass_ldc_tdi_samples(const char *fileName, const char *dataName, int *Nsamples);
int glass_ldc_read_tdi(const char *fileName, const char *dataName, int Nsamples, double *X, double *Y,
                       double *Z, double *A, double *E, double *T, double *delta);

/* Replaces void LISA_Read_HDF5_LDC_RADLER_TDI(struct TDI *tdi,
 *   char *fileName)  (utils/src/glass_lisa.h:336,
 *   utils/src/glass_lisa.c:1456-1506): dataset "/H5LISA/PreProcess/TDIdata",
 *   an Nrow x Ncol matrix (Ncol >= 4, Nrow >= 4) whose columns 1, 2, 3 are
 *   X, Y, Z and column 0 the time.  Nsamples = Nrow/2, the size the
 *   reference gives tdi (its loop writes all Nrow rows past those arrays;
 *   here the first Nrow/2 rows are read); *delta = row 1 col 0 - row 0 col 0. */
int glass_ldc_radler_samples(const char *fileName, int *Nsamples);
int glass_ldc_read_radler_tdi(const char *fileName, int Nsamples, double *X, double *Y, double *Z,
                              double *A, double *E, double *T, double *delta);

#ifdef __cplusplus
}
#endif
#endif
