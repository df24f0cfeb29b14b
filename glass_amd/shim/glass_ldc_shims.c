/*
 * glass_ldc_shims.c -- reference-side shims for the LDC HDF5 readers: drop
 * into the reference tree as utils/src/glass_ldc_shims.c, compile it with
 * -DGLASS_HIP next to glass_lisa.c built with the two originals renamed
 * (*_cpu), and link libglass_ldc.so.
 *
 * Each shim sizes struct TDI through the reference's alloc_tdi (the
 * originals do the same, glass_lisa.c:1421,1484), fills it through the C ABI
 * (include/glass_ldc.h) and turns an error into the reference's exit(1).
 *
 * GLASS_HIP_SHIM_SUBSET: compile against a caller-provided declaration of
 * struct TDI and alloc_tdi instead of the reference headers.
 */
#include <stdio.h>
#include <stdlib.h>

#ifndef GLASS_HIP_SHIM_SUBSET
#include "glass_utils.h"
#endif

#include "glass_ldc.h"

static void ldc_fail(const char *what)
{
    fprintf(stderr, "ERROR: %s: %s\n", what, glass_ldc_last_error());
    exit(1);
}

void LISA_Read_HDF5_LDC_TDI(struct TDI *tdi, char *fileName, const char *dataName)
{
    int N = 0;
    if (glass_ldc_tdi_samples(fileName, dataName, &N)) ldc_fail("LISA_Read_HDF5_LDC_TDI");
    alloc_tdi(tdi, N, 3);
    if (glass_ldc_read_tdi(fileName, dataName, N, tdi->X, tdi->Y, tdi->Z, tdi->A, tdi->E, tdi->T, &tdi->delta))
        ldc_fail("LISA_Read_HDF5_LDC_TDI");
}

void LISA_Read_HDF5_LDC_RADLER_TDI(struct TDI *tdi, char *fileName)
{
    int N = 0;
    if (glass_ldc_radler_samples(fileName, &N)) ldc_fail("LISA_Read_HDF5_LDC_RADLER_TDI");
    alloc_tdi(tdi, N, 3);
    if (glass_ldc_read_radler_tdi(fileName, N, tdi->X, tdi->Y, tdi->Z, tdi->A, tdi->E, tdi->T, &tdi->delta))
        ldc_fail("LISA_Read_HDF5_LDC_RADLER_TDI");
}
