/* Writes HDF5 files in the LDC layouts for the reader tests (the reference
 * holds no HDF5 fixture).  Built by tests/test_ldc.py into a shared library
 * with gcc -lhdf5 and called through ctypes:
 *
 *   ldc_write_compound: a compound dataset of rows {member_0, ..., member_m}
 *     (doubles, names and order given), shape (N) or (N, ncol) for
 *     ncol > 0, at a path whose parent groups are created;
 *   ldc_write_matrix: an Nrow x Ncol double matrix (the RADLER layout);
 *   ldc_write_int_compound: {t, X, Y, Z} stored as 32-bit integers, which
 *     HDF5 converts on read.
 */
#include <hdf5.h>
#include <stdlib.h>
#include <string.h>

static hid_t create_file(const char *path)
{
    return H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
}

static hid_t make_lcpl(void)
{
    hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
    H5Pset_create_intermediate_group(lcpl, 1);
    return lcpl;
}

/* cols: m arrays of N doubles (row-major rows are assembled here) */
int ldc_write_compound(const char *path, const char *dset, int N, int ncol, int m, const char **names,
                       const double **cols)
{
    hid_t f = create_file(path);
    if (f < 0) return -1;
    size_t rowsz = (size_t)m * sizeof(double);
    hid_t t = H5Tcreate(H5T_COMPOUND, rowsz);
    for (int j = 0; j < m; j++) H5Tinsert(t, names[j], (size_t)j * sizeof(double), H5T_IEEE_F64LE);
    hsize_t dims[2] = {(hsize_t)N, (hsize_t)(ncol > 0 ? ncol : 1)};
    int rank = ncol > 0 ? 2 : 1;
    hid_t sp = H5Screate_simple(rank, dims, NULL);
    hid_t lcpl = make_lcpl();
    hid_t d = H5Dcreate2(f, dset, t, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT);
    size_t rows = (size_t)N * (size_t)(ncol > 0 ? ncol : 1);
    double *buf = malloc(rows * rowsz);
    for (size_t r = 0; r < rows; r++)
        for (int j = 0; j < m; j++) buf[r * m + j] = cols[j][r % (size_t)N] + (double)(r / (size_t)N);
    herr_t w = H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    free(buf);
    H5Dclose(d);
    H5Pclose(lcpl);
    H5Sclose(sp);
    H5Tclose(t);
    H5Fclose(f);
    return w < 0 ? -2 : 0;
}

int ldc_write_matrix(const char *path, const char *dset, int Nrow, int Ncol, const double *data)
{
    hid_t f = create_file(path);
    if (f < 0) return -1;
    hsize_t dims[2] = {(hsize_t)Nrow, (hsize_t)Ncol};
    hid_t sp = H5Screate_simple(2, dims, NULL);
    hid_t lcpl = make_lcpl();
    hid_t d = H5Dcreate2(f, dset, H5T_IEEE_F64LE, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT);
    herr_t w = H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(d);
    H5Pclose(lcpl);
    H5Sclose(sp);
    H5Fclose(f);
    return w < 0 ? -2 : 0;
}

typedef struct
{
    int t, X, Y, Z;
} irow;

int ldc_write_int_compound(const char *path, const char *dset, int N, const int *vals)
{
    hid_t f = create_file(path);
    if (f < 0) return -1;
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(irow));
    H5Tinsert(t, "t", HOFFSET(irow, t), H5T_STD_I32LE);
    H5Tinsert(t, "X", HOFFSET(irow, X), H5T_STD_I32LE);
    H5Tinsert(t, "Y", HOFFSET(irow, Y), H5T_STD_I32LE);
    H5Tinsert(t, "Z", HOFFSET(irow, Z), H5T_STD_I32LE);
    hsize_t dims[1] = {(hsize_t)N};
    hid_t sp = H5Screate_simple(1, dims, NULL);
    hid_t lcpl = make_lcpl();
    hid_t d = H5Dcreate2(f, dset, t, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT);
    herr_t w = H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, vals);
    H5Dclose(d);
    H5Pclose(lcpl);
    H5Sclose(sp);
    H5Tclose(t);
    H5Fclose(f);
    return w < 0 ? -2 : 0;
}
