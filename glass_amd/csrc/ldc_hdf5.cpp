// LDC HDF5 readers (host code, libglass_ldc.so, C ABI include/glass_ldc.h).
//
// The two readers the reference's ReadHDF5 (utils/src/glass_data.c:654-660)
// calls before it cuts, windows and Fourier-transforms the data:
//
//   LISA_Read_HDF5_LDC_TDI        utils/src/glass_lisa.c:1394-1454
//     a 1-d dataset of compound rows {t, X, Y, Z} (doubles, matched by
//     member name), e.g. "/obs/tdi", "/sky/dgb/tdi"; delta = t[1] - t[0]
//   LISA_Read_HDF5_LDC_RADLER_TDI  utils/src/glass_lisa.c:1456-1506
//     "/H5LISA/PreProcess/TDIdata", an Nrow x Ncol matrix whose columns
//     1..3 are X, Y, Z; delta = row 1 column 0 - row 0 column 0
//
// Both also fill A, E, T by XYZ2AET (glass_lisa.c:429-455).  Reading is
// split in two calls (sample count, then fill caller arrays) so no memory
// crosses the boundary.  Differences from the reference, all on inputs it
// does not check:
//   * missing files, datasets or members, fewer than 2 rows, a compound
//     dataset with more than one column, or a RADLER matrix with fewer than
//     4 columns return an error instead of reading garbage;
//   * the RADLER reader sizes its arrays for Nrow/2 samples (alloc_tdi at
//     :1484) and then writes all Nrow rows, past those arrays; here the
//     sample count is Nrow/2, the part the reference's arrays hold, and
//     exactly those rows are read.
#include "../../include/glass_ldc.h"

#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// Keeps HDF5's automatic error-stack printing off for the duration of one
// call (the codes returned here report the failure) and restores the
// caller's setting afterwards.
struct QuietH5
{
    H5E_auto2_t func = nullptr;
    void *data = nullptr;
    QuietH5()
    {
        H5Eget_auto2(H5E_DEFAULT, &func, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietH5() { H5Eset_auto2(H5E_DEFAULT, func, data); }
};

// Closes an HDF5 identifier with the matching routine when it leaves scope.
struct Hid
{
    hid_t id;
    herr_t (*close)(hid_t);
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid()
    {
        if (id >= 0) close(id);
    }
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
};

struct Row
{
    double t, X, Y, Z;
};

// XYZ2AET with the reference's constants (glass_lisa.c:429-455): A and E
// as XYZ2AE, T = (X + Y + Z) * 0.577350269189626.
inline void xyz_to_aet(double X, double Y, double Z, double *A, double *E, double *T)
{
    *A = (Z - X) * M_SQRT1_2;
    *E = (X - 2 * Y + Z) * 0.408248290463863;
    *T = (X + Y + Z) * 0.577350269189626;
}

// Opens fileName/dataName and returns the dataset's shape.
int open_dataset(const char *fileName, const char *dataName, Hid &file, Hid &dset,
                 std::vector<hsize_t> &dims)
{
    if (!fileName || !dataName) return fail(GLASS_LDC_EINVAL, "null file or dataset name");
    file.id = H5Fopen(fileName, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file.id < 0) return fail(GLASS_LDC_EIO, std::string("cannot open HDF5 file ") + fileName);
    dset.id = H5Dopen2(file.id, dataName, H5P_DEFAULT);
    if (dset.id < 0)
        return fail(GLASS_LDC_EIO, std::string("no dataset ") + dataName + " in " + fileName);
    Hid space(H5Dget_space(dset.id), H5Sclose);
    if (space.id < 0) return fail(GLASS_LDC_EIO, "cannot read the dataspace");
    int nd = H5Sget_simple_extent_ndims(space.id);
    if (nd < 1) return fail(GLASS_LDC_EFORMAT, std::string(dataName) + " is not an array");
    dims.assign(nd, 0);
    H5Sget_simple_extent_dims(space.id, dims.data(), nullptr);
    return GLASS_LDC_OK;
}

// Row count of an LDC compound TDI dataset, checked for the {t,X,Y,Z}
// members and a single column.
int tdi_rows(const char *fileName, const char *dataName, Hid &file, Hid &dset, hsize_t *rows)
{
    std::vector<hsize_t> dims;
    int rc = open_dataset(fileName, dataName, file, dset, dims);
    if (rc) return rc;
    for (size_t d = 1; d < dims.size(); d++)
        if (dims[d] != 1)
            return fail(GLASS_LDC_EFORMAT, std::string(dataName) + " has more than one column");
    Hid ftype(H5Dget_type(dset.id), H5Tclose);
    if (H5Tget_class(ftype.id) != H5T_COMPOUND)
        return fail(GLASS_LDC_EFORMAT, std::string(dataName) + " is not a compound dataset");
    for (const char *m : {"t", "X", "Y", "Z"})
        if (H5Tget_member_index(ftype.id, m) < 0)
            return fail(GLASS_LDC_EFORMAT, std::string(dataName) + " has no member " + m);
    if (dims[0] < 2) return fail(GLASS_LDC_EFORMAT, std::string(dataName) + " has fewer than 2 rows");
    if (dims[0] > 0x7fffffff) return fail(GLASS_LDC_EFORMAT, "more than 2^31-1 rows");
    *rows = dims[0];
    return GLASS_LDC_OK;
}

const char *kRadler = "/H5LISA/PreProcess/TDIdata";

int radler_shape(const char *fileName, Hid &file, Hid &dset, hsize_t *nrow, hsize_t *ncol)
{
    std::vector<hsize_t> dims;
    int rc = open_dataset(fileName, kRadler, file, dset, dims);
    if (rc) return rc;
    if (dims.size() != 2) return fail(GLASS_LDC_EFORMAT, std::string(kRadler) + " is not a matrix");
    if (dims[1] < 4) return fail(GLASS_LDC_EFORMAT, std::string(kRadler) + " has fewer than 4 columns");
    if (dims[0] < 4) return fail(GLASS_LDC_EFORMAT, std::string(kRadler) + " has fewer than 4 rows");
    if (dims[0] / 2 > 0x7fffffff) return fail(GLASS_LDC_EFORMAT, "more than 2^31-1 samples");
    *nrow = dims[0];
    *ncol = dims[1];
    return GLASS_LDC_OK;
}

int check_out(int Nsamples, int have, double *X, double *Y, double *Z, double *A, double *E, double *T,
              double *delta)
{
    if (Nsamples != have)
        return fail(GLASS_LDC_EINVAL, "Nsamples " + std::to_string(Nsamples) + " but the dataset holds " +
                                          std::to_string(have));
    if (!X || !Y || !Z || !A || !E || !T || !delta) return fail(GLASS_LDC_EINVAL, "null output array");
    return GLASS_LDC_OK;
}

}  // namespace

extern "C" {

const char *glass_ldc_last_error(void) { return g_err.c_str(); }

int glass_ldc_tdi_samples(const char *fileName, const char *dataName, int *Nsamples)
{
    QuietH5 q;
    if (!Nsamples) return fail(GLASS_LDC_EINVAL, "null Nsamples");
    Hid file(-1, H5Fclose), dset(-1, H5Dclose);
    hsize_t rows = 0;
    int rc = tdi_rows(fileName, dataName, file, dset, &rows);
    if (rc) return rc;
    *Nsamples = (int)rows;
    return GLASS_LDC_OK;
}

int glass_ldc_read_tdi(const char *fileName, const char *dataName, int Nsamples, double *X, double *Y,
                       double *Z, double *A, double *E, double *T, double *delta)
{
    QuietH5 q;
    Hid file(-1, H5Fclose), dset(-1, H5Dclose);
    hsize_t rows = 0;
    int rc = tdi_rows(fileName, dataName, file, dset, &rows);
    if (rc) return rc;
    rc = check_out(Nsamples, (int)rows, X, Y, Z, A, E, T, delta);
    if (rc) return rc;
    // memory type: the four members by name, as the reference's H5Tinsert
    // calls (the file may hold more members, in any order)
    Hid mtype(H5Tcreate(H5T_COMPOUND, sizeof(Row)), H5Tclose);
    H5Tinsert(mtype.id, "t", HOFFSET(Row, t), H5T_NATIVE_DOUBLE);
    H5Tinsert(mtype.id, "X", HOFFSET(Row, X), H5T_NATIVE_DOUBLE);
    H5Tinsert(mtype.id, "Y", HOFFSET(Row, Y), H5T_NATIVE_DOUBLE);
    H5Tinsert(mtype.id, "Z", HOFFSET(Row, Z), H5T_NATIVE_DOUBLE);
    std::vector<Row> s(rows);
    if (H5Dread(dset.id, mtype.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, s.data()) < 0)
        return fail(GLASS_LDC_EIO, std::string("cannot read ") + dataName);
    for (hsize_t i = 0; i < rows; i++)
    {
        X[i] = s[i].X;
        Y[i] = s[i].Y;
        Z[i] = s[i].Z;
        xyz_to_aet(s[i].X, s[i].Y, s[i].Z, &A[i], &E[i], &T[i]);
    }
    *delta = s[1].t - s[0].t;
    return GLASS_LDC_OK;
}

int glass_ldc_radler_samples(const char *fileName, int *Nsamples)
{
    QuietH5 q;
    if (!Nsamples) return fail(GLASS_LDC_EINVAL, "null Nsamples");
    Hid file(-1, H5Fclose), dset(-1, H5Dclose);
    hsize_t nrow = 0, ncol = 0;
    int rc = radler_shape(fileName, file, dset, &nrow, &ncol);
    if (rc) return rc;
    *Nsamples = (int)(nrow / 2);
    return GLASS_LDC_OK;
}

int glass_ldc_read_radler_tdi(const char *fileName, int Nsamples, double *X, double *Y, double *Z, double *A,
                              double *E, double *T, double *delta)
{
    QuietH5 q;
    Hid file(-1, H5Fclose), dset(-1, H5Dclose);
    hsize_t nrow = 0, ncol = 0;
    int rc = radler_shape(fileName, file, dset, &nrow, &ncol);
    if (rc) return rc;
    rc = check_out(Nsamples, (int)(nrow / 2), X, Y, Z, A, E, T, delta);
    if (rc) return rc;
    // only the rows that are kept (and row 1 for delta) are read: a
    // hyperslab of the first Nrow/2 rows, all columns
    Hid fspace(H5Dget_space(dset.id), H5Sclose);
    hsize_t start[2] = {0, 0}, count[2] = {nrow / 2, ncol};
    if (H5Sselect_hyperslab(fspace.id, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0)
        return fail(GLASS_LDC_EIO, "cannot select the rows");
    Hid mspace(H5Screate_simple(2, count, nullptr), H5Sclose);
    std::vector<double> d((size_t)count[0] * ncol);
    if (H5Dread(dset.id, H5T_NATIVE_DOUBLE, mspace.id, fspace.id, H5P_DEFAULT, d.data()) < 0)
        return fail(GLASS_LDC_EIO, std::string("cannot read ") + kRadler);
    for (hsize_t i = 0; i < count[0]; i++)
    {
        const double *r = d.data() + i * ncol;
        X[i] = r[1];
        Y[i] = r[2];
        Z[i] = r[3];
        xyz_to_aet(r[1], r[2], r[3], &A[i], &E[i], &T[i]);
    }
    *delta = d[ncol] - d[0];
    return GLASS_LDC_OK;
}

}  // extern "C"
