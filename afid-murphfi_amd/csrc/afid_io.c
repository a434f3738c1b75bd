/* afid_io.c -- the solver's HDF5 files in the reference's layout (serial
 * HDF5 C API).  Declarations and the reference routines each entry point
 * stands for: include/afid_io.h.
 *
 * The reference writes these files collectively through MPI-IO, one
 * hyperslab per rank (HdfRoutines.F90:428-520).  With the whole field in one
 * host array a single contiguous write of the full extent gives the same file
 * content: dataset "var", H5T_NATIVE_DOUBLE, no chunking, no filters. */
#include "afid_io.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static _Thread_local char g_err[512];

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

/* HDF5 prints its error stack to stderr by default; this library reports
 * through return codes instead. */
static void quiet(void)
{
    static int done = 0;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
        done = 1;
    }
}

const char *afid_h5_last_error(void) { return g_err; }

int afid_h5_create_file(const char *path)
{
    quiet();
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(1, "cannot create %s", path);
    H5Fclose(f);
    return 0;
}

static hid_t open_rw(const char *path)
{
    hid_t f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) fail(1, "cannot open %s for writing", path);
    return f;
}

static hid_t open_ro(const char *path)
{
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) fail(1, "cannot open %s", path);
    return f;
}

int afid_h5_create_group(const char *path, const char *name)
{
    quiet();
    hid_t f = open_rw(path);
    if (f < 0) return 1;
    int rc = 0;
    if (H5Lexists(f, name, H5P_DEFAULT) <= 0) {
        hid_t g = H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (g < 0) rc = fail(2, "cannot create group %s in %s", name, path);
        else H5Gclose(g);
    }
    H5Fclose(f);
    return rc;
}

/* Write `rank`-D dataset `name` of type `type` into the open file `f`,
 * replacing a dataset of the same name.  `cdims` are C-order dims. */
static int write_dataset(hid_t f, const char *path, const char *name, hid_t type, int rank,
                         const hsize_t *cdims, const void *data)
{
    if (H5Lexists(f, name, H5P_DEFAULT) > 0 && H5Ldelete(f, name, H5P_DEFAULT) < 0)
        return fail(3, "cannot replace %s in %s", name, path);
    hid_t sp = H5Screate_simple(rank, cdims, NULL);
    if (sp < 0) return fail(3, "cannot make the dataspace of %s", name);
    hid_t d = H5Dcreate2(f, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = 0;
    if (d < 0) rc = fail(3, "cannot create dataset %s in %s", name, path);
    else {
        if (H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
            rc = fail(4, "cannot write dataset %s in %s", name, path);
        H5Dclose(d);
    }
    H5Sclose(sp);
    return rc;
}

/* Read dataset `name` of the open file `f` into `data` after checking that its
 * C-order dims are `cdims`. */
static int read_dataset(hid_t f, const char *path, const char *name, hid_t type, int rank,
                        const hsize_t *cdims, void *data)
{
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0) return fail(5, "no dataset %s in %s", name, path);
    hid_t sp = H5Dget_space(d);
    int rc = 0;
    hsize_t got[8] = {0};
    int r = sp < 0 ? -1 : H5Sget_simple_extent_ndims(sp);
    if (r != rank || r > 8) {
        rc = fail(6, "dataset %s in %s has rank %d, expected %d", name, path, r, rank);
    } else {
        H5Sget_simple_extent_dims(sp, got, NULL);
        for (int i = 0; i < rank; ++i)
            if (got[i] != cdims[i]) {
                /* report in Fortran order (fastest first), as the files are described */
                rc = fail(6, "dataset %s in %s: dim %d (Fortran order) is %llu, expected %llu", name, path,
                          rank - i, (unsigned long long)got[i], (unsigned long long)cdims[i]);
                break;
            }
        if (!rc && H5Dread(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
            rc = fail(7, "cannot read dataset %s in %s", name, path);
    }
    if (sp >= 0) H5Sclose(sp);
    H5Dclose(d);
    return rc;
}

int afid_h5_write_field3d(const char *path, const double *data, int64_t nx, int64_t ny, int64_t nz)
{
    quiet();
    if (!data || nx < 1 || ny < 1 || nz < 1)
        return fail(8, "afid_h5_write_field3d: bad arguments (nx %lld, ny %lld, nz %lld)",
                    (long long)nx, (long long)ny, (long long)nz);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(1, "cannot create %s", path);
    const hsize_t cd[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
    int rc = write_dataset(f, path, "var", H5T_NATIVE_DOUBLE, 3, cd, data);
    H5Fclose(f);
    return rc;
}

int afid_h5_read_field3d(const char *path, double *data, int64_t nx, int64_t ny, int64_t nz)
{
    quiet();
    if (!data || nx < 1 || ny < 1 || nz < 1)
        return fail(8, "afid_h5_read_field3d: bad arguments (nx %lld, ny %lld, nz %lld)",
                    (long long)nx, (long long)ny, (long long)nz);
    hid_t f = open_ro(path);
    if (f < 0) return 1;
    const hsize_t cd[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
    int rc = read_dataset(f, path, "var", H5T_NATIVE_DOUBLE, 3, cd, data);
    H5Fclose(f);
    return rc;
}

int afid_h5_dataset_dims(const char *path, const char *name, int64_t *dims, int maxrank)
{
    quiet();
    hid_t f = open_ro(path);
    if (f < 0) return -1;
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0) {
        H5Fclose(f);
        return -fail(5, "no dataset %s in %s", name, path);
    }
    hid_t sp = H5Dget_space(d);
    int r = sp < 0 ? -1 : H5Sget_simple_extent_ndims(sp);
    int rc = r;
    if (r < 0) rc = -fail(6, "cannot read the dataspace of %s in %s", name, path);
    else if (r > maxrank || r > 8) rc = -fail(6, "dataset %s in %s has rank %d > %d", name, path, r, maxrank);
    else {
        hsize_t got[8];
        H5Sget_simple_extent_dims(sp, got, NULL);
        for (int i = 0; i < r; ++i) dims[i] = (int64_t)got[r - 1 - i];
    }
    if (sp >= 0) H5Sclose(sp);
    H5Dclose(d);
    H5Fclose(f);
    return rc;
}

static int write_1d(const char *path, const char *name, hid_t type, const void *data, int64_t n)
{
    quiet();
    if (!data || n < 1) return fail(8, "bad arguments writing %s", name);
    hid_t f = open_rw(path);
    if (f < 0) return 1;
    const hsize_t cd[1] = {(hsize_t)n};
    int rc = write_dataset(f, path, name, type, 1, cd, data);
    H5Fclose(f);
    return rc;
}

static int read_1d(const char *path, const char *name, hid_t type, void *data, int64_t n)
{
    quiet();
    if (!data || n < 1) return fail(8, "bad arguments reading %s", name);
    hid_t f = open_ro(path);
    if (f < 0) return 1;
    const hsize_t cd[1] = {(hsize_t)n};
    int rc = read_dataset(f, path, name, type, 1, cd, data);
    H5Fclose(f);
    return rc;
}

int afid_h5_write_int_scalar(const char *path, const char *name, int32_t value)
{
    return write_1d(path, name, H5T_NATIVE_INT32, &value, 1);
}

int afid_h5_write_real_scalar(const char *path, const char *name, double value)
{
    return write_1d(path, name, H5T_NATIVE_DOUBLE, &value, 1);
}

int afid_h5_read_int_scalar(const char *path, const char *name, int32_t *value)
{
    return read_1d(path, name, H5T_NATIVE_INT32, value, 1);
}

int afid_h5_read_real_scalar(const char *path, const char *name, double *value)
{
    return read_1d(path, name, H5T_NATIVE_DOUBLE, value, 1);
}

int afid_h5_write_real1d(const char *path, const char *name, const double *data, int64_t n)
{
    return write_1d(path, name, H5T_NATIVE_DOUBLE, data, n);
}

int afid_h5_read_real1d(const char *path, const char *name, double *data, int64_t n)
{
    return read_1d(path, name, H5T_NATIVE_DOUBLE, data, n);
}
