/* afid_io.h -- HDF5 files of the solver: checkpoints (continua_*.h5), their
 * master file and the mean-profile file means.h5, in the reference's layout.
 *
 * Host-only library (libafid_io.so, C, serial HDF5 C API).  It moves host
 * arrays to and from files; afid_download / afid_upload of afid_gpu.h move
 * the same arrays to and from the device.  Every entry point returns 0 on
 * success and a non-zero code on failure, with the message in
 * afid_h5_last_error().  The serial HDF5 library underneath is not
 * thread-safe: call these functions from one thread at a time.
 *
 * Layout of a 3-D field file (HdfWriteRealHalo3D, HdfRoutines.F90:428-520):
 * one dataset "var", 64-bit IEEE doubles, Fortran dims (nx, ny, nz) = C dims
 * {nz, ny, nx}; the host array is [nz][ny][nx] with x fastest, which is
 * exactly what afid_download returns for lbound (1,1,1) and extent
 * (nx, nym, nzm) of a base field (nx = nxm + 1) or (nxr, nymr, nzmr) of a
 * refined one.
 */
#ifndef AFID_IO_H
#define AFID_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call of this library (thread-local). */
const char *afid_h5_last_error(void);

/* Create (truncate) an empty file: HdfCreateBlankFile, HdfRoutines.F90:11-22,
 * and the file part of HdfCreateMeansFile, CalcMeanProfiles.F90:289-300. */
int afid_h5_create_file(const char *path);

/* Create group `name` in an existing file (h5gcreate_f of HdfCreateMeansFile,
 * CalcMeanProfiles.F90:301-332; CreateSalinityH5Groups, salinity.F90:695-719;
 * CreatePhaseH5Groups, phasefield.F90:802-818).  An existing group is kept. */
int afid_h5_create_group(const char *path, const char *name);

/* Write a 3-D field file (creates or truncates `path`):
 * HdfWriteRealHalo3D (HdfRoutines.F90:428-520) and HdfWriteRealHalo3DR.
 * `data` is [nz][ny][nx], x fastest; nx, ny, nz are the Fortran dims. */
int afid_h5_write_field3d(const char *path, const double *data, int64_t nx, int64_t ny, int64_t nz);

/* Read a 3-D field file written as above: HdfReadContinua
 * (HdfReadContinua.F90:11-107).  Fails if the dataset's dims are not
 * (nx, ny, nz) (the reference reads whatever hyperslab it asks for). */
int afid_h5_read_field3d(const char *path, double *data, int64_t nx, int64_t ny, int64_t nz);

/* Rank and Fortran-order dims (fastest first) of dataset `name`; `dims` holds
 * `maxrank` entries.  Returns the rank (>= 0) or a negative code.  Used on
 * restart to compare a checkpoint's grid with the configured one
 * (ReadFlowInterp.F90:23-56 reads the grid of continua_master.h5 and
 * :84-89 compares it). */
int afid_h5_dataset_dims(const char *path, const char *name, int64_t *dims, int maxrank);

/* One-element 1-D datasets of continua_master.h5 in an existing file:
 * HdfSerialWriteIntScalar / HdfSerialWriteRealScalar (HdfRoutines.F90:25-96)
 * as WriteFlowField.F90:52-84 calls them; a dataset of that name is replaced.
 * The reads are HdfSerialReadIntScalar / HdfSerialReadRealScalar
 * (HdfRoutines.F90:257-309). */
int afid_h5_write_int_scalar(const char *path, const char *name, int32_t value);
int afid_h5_write_real_scalar(const char *path, const char *name, double value);
int afid_h5_read_int_scalar(const char *path, const char *name, int32_t *value);
int afid_h5_read_real_scalar(const char *path, const char *name, double *value);

/* 1-D dataset of n doubles in an existing file, e.g. "Tbar/00001" of
 * means.h5: HdfSerialWriteReal1D (HdfRoutines.F90:139-177), which deletes a
 * dataset of the same name first.  The group must exist.  The read is
 * HdfSerialReadReal1D (HdfRoutines.F90:311-336). */
int afid_h5_write_real1d(const char *path, const char *name, const double *data, int64_t n);
int afid_h5_read_real1d(const char *path, const char *name, double *data, int64_t n);

#ifdef __cplusplus
}
#endif

#endif
