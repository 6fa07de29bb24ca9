/*
 * mifdtd HDF5 files: field snapshots and the geometry file (libmifdtd_h5.so, host code).
 *
 * Snapshot files follow K7/node/save_fields.py:10-48: one file per saved step, holding
 * one dataset per field component, named by the component ('ex', 'ey', 'ez', 'hx',
 * 'hy', 'hz'), in the shape of the saved box (K7/util/common.py:155-179, axes of
 * length 1 dropped), C order, gzip-compressed ('compression="gzip"', h5py's level 4).
 * The geometry file follows K7/main.py:70-80 and K7/test/create_sample_h5.py:3-8: root
 * attributes nx, ny, nz (int64 scalars) and coeff_use (a fixed-length NULLPAD string).
 *
 * Every function returns 0 or a negative MIFDTD_H5_E* code and leaves the reason in
 * mifdtd_h5_last_error() (per calling thread). No function touches the GPU.
 */
#ifndef MIFDTD_H5_H
#define MIFDTD_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define MIFDTD_H5_EINVAL (-1)   /* an argument out of range */
#define MIFDTD_H5_EIO (-2)      /* the HDF5 library or zlib failed */
#define MIFDTD_H5_EFORMAT (-3)  /* the file lacks what the format requires */
#define MIFDTD_H5_MAX_RANK 8

const char *mifdtd_h5_last_error(void);

/* Write `ndset` datasets of one shape (rank 1..3, dims[rank]) into a new file `path`
 * (truncating it): names[n] <- data[n] (float32, or float64 with f64 = 1, C order, host
 * memory). Chunked, deflate level gzip_level (0..9), chunks compressed on nthreads
 * host threads (1..256). Replaces h5py's File(fpath, 'w') + create_dataset(str_f,
 * data=..., compression='gzip') of K7/node/save_fields.py:43-48. */
int mifdtd_h5_write_fields(const char *path, int ndset, const char *const *names, const void *const *data, int f64,
                           int rank, const long long *dims, int gzip_level, int nthreads);

/* The geometry file of K7/test/create_sample_h5.py:3-8 (attributes only). */
int mifdtd_h5_write_geometry(const char *path, long long nx, long long ny, long long nz, const char *coeff_use);

/* Read nx, ny, nz and coeff_use (into a buffer of `cap` bytes, NUL-terminated) from a
 * geometry file, as K7/main.py:70-75 reads them. Integer attributes of any integer type
 * and fixed- or variable-length strings are accepted. */
int mifdtd_h5_read_geometry(const char *path, long long *nx, long long *ny, long long *nz, char *coeff_use, int cap);

/* Shape (rank <= MIFDTD_H5_MAX_RANK, dims), element type (f64 = 1 for 8-byte floats) and
 * deflate level (-1 when the dataset has no deflate filter) of dataset `name`. */
int mifdtd_h5_dataset_info(const char *path, const char *name, int *rank, long long *dims, int *f64,
                           int *gzip_level);

/* Read dataset `name` whole into `out` (nbytes must equal its element count times 4,
 * or 8 with f64 = 1), converting to the requested float type. */
int mifdtd_h5_read_dataset(const char *path, const char *name, void *out, long long nbytes, int f64);

#ifdef __cplusplus
}
#endif
#endif /* MIFDTD_H5_H */
