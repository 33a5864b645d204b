/* sparc_h5.h -- the reference's HDF5 file layouts, host side.
 *
 * SPARC2 writes and reads three kinds of file through the HDF5 Fortran API
 * (all_modules.f90): the background bkg.h5 (READ_IN_HDF5, 179-282; written by
 * the background output, 1425-1560), the full snapshot / restart
 * full_NNNN.h5 (WRITE_OUT_FULL_STATE, 1072-1423; READ_IN_FULL_STATE,
 * 741-1069) and the chunk output chunk_NNNN.h5 (WRITE_OUT_CHUNK, 1566-1799).
 * Each file carries the root attributes "xyzdims" (3 ints: nx, ny, nz of the
 * file's arrays) and "time" (1 double, hours of simulated time; bkg.h5 0) and
 * flat datasets of doubles.
 *
 * These primitives make those files from C: arrays are passed in Fortran
 * order (x fastest) with their Fortran extents; the file's dataspace lists
 * the extents in reverse, exactly as the Fortran API stores a Fortran array,
 * so files written here read in the reference and vice versa.  Every entry
 * returns 0 or a negative code; sparc_h5_last_error() gives the reason.
 * No counterpart of these functions exists in the reference's C interface
 * (it has none): they replace the Fortran HDF5 calls of the routines above.
 */
#ifndef SPARC_H5_H
#define SPARC_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define SPARC_H5_OK 0
#define SPARC_H5_E_INVALID -1
#define SPARC_H5_E_IO -2
#define SPARC_H5_E_SHAPE -3

/* Create (truncate) a file with the root attributes xyzdims and time
 * (h5fcreate_f + h5ltset_attribute_int_f/_double_f, all_modules.f90:1112-1120). */
int sparc_h5_create(const char* path, const int xyzdims[3], double time_hours);

/* Root attributes of an existing file (h5ltget_attribute_int_f,
 * all_modules.f90:203,765). */
int sparc_h5_attrs(const char* path, int xyzdims[3], double* time_hours);

/* 1 if the file has a dataset of that name, 0 if not, < 0 on error. */
int sparc_h5_has(const char* path, const char* name);

/* Add dataset `name` of rank 1..4 with Fortran extents dims[0..rank-1]
 * (x first), values data[i] * scale (h5dcreate_f + h5dwrite_f of
 * `array * scale`, e.g. all_modules.f90:1137-1141). */
int sparc_h5_write(const char* path, const char* name, int rank, const long* dims,
                   const double* data, double scale);

/* Read dataset `name`, which must have exactly the Fortran extents given,
 * into data as values / scale (h5dread_f, then the division of
 * all_modules.f90:786). */
int sparc_h5_read(const char* path, const char* name, int rank, const long* dims, double* data,
                  double scale);

const char* sparc_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SPARC_H5_H */
