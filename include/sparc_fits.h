/* sparc_fits.h -- the forcing cube of the PULSE = 0 driver, host side.
 *
 * readfits_xy_distribute (all_modules.f90:470-517) reads the forcing
 * function `forcingfunc`: a FITS primary image of nx x ny x nsteps doubles
 * (ftgpvd, null value -999), Fortran order.  These functions read and write
 * such a cube with CFITSIO.  Every entry returns 0 or the CFITSIO status
 * (> 0) or -1 for a bad argument; sparc_fits_last_error() gives the text.
 * No counterpart exists in the reference's C interface (it has none): they
 * replace its Fortran CFITSIO calls.
 */
#ifndef SPARC_FITS_H
#define SPARC_FITS_H

#ifdef __cplusplus
extern "C" {
#endif

/* Read the first nx*ny*nsteps values of the primary image (ftgpvd); the
 * image must have at least that many. */
int sparc_fits_read_cube(const char* path, long nx, long ny, long nsteps, double* out);

/* Write a cube as a new FITS file (BITPIX -64, NAXIS 3), replacing any file
 * of that name. */
int sparc_fits_write_cube(const char* path, long nx, long ny, long nsteps, const double* data);

const char* sparc_fits_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SPARC_FITS_H */
