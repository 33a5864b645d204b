// fitsio.cpp -- the forcing cube of the PULSE = 0 driver over CFITSIO; host
// code only, its own shared library.
#include "../../include/sparc_fits.h"

#include <fitsio.h>

#include <string>

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

std::string text(int status) {
  char buf[FLEN_STATUS] = {};
  fits_get_errstatus(status, buf);
  return buf;
}
}  // namespace

extern "C" {

const char* sparc_fits_last_error(void) { return g_err.c_str(); }

int sparc_fits_read_cube(const char* path, long nx, long ny, long nsteps, double* out) {
  if (!path || !out || nx < 1 || ny < 1 || nsteps < 1) return fail(-1, "bad argument");
  fitsfile* f = nullptr;
  int status = 0;
  if (fits_open_file(&f, path, READONLY, &status)) return fail(status, std::string(path) + ": " + text(status));
  int naxis = 0;
  long naxes[3] = {1, 1, 1};
  fits_get_img_dim(f, &naxis, &status);
  fits_get_img_size(f, 3, naxes, &status);
  const long have = naxis < 1 ? 0 : naxes[0] * (naxis > 1 ? naxes[1] : 1) * (naxis > 2 ? naxes[2] : 1);
  const long want = nx * ny * nsteps;
  if (!status && have < want) {
    int s2 = 0;
    fits_close_file(f, &s2);
    return fail(-1, std::string(path) + ": the image holds fewer values than nx*ny*nsteps");
  }
  double nullval = -999.0;  // as readfits_xy_distribute
  int anynull = 0;
  fits_read_img(f, TDOUBLE, 1, want, &nullval, out, &anynull, &status);
  int s2 = 0;
  fits_close_file(f, &s2);
  return status ? fail(status, std::string(path) + ": " + text(status)) : 0;
}

int sparc_fits_write_cube(const char* path, long nx, long ny, long nsteps, const double* data) {
  if (!path || !data || nx < 1 || ny < 1 || nsteps < 1) return fail(-1, "bad argument");
  fitsfile* f = nullptr;
  int status = 0;
  const std::string name = std::string("!") + path;  // '!' replaces an existing file
  if (fits_create_file(&f, name.c_str(), &status)) return fail(status, std::string(path) + ": " + text(status));
  long naxes[3] = {nx, ny, nsteps};
  fits_create_img(f, DOUBLE_IMG, 3, naxes, &status);
  fits_write_img(f, TDOUBLE, 1, nx * ny * nsteps, const_cast<double*>(data), &status);
  int s2 = 0;
  fits_close_file(f, &s2);
  return status ? fail(status, std::string(path) + ": " + text(status)) : 0;
}

}  // extern "C"
