// h5io.cpp -- the reference's HDF5 file layouts (bkg.h5, full_NNNN.h5,
// chunk_NNNN.h5) over the HDF5 C library; host code only, built into its
// own shared library so the HIP library does not depend on HDF5.
//
// A Fortran array a(n1, n2, ..., nr) written with the Fortran API appears in
// the file with the dataspace (nr, ..., n2, n1) and the same bytes; these
// functions take Fortran extents and reverse them, so their files and the
// reference's are the same.
#include "../../include/sparc_h5.h"

#include <hdf5.h>
#include <hdf5_hl.h>

#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// HDF5 prints its own error stack by default; the codes returned here carry
// the reason instead.
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

struct File {
  hid_t id = -1;
  File(const char* path, unsigned mode) { id = H5Fopen(path, mode, H5P_DEFAULT); }
  ~File() {
    if (id >= 0) H5Fclose(id);
  }
};

bool dims_ok(int rank, const long* dims) {
  if (rank < 1 || rank > 4 || !dims) return false;
  for (int r = 0; r < rank; ++r)
    if (dims[r] < 1) return false;
  return true;
}

}  // namespace

extern "C" {

const char* sparc_h5_last_error(void) { return g_err.c_str(); }

int sparc_h5_create(const char* path, const int xyzdims[3], double time_hours) {
  if (!path || !xyzdims) return fail(SPARC_H5_E_INVALID, "null argument");
  quiet();
  const hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail(SPARC_H5_E_IO, std::string("cannot create ") + path);
  int rc = SPARC_H5_OK;
  if (H5LTset_attribute_int(f, "/", "xyzdims", xyzdims, 3) < 0 ||
      H5LTset_attribute_double(f, "/", "time", &time_hours, 1) < 0)
    rc = fail(SPARC_H5_E_IO, "cannot write the root attributes");
  H5Fclose(f);
  return rc;
}

int sparc_h5_attrs(const char* path, int xyzdims[3], double* time_hours) {
  if (!path || !xyzdims || !time_hours) return fail(SPARC_H5_E_INVALID, "null argument");
  quiet();
  File f(path, H5F_ACC_RDONLY);
  if (f.id < 0) return fail(SPARC_H5_E_IO, std::string("cannot open ") + path);
  if (H5LTget_attribute_int(f.id, "/", "xyzdims", xyzdims) < 0)
    return fail(SPARC_H5_E_IO, "no xyzdims attribute");
  if (H5LTget_attribute_double(f.id, "/", "time", time_hours) < 0) *time_hours = 0.0;
  return SPARC_H5_OK;
}

int sparc_h5_has(const char* path, const char* name) {
  if (!path || !name) return fail(SPARC_H5_E_INVALID, "null argument");
  quiet();
  File f(path, H5F_ACC_RDONLY);
  if (f.id < 0) return fail(SPARC_H5_E_IO, std::string("cannot open ") + path);
  const htri_t e = H5Lexists(f.id, name, H5P_DEFAULT);
  return e < 0 ? fail(SPARC_H5_E_IO, "link query failed") : (e > 0 ? 1 : 0);
}

int sparc_h5_write(const char* path, const char* name, int rank, const long* dims,
                   const double* data, double scale) {
  if (!path || !name || !data || !dims_ok(rank, dims))
    return fail(SPARC_H5_E_INVALID, "bad argument");
  quiet();
  File f(path, H5F_ACC_RDWR);
  if (f.id < 0) return fail(SPARC_H5_E_IO, std::string("cannot open ") + path);
  hsize_t fd[4];
  size_t n = 1;
  for (int r = 0; r < rank; ++r) {
    fd[r] = (hsize_t)dims[rank - 1 - r];
    n *= (size_t)dims[r];
  }
  std::vector<double> tmp;
  const double* src = data;
  if (scale != 1.0) {
    tmp.resize(n);
    for (size_t q = 0; q < n; ++q) tmp[q] = data[q] * scale;
    src = tmp.data();
  }
  const hid_t sp = H5Screate_simple(rank, fd, nullptr);
  if (sp < 0) return fail(SPARC_H5_E_IO, "dataspace");
  const hid_t ds = H5Dcreate2(f.id, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  int rc = SPARC_H5_OK;
  if (ds < 0) {
    rc = fail(SPARC_H5_E_IO, std::string("cannot create dataset ") + name);
  } else {
    if (H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, src) < 0)
      rc = fail(SPARC_H5_E_IO, std::string("cannot write dataset ") + name);
    H5Dclose(ds);
  }
  H5Sclose(sp);
  return rc;
}

int sparc_h5_read(const char* path, const char* name, int rank, const long* dims, double* data,
                  double scale) {
  if (!path || !name || !data || !dims_ok(rank, dims) || scale == 0.0)
    return fail(SPARC_H5_E_INVALID, "bad argument");
  quiet();
  File f(path, H5F_ACC_RDONLY);
  if (f.id < 0) return fail(SPARC_H5_E_IO, std::string("cannot open ") + path);
  const hid_t ds = H5Dopen2(f.id, name, H5P_DEFAULT);
  if (ds < 0) return fail(SPARC_H5_E_IO, std::string("no dataset ") + name);
  const hid_t sp = H5Dget_space(ds);
  int rc = SPARC_H5_OK;
  hsize_t fd[8] = {};
  const int fr = sp < 0 ? -1 : H5Sget_simple_extent_ndims(sp);
  if (fr != rank || H5Sget_simple_extent_dims(sp, fd, nullptr) < 0) {
    rc = fail(SPARC_H5_E_SHAPE, std::string("dataset ") + name + " has another rank");
  } else {
    size_t n = 1;
    for (int r = 0; r < rank; ++r) {
      if (fd[r] != (hsize_t)dims[rank - 1 - r]) {
        rc = fail(SPARC_H5_E_SHAPE, std::string("dataset ") + name + " has other extents");
        break;
      }
      n *= (size_t)dims[r];
    }
    if (rc == SPARC_H5_OK) {
      if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) {
        rc = fail(SPARC_H5_E_IO, std::string("cannot read dataset ") + name);
      } else if (scale != 1.0) {
        for (size_t q = 0; q < n; ++q) data[q] = data[q] / scale;
      }
    }
  }
  if (sp >= 0) H5Sclose(sp);
  H5Dclose(ds);
  return rc;
}

}  // extern "C"
