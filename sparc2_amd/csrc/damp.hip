// damp.hip -- FFT damping of the velocities (DAMPING = 1).
//
// damp_velocity (all_modules.f90:1802-1888) takes, per z plane, the
// unnormalised r2c transform of a velocity field, multiplies it by
// damping_rates(kx, ky) (INIT_DAMPING, initialize.f90:1276-1336; the rates
// carry the 1/(nx*ny) of the transform pair) and transforms back (c2r); the
// RHS subtracts the result (RHS.f90:166-172, 659-665).  Here the nz planes of
// a field go through one batched hipFFT plan each way, the multiply is one
// kernel over the half spectra, and the stage kernel subtracts the three
// terms at the point after its boundary zeroing.
#include <hipfft/hipfft.h>

#include <vector>

#include "sparc_dev.h"

namespace sparc {

struct DampPlan {
  hipfftHandle fwd = 0, inv = 0;
  bool f32 = false;
  int nx = 0, ny = 0, nz = 0;
  long fs = 0;
  void* spec = nullptr;   // half spectra of one field, (nx/2+1)*ny*nz complex
  void* rates = nullptr;  // (nx/2+1)*ny, element type
  void* term = nullptr;   // 3 slabs of fs: the terms of vx, vy, vz
};

// w(kx, ky, k) *= rates(kx, ky); grid.y = plane
template <typename C, typename T>
__global__ __launch_bounds__(256) void k_damp_mul(C* __restrict__ w, const T* __restrict__ rates, int m) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= m) return;
  const long i = (long)blockIdx.y * m + q;
  const T r = rates[q];
  C v = w[i];
  v.x *= r;
  v.y *= r;
  w[i] = v;
}

void damp_destroy(DampPlan* d) {
  if (!d) return;
  if (d->fwd) (void)hipfftDestroy(d->fwd);
  if (d->inv) (void)hipfftDestroy(d->inv);
  if (d->spec) (void)hipFree(d->spec);
  if (d->rates) (void)hipFree(d->rates);
  if (d->term) (void)hipFree(d->term);
  delete d;
}

// rates: (nx/2+1, ny) Fortran order = C (ny, nx/2+1), FP64 from the host.
// Returns nullptr on failure (and frees what it made).
DampPlan* damp_create(int nx, int ny, int nz, long fs, bool f32, const double* rates, hipStream_t st) {
  DampPlan* d = new DampPlan();
  d->f32 = f32;
  d->nx = nx;
  d->ny = ny;
  d->nz = nz;
  d->fs = fs;
  const int mx = nx / 2 + 1;
  const long m = (long)mx * ny;
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  bool ok = hipMalloc(&d->spec, 2 * esz * m * nz) == hipSuccess &&
            hipMalloc(&d->rates, esz * m) == hipSuccess &&
            hipMalloc(&d->term, esz * 3 * fs) == hipSuccess;
  if (ok) {
    if (f32) {
      std::vector<float> h(rates, rates + m);
      ok = hipMemcpy(d->rates, h.data(), esz * m, hipMemcpyHostToDevice) == hipSuccess;
    } else {
      ok = hipMemcpy(d->rates, rates, esz * m, hipMemcpyHostToDevice) == hipSuccess;
    }
  }
  // n, embeddings in C order (slowest first): a plane is ny rows of nx reals
  // or of nx/2+1 complex values; batch = the nz planes of one field
  int n[2] = {ny, nx}, real_embed[2] = {ny, nx}, cplx_embed[2] = {ny, mx};
  ok = ok && hipfftPlanMany(&d->fwd, 2, n, real_embed, 1, nx * ny, cplx_embed, 1, (int)m,
                            f32 ? HIPFFT_R2C : HIPFFT_D2Z, nz) == HIPFFT_SUCCESS;
  ok = ok && hipfftPlanMany(&d->inv, 2, n, cplx_embed, 1, (int)m, real_embed, 1, nx * ny,
                            f32 ? HIPFFT_C2R : HIPFFT_Z2D, nz) == HIPFFT_SUCCESS;
  ok = ok && hipfftSetStream(d->fwd, st) == HIPFFT_SUCCESS && hipfftSetStream(d->inv, st) == HIPFFT_SUCCESS;
  if (!ok) {
    damp_destroy(d);
    return nullptr;
  }
  return d;
}

const void* damp_term_slab(const DampPlan* d, int f) {
  return (const char*)d->term + (size_t)f * d->fs * (d->f32 ? sizeof(float) : sizeof(double));
}

// The terms of vx, vy, vz of the state buffer `u` (nv field slabs of fs).
// The r2c transforms read u out of place and leave it unchanged.
hipError_t damp_terms(DampPlan* d, const void* u, hipStream_t st) {
  const int mx = d->nx / 2 + 1;
  const int m = mx * d->ny;
  const size_t esz = d->f32 ? sizeof(float) : sizeof(double);
  dim3 grid((unsigned)((m + 255) / 256), (unsigned)d->nz);
  for (int f = 0; f < 3; ++f) {
    const char* in = (const char*)u + (size_t)(VX + f) * d->fs * esz;
    char* out = (char*)d->term + (size_t)f * d->fs * esz;
    hipfftResult r;
    if (d->f32) {
      r = hipfftExecR2C(d->fwd, (hipfftReal*)in, (hipfftComplex*)d->spec);
      if (r != HIPFFT_SUCCESS) return hipErrorLaunchFailure;
      hipLaunchKernelGGL((k_damp_mul<float2, float>), grid, dim3(256), 0, st, (float2*)d->spec,
                         (const float*)d->rates, m);
      r = hipfftExecC2R(d->inv, (hipfftComplex*)d->spec, (hipfftReal*)out);
    } else {
      r = hipfftExecD2Z(d->fwd, (hipfftDoubleReal*)in, (hipfftDoubleComplex*)d->spec);
      if (r != HIPFFT_SUCCESS) return hipErrorLaunchFailure;
      hipLaunchKernelGGL((k_damp_mul<double2, double>), grid, dim3(256), 0, st, (double2*)d->spec,
                         (const double*)d->rates, m);
      r = hipfftExecZ2D(d->inv, (hipfftDoubleComplex*)d->spec, (hipfftDoubleReal*)out);
    }
    if (r != HIPFFT_SUCCESS) return hipErrorLaunchFailure;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace sparc
