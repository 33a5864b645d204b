// ops.hip -- digital filters, standalone derivative operators and the
// diagnostic reduction.
//
// Filters (FILTER_VARS, derivatives.f90:659-703): one pass per direction,
// y then x then z as the reference orders them.  Each pass reads the field,
// applies the edge smoothing on the fly (Jacobi form: both reference edge
// statements read the unsmoothed values), filters, and writes a second
// buffer; the host swaps the buffer roles afterwards.
//
// Standalone derivatives (ddx/ddy/ddz, derivatives.f90:10-116,226-341,
// 451-562) evaluate the 11-point sums in the reference's index order; they
// serve the background set-up (initialize.f90:993-1016,1219-1228) and the
// drop-in operator entry points.
#include <algorithm>

#include "sparc_dev.h"

namespace sparc {

// ---------------------------------------------------------------------------
// Filters
// ---------------------------------------------------------------------------
// A block filters one segment of one x row (j, k) of one field.  Rows are
// numbered so that the rows a point's filter reads (same x, the filtered axis
// +-G) are consecutive (y filter: j fastest; z filter: k fastest), and the
// block ids are remapped so each XCD takes a contiguous range of rows: the
// 2G+1 reads of a point then hit the XCD's L2 and HBM sees about one read per
// value.  The row decode is a handful of uniform 32-bit divisions per block.
template <typename T, int AXIS, int G>
__global__ __launch_bounds__(256) void k_filter(const T* __restrict__ in, T* __restrict__ out,
                                                int nx, int ny, int nz, long fs, int nseg,
                                                int periodic, T s, int bare,
                                                const int* __restrict__ halt) {
  // a run stopped by its failure guard keeps its state: the pass copies
  // (the host swaps the buffer roles after every pass)
  const bool stop = halted(halt);
  const int nblk = gridDim.x, b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3, base = nblk >> 3, rem = nblk & 7;
  const int lg = xcd * base + (xcd < rem ? xcd : rem) + loc;
  const int seg = lg % nseg;
  int row = lg / nseg, j, k;
  if (AXIS == 2) {
    k = row % nz;
    row /= nz;
    j = row % ny;
    row /= ny;
  } else {
    j = row % ny;
    row /= ny;
    k = row % nz;
    row /= nz;
  }
  const int f = row;
  const int i = seg * (int)blockDim.x + (int)threadIdx.x;
  if (i >= nx) return;
  const long idx = (long)i + (long)nx * (j + (long)ny * k);
  const int n = AXIS == 0 ? nx : AXIS == 1 ? ny : nz;
  const int pos = AXIS == 0 ? i : AXIS == 1 ? j : k;
  const long stride = AXIS == 0 ? 1 : AXIS == 1 ? (long)nx : (long)nx * ny;
  const T* line = in + (long)f * fs + (idx - (long)pos * stride);
  if (stop) {
    out[(long)f * fs + idx] = line[(long)pos * stride];
    return;
  }
  const bool wrapped = (AXIS < 2) && periodic;
  // edge smoothing: derivatives.f90:671-675 (y), 680-684 (x), 690-693 (z)
  // (bare = the filtx/filty/filtz operators alone: no smoothing, result filt)
  auto sm = [&](int r) -> T {
    if (!bare && !wrapped && ((r >= 1 && r <= G - 1) || (r >= n - G && r <= n - 2)))
      return (line[(long)(r - 1) * stride] + line[(long)(r + 1) * stride]) / T(2);
    return line[(long)r * stride];
  };
  T filt;
  if (wrapped) {
    T acc = 0;
#pragma unroll
    for (int m = -G; m <= G; ++m) acc += line[(long)wrap(pos + m, n) * stride] * filt_w<T>(G, m);
    filt = acc;
  } else if (pos < G || pos >= n - G) {
    // x/y: copy (g = 5) or zero (g = 3); z: copy for both widths
    // (derivatives.f90:190-192,215-217,414-416,439-441,624,635)
    filt = (AXIS == 2 || G == 5) ? sm(pos) : T(0);
  } else {
    T acc = 0;
#pragma unroll
    for (int m = -G; m <= G; ++m) acc += sm(pos + m) * filt_w<T>(G, m);
    filt = acc;
  }
  // g = 5: a = filt; g = 3: a = a - s*filt (derivatives.f90:677-678,686-687,695-696)
  out[(long)f * fs + idx] = (G == 5 || bare) ? filt : sm(pos) - s * filt;
}

// FILTER_VARS's y and x passes in one kernel (derivatives.f90:669-687):
// y filter, then x filter of its result, each with its edge smoothing, one
// read and one write per value.  A wave owns 64 consecutive x points (one per
// lane, so loads and stores are coalesced) of one row segment and marches up
// a chunk of KY rows with a register queue of the raw rows j-H .. j+H, loaded
// PD rows ahead: the y stencil is read from registers.  The x stencil of the
// y-filtered row is read from the neighbouring lanes (wavefront shuffles), so
// the inner 64 - 2H lanes write results; the outer H lanes on each side only
// supply the stencil (H = G + 1: the filter's reach plus the one-point reach
// of the smoothing).  The per-point arithmetic is that of the separate passes,
// in the same order, so both give the same bits.
constexpr int FXY_KY = 64;
constexpr int FXY_PD = 4;

// S: the stored type -- T, or uint16_t for the FP32 form's bfloat16
// compensation (filtered in FP32 with the same operator, stored rounded).
template <typename S> __device__ __forceinline__ auto from_s(S v) {
  if constexpr (sizeof(S) == 2) return bf2f(v);
  else return v;
}
template <typename T, typename S> __device__ __forceinline__ S to_s(T v) {
  if constexpr (sizeof(S) == 2) return f2bf(v);
  else return v;
}

template <typename T, int G, typename S = T>
__global__ __launch_bounds__(256) void k_filter_xy(const S* __restrict__ in, S* __restrict__ out,
                                                   int nx, int ny, int nz, long fs, int nseg, int nych,
                                                   long nwave, int periodic, T s,
                                                   const int* __restrict__ halt) {
  constexpr int H = G + 1, OW = 64 - 2 * H, Q = 2 * H + 1;
  const int nblk = gridDim.x, b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3, base = nblk >> 3, rem = nblk & 7;
  const long lg = xcd * base + (xcd < rem ? xcd : rem) + loc;
  const long wv = lg * 4 + (threadIdx.x >> 6);
  if (wv >= nwave) return;  // whole waves only: the shuffles below see full waves
  const int lane = threadIdx.x & 63;
  const int seg = (int)(wv % nseg);
  long rest = wv / nseg;
  const int ych = (int)(rest % nych);
  rest /= nych;
  const int k = (int)(rest % nz);
  const int f = (int)(rest / nz);
  const int x0 = seg * OW;
  const int xl = x0 - H + lane;  // this lane's x (unwrapped)
  const bool xin = periodic || (xl >= 0 && xl < nx);
  const int xg = periodic ? wrap(xl, nx) : xl;
  const bool writer = lane >= H && lane < 64 - H && xl < nx;
  const int j0 = ych * FXY_KY, j1 = min(j0 + FXY_KY, ny);
  const long plane = (long)f * fs + (long)nx * ny * k;
  const S* col = in + plane + (xin ? xg : 0);
  S* ocol = out + plane + xl;
  if (halted(halt)) {  // a run stopped by its failure guard keeps its state
    if (writer)
      for (int j = j0; j < j1; ++j) ocol[(long)j * nx] = col[(long)j * nx];
    return;
  }
  auto ld = [&](int r) -> T {
    if (periodic) return (T)from_s(col[(long)wrap(r, ny) * nx]);
    return (xin && r >= 0 && r < ny) ? (T)from_s(col[(long)r * nx]) : T(0);
  };
  T q[Q], pf[FXY_PD];
#pragma unroll
  for (int m = 0; m < Q - 1; ++m) q[m] = ld(j0 - H + m);
#pragma unroll
  for (int d = 0; d < FXY_PD; ++d) pf[d] = ld(j0 + H + d);
  const bool xedge = !periodic && ((xl >= 1 && xl <= G - 1) || (xl >= nx - G && xl <= nx - 2));
  const bool xcopy = !periodic && (xl < G || xl >= nx - G);
  for (int j = j0; j < j1; ++j) {
    q[Q - 1] = pf[0];
#pragma unroll
    for (int d = 0; d < FXY_PD - 1; ++d) pf[d] = pf[d + 1];
    pf[FXY_PD - 1] = ld(j + H + FXY_PD);
    // y pass: smoothed raw value of row j+d (derivatives.f90:671-675)
    auto smy = [&](int d) -> T {
      const int r = j + d;
      if (!periodic && ((r >= 1 && r <= G - 1) || (r >= ny - G && r <= ny - 2)))
        return (q[d + H - 1] + q[d + H + 1]) / T(2);
      return q[d + H];
    };
    T yf, y;
    if (periodic) {
      T acc = 0;
#pragma unroll
      for (int m = -G; m <= G; ++m) acc += q[m + H] * filt_w<T>(G, m);
      yf = acc;
      y = G == 5 ? yf : q[H] - s * yf;
    } else if (j < G || j >= ny - G) {
      yf = G == 5 ? smy(0) : T(0);  // copy (g = 5) or zero (g = 3)
      y = G == 5 ? yf : smy(0) - s * yf;
    } else {
      T acc = 0;
#pragma unroll
      for (int m = -G; m <= G; ++m) acc += smy(m) * filt_w<T>(G, m);
      yf = acc;
      y = G == 5 ? yf : smy(0) - s * yf;
    }
    // x pass on the y-filtered row across the lanes (derivatives.f90:680-687)
    const T ym = __shfl_up(y, 1), yp = __shfl_down(y, 1);
    const T sx = xedge ? (ym + yp) / T(2) : y;
    T acc = 0;
#pragma unroll
    for (int m = -G; m <= G; ++m) {
      const T v = m < 0 ? __shfl_up(sx, -m) : (m > 0 ? __shfl_down(sx, m) : sx);
      acc += v * filt_w<T>(G, m);
    }
    T xf = acc;
    if (xcopy) xf = G == 5 ? sx : T(0);
    if (writer) ocol[(long)j * nx] = to_s<T, S>(G == 5 ? xf : sx - s * xf);
#pragma unroll
    for (int m = 0; m < Q - 1; ++m) q[m] = q[m + 1];
  }
}

// z filter: one column per thread marching up a chunk of planes with a
// register queue of the raw values (planes pos-G-1 .. pos+G+1: the filter
// reads smoothed values at pos-G .. pos+G, and a smoothed value reads its two
// neighbours), so each value is loaded once; loads run ZPD planes ahead.
// Same per-point arithmetic, in the same order, as k_filter along z.
constexpr int ZPD = 4;

template <typename T, int G, typename S = T>
__global__ __launch_bounds__(256) void k_filter_z(const S* in, S* out,
                                                  int nx, int ny, int nz, long fs, int ncb, int kc,
                                                  T s, int bare, const int* __restrict__ halt) {
  const bool stop = halted(halt);  // guard stop: copy, as in k_filter
  const int nch = (nz + kc - 1) / kc;
  const int lg = blockIdx.x;
  const int cb = lg % ncb, rest = lg / ncb;
  const int chunk = rest % nch, f = rest / nch;
  const long nxy = (long)nx * ny;
  const long col = (long)cb * 256 + threadIdx.x;
  if (col >= nxy) return;
  const S* line = in + (long)f * fs + col;
  S* o = out + (long)f * fs + col;
  const int k0 = chunk * kc, k1 = min(k0 + kc, nz);
  if (stop) {
    for (int pos = k0; pos < k1; ++pos) o[(long)pos * nxy] = line[(long)pos * nxy];
    return;
  }
  auto ld = [&](int r) -> T { return (r >= 0 && r < nz) ? (T)from_s(line[(long)r * nxy]) : T(0); };
  constexpr int Q = 2 * G + 3;
  T q[Q], pf[ZPD];
#pragma unroll
  for (int m = 0; m < Q - 1; ++m) q[m] = ld(k0 - G - 1 + m);
#pragma unroll
  for (int d = 0; d < ZPD; ++d) pf[d] = ld(k0 + G + 1 + d);
  for (int pos = k0; pos < k1; ++pos) {
    q[Q - 1] = pf[0];
#pragma unroll
    for (int d = 0; d < ZPD - 1; ++d) pf[d] = pf[d + 1];
    pf[ZPD - 1] = ld(pos + G + 1 + ZPD);
    // smoothed value of plane pos+d (derivatives.f90:690-693)
    auto sm = [&](int d) -> T {
      const int r = pos + d;
      if (!bare && ((r >= 1 && r <= G - 1) || (r >= nz - G && r <= nz - 2)))
        return (q[d + G] + q[d + G + 2]) / T(2);
      return q[d + G + 1];
    };
    T filt;
    if (pos < G || pos >= nz - G) {
      filt = sm(0);  // z: copy for both widths (derivatives.f90:624,635)
    } else {
      T acc = 0;
#pragma unroll
      for (int m = -G; m <= G; ++m) acc += sm(m) * filt_w<T>(G, m);
      filt = acc;
    }
    o[(long)pos * nxy] = to_s<T, S>((G == 5 || bare) ? filt : sm(0) - s * filt);
#pragma unroll
    for (int m = 0; m < Q - 1; ++m) q[m] = q[m + 1];
  }
}

// z pass (FILTER_VARS, derivatives.f90:690-696) on stored type T or, cbf,
// on bfloat16 compensation terms.  inplace: in may equal out -- one chunk per
// column, whose register queue holds every raw value it still needs before
// that plane is overwritten.
template <typename T>
hipError_t launch_filter_zc(int g, const void* in, void* out, int nx, int ny, int nz, long fs,
                            int nfield, T s, int bare, hipStream_t st, const int* halt, bool inplace,
                            bool cbf) {
  if (cbf && sizeof(T) != 4) return hipErrorInvalidValue;
  // column blocks x z chunks x fields: >= ~2048 blocks, chunks >= 16 deep
  const long nxy = (long)nx * ny;
  const int ncb = (int)((nxy + 255) / 256);
  const long want = (2048 + (long)ncb * nfield - 1) / ((long)ncb * nfield);
  const int nch = inplace ? 1 : (int)std::max(1L, std::min(want, (long)nz / 16));
  const int kc = (nz + nch - 1) / nch;
  const long nb = (long)ncb * ((nz + kc - 1) / kc) * nfield;
  if (nb >= (1L << 31)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)nb);
#define SPARC_FZ(GG, S) \
  hipLaunchKernelGGL((k_filter_z<T, GG, S>), grid, dim3(256), 0, st, (const S*)in, (S*)out, nx, ny, nz, fs, ncb, kc, s, bare, halt)
  if (cbf) {
    if (g == 5) SPARC_FZ(5, uint16_t);
    else SPARC_FZ(3, uint16_t);
  } else {
    if (g == 5) SPARC_FZ(5, T);
    else SPARC_FZ(3, T);
  }
#undef SPARC_FZ
  return hipGetLastError();
}
template hipError_t launch_filter_zc<double>(int, const void*, void*, int, int, int, long, int, double,
                                             int, hipStream_t, const int*, bool, bool);
template hipError_t launch_filter_zc<float>(int, const void*, void*, int, int, int, long, int, float,
                                            int, hipStream_t, const int*, bool, bool);

// cbf (FP32 only): the arrays hold bfloat16 compensation terms.
template <typename T>
hipError_t launch_filter_xy(int g, const void* in, void* out, int nx, int ny, int nz, long fs,
                            int nfield, int periodic, T s, hipStream_t st, const int* halt, bool cbf) {
  const int ow = 64 - 2 * (g + 1);
  const int nseg = (nx + ow - 1) / ow, nych = (ny + FXY_KY - 1) / FXY_KY;
  const long nwave = (long)nseg * nych * nz * nfield;
  const long nb = (nwave + 3) / 4;
  if (nb >= (1L << 31)) return hipErrorInvalidValue;
  if (cbf && sizeof(T) != 4) return hipErrorInvalidValue;
  const dim3 grid((unsigned)nb);
#define SPARC_FXY(GG, S)                                                                              \
  hipLaunchKernelGGL((k_filter_xy<T, GG, S>), grid, dim3(256), 0, st, (const S*)in, (S*)out, nx, ny, \
                     nz, fs, nseg, nych, nwave, periodic, s, halt)
  if (cbf) {
    if (g == 5) SPARC_FXY(5, uint16_t);
    else SPARC_FXY(3, uint16_t);
  } else {
    if (g == 5) SPARC_FXY(5, T);
    else SPARC_FXY(3, T);
  }
#undef SPARC_FXY
  return hipGetLastError();
}
template hipError_t launch_filter_xy<double>(int, const void*, void*, int, int, int, long, int, int,
                                             double, hipStream_t, const int*, bool);
template hipError_t launch_filter_xy<float>(int, const void*, void*, int, int, int, long, int, int,
                                            float, hipStream_t, const int*, bool);

template <typename T>
hipError_t launch_filter_zc(int g, const void* in, void* out, int nx, int ny, int nz, long fs,
                            int nfield, T s, int bare, hipStream_t st, const int* halt, bool inplace,
                            bool cbf);

// One pass along x, y or z.  bare: the filtx/filty/filtz operators alone
// (no edge smoothing); inplace: z only, see launch_filter_zc.
template <typename T>
hipError_t launch_filter(int axis, int g, const T* in, T* out, int nx, int ny, int nz, long fs,
                         int nfield, int periodic, T s, int bare, hipStream_t st, const int* halt,
                         bool inplace) {
  if (inplace && axis != 2) return hipErrorInvalidValue;
  if (axis == 2) return launch_filter_zc<T>(g, in, out, nx, ny, nz, fs, nfield, s, bare, st, halt, inplace, false);
  // one block per x-row segment: 256 threads, or the row rounded up to
  // whole waves when it is shorter
  const int nt = nx >= 256 ? 256 : (nx + 63) / 64 * 64;
  const int nseg = (nx + nt - 1) / nt;
  const long nb = (long)nseg * ny * nz * nfield;
  if (nb >= (1L << 31)) return hipErrorInvalidValue;
  dim3 grid((unsigned)nb);
#define SPARC_F(AX, GG) \
  hipLaunchKernelGGL((k_filter<T, AX, GG>), grid, dim3(nt), 0, st, in, out, nx, ny, nz, fs, nseg, periodic, s, bare, halt)
  if (g == 5) {
    if (axis == 0) SPARC_F(0, 5);
    else if (axis == 1) SPARC_F(1, 5);
    else SPARC_F(2, 5);
  } else {
    if (axis == 0) SPARC_F(0, 3);
    else if (axis == 1) SPARC_F(1, 3);
    else SPARC_F(2, 3);
  }
#undef SPARC_F
  return hipGetLastError();
}
template hipError_t launch_filter<double>(int, int, const double*, double*, int, int, int, long,
                                          int, int, double, int, hipStream_t, const int*, bool);
template hipError_t launch_filter<float>(int, int, const float*, float*, int, int, int, long, int,
                                         int, float, int, hipStream_t, const int*, bool);

// ---------------------------------------------------------------------------
// Standalone derivatives (FP64, the reference's summation order)
// ---------------------------------------------------------------------------
template <int AXIS>
__global__ __launch_bounds__(256) void k_deriv(const double* __restrict__ in, double* __restrict__ out,
                                               int nx, int ny, int nz, double sxy,
                                               const double* __restrict__ stretch, int bc) {
  const long N = (long)nx * ny * nz;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int i = (int)(idx % nx);
  const int j = (int)((idx / nx) % ny);
  const int k = (int)(idx / ((long)nx * ny));
  const int n = AXIS == 0 ? nx : AXIS == 1 ? ny : nz;
  const int pos = AXIS == 0 ? i : AXIS == 1 ? j : k;
  const long stride = AXIS == 0 ? 1 : AXIS == 1 ? (long)nx : (long)nx * ny;
  const double* line = in + (idx - (long)pos * stride);
  const double s = AXIS == 2 ? stretch[k] : sxy;
  constexpr double cc[11] = {-0.002484594688, +0.020779405824, -0.090320001280, +0.286511173973,
                             -0.872756993962, 0.0, +0.872756993962, -0.286511173973,
                             +0.090320001280, -0.020779405824, +0.002484594688};
  double acc = 0.0;
  if (AXIS < 2 && bc == 5) {
    for (int m = 0; m < 11; ++m) acc += line[(long)wrap(pos + m - 5, n) * stride] * cc[m];
    out[idx] = s * acc;
  } else if (pos >= n - 5) {
    const int r = n - 1 - pos;
    for (int m = 0; m < 11; ++m) acc += line[(long)(n - 1 - m) * stride] * DPB_D[r][m];
    out[idx] = -s * acc;
  } else if (AXIS < 2 && pos < 5) {
    for (int m = 0; m < 11; ++m) acc += line[(long)m * stride] * DPB_D[pos][m];
    out[idx] = s * acc;
  } else {
    // interior; for z this includes the bottom rows over zero ghost cells
    for (int m = 0; m < 11; ++m) {
      const int p = pos + m - 5;
      acc += (p >= 0 ? line[(long)p * stride] : 0.0) * cc[m];
    }
    out[idx] = s * acc;
  }
}

hipError_t launch_deriv(int axis, const double* in, double* out, int nx, int ny, int nz, double sxy,
                        const double* stretch, int bc, hipStream_t st) {
  const long N = (long)nx * ny * nz;
  dim3 grid((unsigned)((N + 255) / 256));
  if (axis == 0)
    hipLaunchKernelGGL(k_deriv<0>, grid, dim3(256), 0, st, in, out, nx, ny, nz, sxy, stretch, bc);
  else if (axis == 1)
    hipLaunchKernelGGL(k_deriv<1>, grid, dim3(256), 0, st, in, out, nx, ny, nz, sxy, stretch, bc);
  else
    hipLaunchKernelGGL(k_deriv<2>, grid, dim3(256), 0, st, in, out, nx, ny, nz, sxy, stretch, bc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Diagnostic reduction of the per-block partials of the last RK stage.
// One block, fixed order -> deterministic.  out = {sum m^2, min rho, min p}.
// Thread 0 then counts the step and evaluates the failure guard.
// ---------------------------------------------------------------------------
__device__ void reduce_partials(const double* __restrict__ partial, int nblk, double* __restrict__ out,
                                const GuardArgs& G) {
  __shared__ double s0[256], s1[256], s2[256];
  const int t = threadIdx.x;
  double a = 0.0, r = 1.0e300, p = 1.0e300;
  for (int b = t; b < nblk; b += 256) {
    a += partial[3 * (long)b];
    r = fmin(r, partial[3 * (long)b + 1]);
    p = fmin(p, partial[3 * (long)b + 2]);
  }
  s0[t] = a;
  s1[t] = r;
  s2[t] = p;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      s0[t] += s0[t + o];
      s1[t] = fmin(s1[t], s1[t + o]);
      s2[t] = fmin(s2[t], s2[t + o]);
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = s0[0];
    out[1] = s1[0];
    out[2] = s2[0];
    if (G.guard) {
      int* g = G.guard;
      const int count = g[G_COUNT] + 1;
      g[G_COUNT] = count;
      if (g[G_MODE] != GUARD_OFF && g[G_TRIPPED] == 0) {
        // driver.f90:175-183; a NaN norm or minimum also counts as failed
        const double ke = sqrt(s0[0] / G.npts) * G.ke_scale;
        if (!(ke <= 1.0e3) || !(s1[0] > 0.0) || !(s2[0] > 0.0)) {
          G.gdiag[0] = s0[0];
          G.gdiag[1] = s1[0];
          G.gdiag[2] = s2[0];
          g[G_AT] = count;
          g[G_TRIPPED] = 1;
          if (g[G_MODE] == GUARD_HALT) g[G_HALT] = 1;
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_diag_reduce(const double* __restrict__ partial, int nblk,
                                                     double* __restrict__ out, GuardArgs G) {
  if (halted(G.guard)) {  // the step is counted, nothing else
    if (threadIdx.x == 0) G.guard[G_COUNT] += 1;
    return;
  }
  reduce_partials(partial, nblk, out, G);
}

// Batched form: block r reduces run r's partials into its own out[3].
__global__ __launch_bounds__(256) void k_diag_reduce_batch(const DiagJob* __restrict__ jobs) {
  const DiagJob& J = jobs[blockIdx.x];
  if (halted(J.g.guard)) {  // the step is counted, nothing else
    if (threadIdx.x == 0) J.g.guard[G_COUNT] += 1;
    return;
  }
  const GuardArgs G = J.g;
  reduce_partials(J.partial, J.nblk, J.out, G);
}

hipError_t launch_diag_reduce_batch(const void* jobs, int njobs, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_reduce_batch, dim3(njobs), dim3(256), 0, st, (const DiagJob*)jobs);
  return hipGetLastError();
}

// Device step counter of the batched form: +1 after each step, 0 on reset.
__global__ void k_tick(int* c, int reset) {
  if (threadIdx.x == 0) c[0] = reset ? 0 : c[0] + 1;
}

hipError_t launch_tick(int* c, int reset, hipStream_t st) {
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, c, reset);
  return hipGetLastError();
}

// Plain copy (the in-place filter of the batched form copies its result back).
template <typename T>
__global__ __launch_bounds__(256) void k_copy(const T* __restrict__ in, T* __restrict__ out, long n,
                                              const int* __restrict__ halt) {
  if (halted(halt)) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

hipError_t launch_copy(const void* in, void* out, long n, bool f32, hipStream_t st, const int* halt) {
  dim3 grid((unsigned)((n + 255) / 256));
  if (f32)
    hipLaunchKernelGGL(k_copy<float>, grid, dim3(256), 0, st, (const float*)in, (float*)out, n, halt);
  else
    hipLaunchKernelGGL(k_copy<double>, grid, dim3(256), 0, st, (const double*)in, (double*)out, n, halt);
  return hipGetLastError();
}

hipError_t launch_diag_reduce(const double* partial, int nblk, double* out, const GuardArgs& g,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_diag_reduce, dim3(1), dim3(256), 0, st, partial, nblk, out, g);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Small helpers used by the host layer
// ---------------------------------------------------------------------------
__global__ void k_fill_zero_d(double* p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
__global__ void k_fill_zero_f(float* p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

hipError_t launch_zero(void* p, long n, bool f32, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  dim3 grid((unsigned)((n + 255) / 256));
  if (f32)
    hipLaunchKernelGGL(k_fill_zero_f, grid, dim3(256), 0, st, (float*)p, n);
  else
    hipLaunchKernelGGL(k_fill_zero_d, grid, dim3(256), 0, st, (double*)p, n);
  return hipGetLastError();
}

}  // namespace sparc

namespace sparc {
// ---------------------------------------------------------------------------
// Alfven limiter search (initialize.f90:966-979): one evaluation of
// va(T) = max_p sqrt(dimb^2 T/(T+red_p) b2_p / (4 pi dimrho rho0_p)) per
// launch, in the reference's operation order; NaN terms are skipped like
// Fortran MAXVAL.  Blocks write partial maxima; the host finishes the max.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_alfven_max(const double* __restrict__ red,
                                                    const double* __restrict__ b2,
                                                    const double* __restrict__ rho0, long n,
                                                    double dimb2, double T, double den0,
                                                    double* __restrict__ part) {
  __shared__ double sm[256];
  double m = -1.0e300;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const double v = sqrt(dimb2 * T / (T + red[i]) * b2[i] / (den0 * rho0[i]));
    if (v > m) m = v;
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o && sm[threadIdx.x + o] > sm[threadIdx.x]) sm[threadIdx.x] = sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

hipError_t launch_alfven_max(const double* red, const double* b2, const double* rho0, long n,
                             double dimb2, double T, double den0, double* part, int nblk,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_alfven_max, dim3(nblk), dim3(256), 0, st, red, b2, rho0, n, dimb2, T, den0, part);
  return hipGetLastError();
}
}  // namespace sparc

namespace sparc {
// ---------------------------------------------------------------------------
// Stream probe: out[i] = sum_r in_r[i] over nread arrays, 16-byte loads and
// stores, grid-stride with four independent elements per thread per trip.
// Its achieved bandwidth at the read-to-write ratios of the stage kernels is
// the measured ceiling their roofline fractions are read against (the HBM
// peak of the data sheet is not reached by any access pattern).
// ---------------------------------------------------------------------------
typedef double sd2 __attribute__((ext_vector_type(2)));

struct StreamArgs {
  const sd2* in[48];
  sd2* out;
  long n2;  // 16-byte elements per array
  int nread;
};

__global__ __launch_bounds__(256) void k_stream(StreamArgs A) {
  const long stride = (long)gridDim.x * 256;
  for (long i0 = (long)blockIdx.x * 256 + threadIdx.x; i0 < A.n2; i0 += 4 * stride) {
    sd2 acc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = (sd2)(0.0);
    for (int r = 0; r < A.nread; ++r) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const long i = i0 + u * stride;
        if (i < A.n2) acc[u] += __builtin_nontemporal_load(&A.in[r][i]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long i = i0 + u * stride;
      if (i < A.n2) __builtin_nontemporal_store(acc[u], &A.out[i]);
    }
  }
}

hipError_t launch_stream(const double* const* in, int nread, double* out, long n, int nblk,
                         hipStream_t st) {
  if (nread < 1 || nread > 48 || n % 2) return hipErrorInvalidValue;
  StreamArgs A{};
  for (int r = 0; r < nread; ++r) A.in[r] = (const sd2*)in[r];
  A.out = (sd2*)out;
  A.n2 = n / 2;
  A.nread = nread;
  hipLaunchKernelGGL(k_stream, dim3(nblk), dim3(256), 0, st, A);
  return hipGetLastError();
}
}  // namespace sparc
