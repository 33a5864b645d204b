// sparc_dev.h -- device-side definitions shared by the SPARC2 HIP kernels.
//
// Layout in HBM: every field is one contiguous slab of `fs` elements (fs =
// nx*ny*nz rounded up to a multiple of 256), x fastest, then y, then z -- the
// reference's Fortran (x, y, z) order, so host <-> device copies are flat.
// A state buffer holds nv such slabs in the reference's field order
// (all_modules.f90:104, initialize.f90:780-826): rho, vx, vy, vz, p, bx, by, bz.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sparc {

enum Field { RHO = 0, VX = 1, VY = 2, VZ = 3, P = 4, BX = 5, BY = 6, BZ = 7 };
enum FaceId { XBOT = 0, XTOP = 1, YBOT = 2, YTOP = 3, ZBOT = 4, ZTOP = 5 };

// ---------------------------------------------------------------------------
// Coefficient tables (initialize.f90:1927-2034).  The centred 11-point DRP row
// is antisymmetric with a zero centre, so only its right half is kept:
// DPC[m-1] multiplies (f(i+m) - f(i-m)), m = 1..5.  DPB[r][m] are the five
// one-sided rows dpz(r+1, m+1), including the literal corrections the
// reference adds to some entries.
// ---------------------------------------------------------------------------
#define SPARC_DPC {0.872756993962, -0.286511173973, 0.090320001280, \
                   -0.020779405824, 0.002484594688}

#define SPARC_DPB { \
    {-2.391602219538, 5.832490322294, -7.650218001182 + 0.000000000001, \
    7.907810563576, -5.922599052629, 3.071037015445, -1.014956769726, \
    0.170022256519, 0.002819958377, -0.004791009708, -0.000013063429}, \
    {-0.180022054228, -1.237550583044, 2.484731692990, -1.810320814061, \
    1.112990048440, -0.481086916514, 0.126598690230, -0.015510730165, \
    0.000024609059 - 0.000003, 0.000156447571 - 0.000000000001, \
    -0.000007390277}, \
    {0.057982271137, -0.536135360383, -0.264089548967 + 0.000000000002, \
    0.917445877606 - 0.000000000002, -0.169688364841, -0.029716326170, \
    0.029681617641, -0.005222483773, -0.000118806260, -0.000118806260, \
    -0.000020069730}, \
    {-0.013277273810, 0.115976072920, -0.617479187931, \
    -0.274113948206 + 0.000000000002, 1.086208764655 - 0.000000000002, \
    -0.402951626982, 0.131066986242, -0.028154858354, 0.002596328316, \
    0.000128743150, 0.0}, \
    {0.016756572303, -0.117478455239, 0.411034935097, -1.130286765151, \
    0.341435872100 - 0.000000000001, 0.556396830543, -0.082525734207, \
    0.003565834658, 0.001173034777, -0.000071772671 + 0.000000000064, \
    -0.000000352273}}

__constant__ static const double DPB_D[5][11] = SPARC_DPB;

// The rows as pairs of floats, hi = fl32(b) and lo = fl32(b - hi), for the
// FP32 form (see RowSum): formed at compile time, so a term costs two table
// reads and no conversions.
struct DpbF {
  float hi[5][11], lo[5][11];
};
constexpr DpbF make_dpbf() {
  constexpr double d[5][11] = SPARC_DPB;
  DpbF t{};
  for (int r = 0; r < 5; ++r)
    for (int m = 0; m < 11; ++m) {
      t.hi[r][m] = (float)d[r][m];
      t.lo[r][m] = (float)(d[r][m] - (double)t.hi[r][m]);
    }
  return t;
}
__constant__ static const DpbF DPB_F = make_dpbf();

template <typename T> struct Coef;
template <> struct Coef<double> {
  __device__ static double c(int m) {
    constexpr double v[5] = SPARC_DPC;
    return v[m];
  }
};
template <> struct Coef<float> {
  __device__ static float c(int m) {
    constexpr double v[5] = SPARC_DPC;
    return (float)v[m];
  }
  // the rest of the FP64 coefficient: c + cl carries it to ~1e-15
  __device__ static float cl(int m) {
    constexpr double v[5] = SPARC_DPC;
    return (float)(v[m] - (double)(float)v[m]);
  }
};

// FP32 keeps each uniform scale of the scheme -- the derivative metric
// (stretchx, stretchy, stretch(k)), dt and the stage weights, the centred
// coefficients -- as a pair hi + lo of its FP64 value.  Rounded to one float
// each would change the discrete wave speed and the time step by a fixed
// ~1e-8 relative, a systematic phase error that grows linearly over
// thousands of steps (about 1e-5 relative L2 after 2000); the lo parts add
// one multiply-add each, in kernels bound by memory traffic.  x * (s + lo):
template <typename T, typename X> __device__ __forceinline__ X scl(T s, T lo, X x) {
  if constexpr (sizeof(T) == 4) {
    const X h = s * x;
    return h + lo * x;
  } else {
    (void)lo;
    return s * x;
  }
}

// Digital filters (initialize.f90:1932-1951): the 11-point filter is
// 1 - sin^10(theta/2) (binomial(10,k)/1024), the 7-point one sin^6(theta/2).
// Both are symmetric; W(g, m) is the weight at offset m in [-g, g].
template <typename T> __device__ inline T filt_w(int g, int m) {
  constexpr double w11[6] = {0.75390625000000, 0.5 * 0.41015625000000,
                             -0.5 * 0.23437500000000, 0.5 * 0.08789062500000,
                             -0.5 * 0.01953125000000, 0.5 * 0.00195312500000};
  constexpr double w7[4] = {0.312500000, -0.234375000, 0.093750000, -0.015625000};
  int am = m < 0 ? -m : m;
  return (T)(g == 5 ? w11[am] : w7[am]);
}

// ---------------------------------------------------------------------------
// One CPML face (RHS.f90:50-93, 424-546; storage initialize.f90:541-585).
// Face arrays are (n, ny, nz) for x faces, (nx, n, nz) for y faces and
// (nx, ny, n) for z faces, x fastest.  psi holds the RK base values
// (zpmlvars & co.), pst the stage values (psizpml & co.), each npsi slabs of
// `fsize`; kinv is 1/kappa.
// ---------------------------------------------------------------------------
template <typename T> struct Face {
  int active;
  int n;       // width along the face's axis
  int off;     // first global index along that axis
  int pad_;
  long fsize;
  const T* a;
  const T* b;
  const T* kinv;
  T* psi;
  T* pst;
};

// One RK stage.  u is the stage value (read with stencil halos), a the RK
// base (read at the point only).  A middle stage writes the next stage value
// a + beta_next*K into un; the last stage overwrites a with a + dt*K
// (step.f90:58) -- safe in place because a is only read at the point.
// State arrays are given per field (the slab base of each field).
template <typename T> struct StageArgs {
  int nx, ny, nz, kc;
  long fs;
  int ntx, nty, nblk_xy, nchunk;
  int zkc, znchunk;  // MHD z pass: its own z chunk (deeper: the queue warm-up re-reads 10 planes)
  int periodic, first, last, pulse_dir;
  // rhoinv holds 1/rho0 at every point (checked bitwise on the host), so the
  // FP64 MHD xy kernel divides (correctly rounded, the same bits) instead of
  // reading the array
  int rinv_rho0;
  // FP64 MHD xy kernel on 16x16 tiles with even nx: stage the planes with
  // direct-to-LDS buffer loads (k_stage_dma)
  int dma;
  T beta_next, dt, sxs, sys, src_amp;
  // FP32: the remainders of the FP64 values (scl above); 0 in FP64
  T beta_lo, dt_lo, sxs_lo, sys_lo;
  const T* stretch_lo;  // nz remainders of stretch(k) (FP32 only)
  // when amp_tab is set the source amplitude is amp_tab[*tstep] (a device
  // table and step counter, so a captured step replays unchanged)
  const T* amp_tab;
  const int* tstep;
  const T* u[8];
  T* a[8];
  T* un[8];           // what the stage writes: the next stage value, or a itself on the last stage
  uint16_t* comp[8];  // Kahan compensation (fp32 only, bfloat16), updated in place on the last stage
  T* dz[6];    // MHD z-pass output: z derivatives in CPML family order
  const T* damp[3];  // FFT damping terms of vx, vy, vz (null: no damping)
  // PULSE = 0 driver: this step's forcing(x, y) (nx*ny) and source_dep(z);
  // null fmap: the pulse (or no source)
  const T* fmap;
  const T* sdep;
  // backgrounds: 3-D (bg3d) or nz-long z profiles
  const T *rho0, *rhoinv, *c2rho0, *gr0z, *gp0z, *sponge, *p0;
  // MHD backgrounds (always 3-D)
  const T *box, *boy, *boz, *cbox, *cboy, *cboz, *gr0x, *gr0y, *gp0x, *gp0y, *red;
  const T *g, *stretch, *sx, *sy, *sz;
  Face<T> face[6];
  double* partial;  // last stage: 3 doubles per block (KE sum, min rho, min p)
  // the run's guard record (Guard below); null: no guard
  const int* halt;
};

// The stage kernels read their arguments through a pointer into the constant
// address space (the kernel-argument segment, or a device array that no
// kernel writes while it runs), so uniform fields become scalar loads.
#define SPARC_AS4 __attribute__((address_space(4)))
template <typename T> using CArgs = const SPARC_AS4 StageArgs<T>*;
template <typename T> using CPtr = const SPARC_AS4 T*;
// Opaque redefinition of the argument pointer: values derived from it
// (array bases, buffer descriptors) are reloaded after this point instead of
// being held in scalar registers across a loop.
#define SPARC_REFRESH(p) asm volatile("" : "+s"(p))

// Failure guard of one run (the reference's per-step checks, driver.f90:
// 175-196): norm_KE > 1e3, min(rho+rho0) <= 0 or min(p+p0) <= 0.  The
// diagnostic reduction of every step evaluates them on the device, so the
// host needs no per-step synchronisation.  Layout of the int words:
//   [0] halt: nonzero once a check has failed in halt mode; every later
//       kernel of the run (stages, z pass, filters, reduction) then returns
//       at once, so the state stays that of the failing step, before its
//       filter -- the state the reference dumps before it stops
//   [1] tripped (first failure recorded)   [2] step count at the failure
//   [3] steps taken (counted by the reduction, halted ones included, so
//       the count matches the host's)   [4] mode
// and gdiag holds the raw reduction (sum of m^2, min rho, min p) of the
// failing step.
enum GuardWord { G_HALT = 0, G_TRIPPED = 1, G_AT = 2, G_COUNT = 3, G_MODE = 4, G_WORDS = 8 };
enum GuardMode { GUARD_OFF = 0, GUARD_RECORD = 1, GUARD_HALT = 2 };

__device__ __forceinline__ bool halted(const int* g) {
  return g != nullptr && __atomic_load_n(g + G_HALT, __ATOMIC_RELAXED) != 0;
}

// Guard arguments of one run's reduction: the KE norm is
// sqrt(sum / npts) * dimc^2 * dimrho in cgs (driver.f90:161).
struct GuardArgs {
  int* guard;
  double* gdiag;
  double npts, ke_scale;
};

// One run's diagnostic reduction in the batched form.
struct DiagJob {
  const double* partial;
  double* out;
  int nblk;
  int pad_;
  GuardArgs g;
};

// ---------------------------------------------------------------------------
// Raw buffer view of one array: a uniform descriptor (SGPRs) plus a 32-bit
// per-lane byte offset, so every array indexed the same way shares one VGPR
// offset instead of a 64-bit address per array.  Loads past `n` elements
// return 0 and stores past it are dropped (the descriptor's range check).
// dword3 = 0x00020000 is the raw-buffer format word of the gfx9 family.
// ---------------------------------------------------------------------------
template <typename T> struct Buf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Buf(const T* p, uint32_t n)
      : r(__builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * (uint32_t)sizeof(T)),
                                            0x00020000)) {}
  __device__ __forceinline__ T operator[](uint32_t i) const {
    if constexpr (sizeof(T) == 8) {
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(i * 8u), 0, 0));
    } else {
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4u), 0, 0));
    }
  }
  // Load at a byte offset.  kOut lies past the range of every field-slab
  // view (a slab is a multiple of 1 KiB below 4 GiB), so a load there
  // returns 0 without a memory access: callers select it instead of
  // branching around a load.
  static constexpr uint32_t kOut = 0xFFFFFF00u;
  __device__ __forceinline__ T ldb(uint32_t boff) const {
    if constexpr (sizeof(T) == 8) {
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)boff, 0, 0));
    } else {
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)boff, 0, 0));
    }
  }
  __device__ __forceinline__ void st(uint32_t i, T v) const {
    if constexpr (sizeof(T) == 8) {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, (int)(i * 8u), 0, 0);
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)(i * 4u), 0, 0);
    }
  }
};

// ---------------------------------------------------------------------------
// Lane values.  The FP32 stage kernels on grids with even nx give each lane
// two adjacent x points: one 8-byte load or store moves both, so a wave
// moves as many bytes per memory instruction as the FP64 kernels do.  F2 is
// that pair; its arithmetic is elementwise, in the operation order of the
// scalar code, so a two-point lane computes what two one-point lanes would.
// ---------------------------------------------------------------------------
struct F2 {
  float x, y;
  F2() = default;
  __host__ __device__ constexpr F2(float s) : x(s), y(s) {}
  __host__ __device__ constexpr F2(float a, float b) : x(a), y(b) {}
};
__device__ __forceinline__ F2 operator+(F2 a, F2 b) { return F2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ F2 operator-(F2 a, F2 b) { return F2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ F2 operator*(F2 a, F2 b) { return F2(a.x * b.x, a.y * b.y); }
__device__ __forceinline__ F2 operator-(F2 a) { return F2(-a.x, -a.y); }
__device__ __forceinline__ F2& operator+=(F2& a, F2 b) { return a = a + b; }

// V = the value one lane holds: T (one point) or F2 (two points, FP32 only)
template <typename T, int VW> struct Lane { using V = T; };
template <> struct Lane<float, 2> { using V = F2; };

// The FP32 form's Kahan compensation is stored as bfloat16: the upper half
// of the float's bits, rounded to nearest even.  The compensation is the
// rounding error of the FP32 state (at most half its ulp), so 8 significant
// bits of it keep the pair (a, c) to ~2^-33 of |a| -- far below what FP32
// arithmetic in the right-hand side already leaves -- with FP32's exponent
// range, at half the bytes on the last RK stage and in the filters.
__host__ __device__ inline uint16_t f2bf(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// One-sided boundary row sum_m b(r, m) v_m (derivatives.f90:20-30: the five
// rows next to a non-periodic x or y face and below the top in z).  FP64:
// multiply-adds in row order.  FP32: the rows' coefficients reach 7.9 in
// magnitude and sum to zero, so each rounded coefficient's error, the same
// at every step, adds up coherently over thousands of steps: in the small
// flux-tube runs, whose edges are a large share of the grid, it set the
// FP32 error (population run 48x48x216 after 2000 steps: 1.4e-5 relative L2
// with float coefficients, 3.7e-6 with the pairs).  Each coefficient is
// therefore the pair hi + lo of its FP64 value, with two accumulators, as
// for the centred rows.  (A fully compensated sum -- Dot2, every product and
// addition split exactly -- came out no more accurate here and slower.)
template <typename T> struct RowSum {
  T p = 0, s = 0;
  __device__ __forceinline__ void add(T v, int r, int m) {
    if constexpr (sizeof(T) == 8) {
      p += v * DPB_D[r][m];
    } else {
      p += v * DPB_F.hi[r][m];
      s += v * DPB_F.lo[r][m];
    }
  }
  __device__ __forceinline__ T sum() const {
    if constexpr (sizeof(T) == 8) return p;
    else return p + s;
  }
};
struct RowSumF2 {
  RowSum<float> a, b;
  __device__ __forceinline__ void add(F2 v, int r, int m) {
    a.add(v.x, r, m);
    b.add(v.y, r, m);
  }
  __device__ __forceinline__ F2 sum() const { return F2(a.sum(), b.sum()); }
};
template <typename V> struct RowSumOf { using type = RowSum<V>; };
template <> struct RowSumOf<F2> { using type = RowSumF2; };

// The row tables copied into LDS, for the stage kernels: there the row of a
// one-sided closure differs between the lanes of a wave, so reading the
// constant tables takes vector memory loads, and their wait also covers
// every load in flight (the point values of the plane); LDS reads do not.
// FP32: hi and lo of the coefficient pairs.
template <typename T> struct RowTab {
  T c[5 * 11];
  T l[sizeof(T) == 4 ? 5 * 11 : 1];
  // every thread of the block calls this; a barrier must follow
  __device__ __forceinline__ void fill(int tid, int nt) {
    for (int q = tid; q < 5 * 11; q += nt) {
      if constexpr (sizeof(T) == 8) {
        c[q] = DPB_D[q / 11][q % 11];
      } else {
        c[q] = DPB_F.hi[q / 11][q % 11];
        l[q] = DPB_F.lo[q / 11][q % 11];
      }
    }
  }
};
// RowSum over an LDS copy of the tables (the same products and sums)
template <typename T> struct RowSumT {
  const RowTab<T>& t;
  T p = 0, s = 0;
  __device__ __forceinline__ explicit RowSumT(const RowTab<T>& tab) : t(tab) {}
  __device__ __forceinline__ void add(T v, int r, int m) {
    if constexpr (sizeof(T) == 8) {
      p += v * t.c[r * 11 + m];
    } else {
      p += v * t.c[r * 11 + m];
      s += v * t.l[r * 11 + m];
    }
  }
  __device__ __forceinline__ T sum() const {
    if constexpr (sizeof(T) == 8) return p;
    else return p + s;
  }
};
struct RowSumTF2 {
  RowSumT<float> a, b;
  __device__ __forceinline__ explicit RowSumTF2(const RowTab<float>& tab) : a(tab), b(tab) {}
  __device__ __forceinline__ void add(F2 v, int r, int m) {
    a.add(v.x, r, m);
    b.add(v.y, r, m);
  }
  __device__ __forceinline__ F2 sum() const { return F2(a.sum(), b.sum()); }
};
template <typename V> struct RowSumTOf { using type = RowSumT<V>; };
template <> struct RowSumTOf<F2> { using type = RowSumTF2; };

// point c of a lane value
template <typename T> __device__ __forceinline__ T lane_pt(T v, int) { return v; }
__device__ __forceinline__ float lane_pt(F2 v, int c) { return c ? v.y : v.x; }
template <typename T> __device__ __forceinline__ void set_lane_pt(T& v, int, T s) { v = s; }
__device__ __forceinline__ void set_lane_pt(F2& v, int c, float s) {
  if (c) v.y = s;
  else v.x = s;
}

// Lane-value load / store through a buffer view of T elements: byte offset
// bo (Buf::kOut reads 0) or element index e of the lane's first point.  CP
// is the cache policy: 0, or kStream (gfx950 `nt`, non-temporal) for values
// that no other block of the launch touches -- the point values, the RK base
// and the outputs -- so they do not evict the staged planes whose halos the
// neighbouring tiles read again from L2.
constexpr int kStream = 2;
template <typename V, int CP = 0, typename T>
__device__ __forceinline__ V ldv(const Buf<T>& b, uint32_t bo) {
  if constexpr (sizeof(V) == 8) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)bo, 0, CP));
  } else {
    static_assert(sizeof(V) == 4, "a lane value of 4 or 8 bytes");
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)bo, 0, CP));
  }
}
template <int CP = 0, typename V, typename T>
__device__ __forceinline__ void stv(const Buf<T>& b, uint32_t e, V v) {
  if constexpr (sizeof(V) == 8) {
    using R = decltype(__builtin_amdgcn_raw_buffer_load_b64(b.r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(R, v), b.r, (int)(e * (uint32_t)sizeof(T)), 0, CP);
  } else {
    static_assert(sizeof(V) == 4, "a lane value of 4 or 8 bytes");
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), b.r, (int)(e * 4u), 0, CP);
  }
}

// 16 bytes per lane from a buffer straight into LDS, lane l writing bytes
// lds + 16 l .. lds + 16 l + 15 (gfx950 buffer_load_dwordx4 with lds; no
// registers).  lds is uniform over the wave and points into shared memory.
// Counted by vmcnt like any load.
__device__ __forceinline__ void lds_load16(__amdgpu_buffer_rsrc_t r, void* lds, uint32_t voff) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 16, voff, 0, 0, 0);
}

// Compensation (bfloat16) of the lane's point(s): V = float (one point, a
// 2-byte access) or F2 (two points, 4 bytes).  ldcomp takes the byte offset
// of the lane's first point in a float field, halved here (Buf<float>::kOut
// stays past the range); stcomp the element index.
template <typename V, int CP = 0>
__device__ __forceinline__ V ldcomp(const Buf<uint16_t>& b, uint32_t fbo) {
  const int bo = (int)(fbo == Buf<float>::kOut ? fbo : fbo >> 1);
  if constexpr (sizeof(V) == 8) {
    const uint32_t u = __builtin_amdgcn_raw_buffer_load_b32(b.r, bo, 0, CP);
    return V(bf2f((uint16_t)(u & 0xFFFFu)), bf2f((uint16_t)(u >> 16)));
  } else {
    static_assert(sizeof(V) == 4, "a compensation lane of one or two FP32 points");
    return bf2f(__builtin_amdgcn_raw_buffer_load_b16(b.r, bo, 0, CP));
  }
}
// The same load with the bits kept raw (in .x for F2, zero-extended for
// float), widened by comp_of_raw where the value is used: widening at the
// load would wait for it there, and a prefetch must not be waited on early.
template <typename V, int CP = 0>
__device__ __forceinline__ V ldcomp_raw(const Buf<uint16_t>& b, uint32_t fbo) {
  const int bo = (int)(fbo == Buf<float>::kOut ? fbo : fbo >> 1);
  if constexpr (sizeof(V) == 8) {
    return V(__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, bo, 0, CP)), 0.0f);
  } else {
    static_assert(sizeof(V) == 4, "a compensation lane of one or two FP32 points");
    return __builtin_bit_cast(float, (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(b.r, bo, 0, CP));
  }
}
template <typename V> __device__ __forceinline__ V comp_of_raw(V r) {
  if constexpr (sizeof(V) == 8) {
    const uint32_t u = __builtin_bit_cast(uint32_t, r.x);
    return V(bf2f((uint16_t)(u & 0xFFFFu)), bf2f((uint16_t)(u >> 16)));
  } else {
    return bf2f((uint16_t)__builtin_bit_cast(uint32_t, r));
  }
}
template <int CP = 0, typename V>
__device__ __forceinline__ void stcomp(const Buf<uint16_t>& b, uint32_t e, V v) {
  if constexpr (sizeof(V) == 8) {
    const uint32_t u = (uint32_t)f2bf(v.x) | ((uint32_t)f2bf(v.y) << 16);
    __builtin_amdgcn_raw_buffer_store_b32(u, b.r, (int)(e * 2u), 0, CP);
  } else {
    static_assert(sizeof(V) == 4, "a compensation lane of one or two FP32 points");
    __builtin_amdgcn_raw_buffer_store_b16(f2bf(v), b.r, (int)(e * 2u), 0, CP);
  }
}

// CPML update at one point (or one lane's point pair) of one face
// (RHS.f90:52-55 per derivative family): the psi RHS comes from the raw
// derivative, then the derivative is replaced by d/kappa + psi; the next psi
// stage value (step.f90:34-36) or, on the last stage, the advanced base
// (step.f90:59-61) is written in place.  v[i] is the derivative of psi
// family fam[i]; fp the face element of the lane's first point.  All loads
// of the face point (a, b, 1/kappa once, then every family's base and stage
// value) are issued before any store, so they share one memory wait.  FR is
// a Face<T> in any address space; CPL and CPW the cache policies of its
// loads and stores.
template <int CPL = 0, int CPW = CPL, typename T, typename V, int NF, typename FR>
__device__ __forceinline__ void pml_face(FR& F, uint32_t fp, V (&v)[NF], const int (&fam)[NF],
                                         bool first, bool last, T beta_next, T dt) {
  const uint32_t n = (uint32_t)F.fsize;
  const uint32_t bo = fp * (uint32_t)sizeof(T);
  const V ca = ldv<V, CPL>(Buf<T>(F.a, n), bo), cb = ldv<V, CPL>(Buf<T>(F.b, n), bo),
          ck = ldv<V, CPL>(Buf<T>(F.kinv, n), bo);
  // the stage value is the base itself on the first stage: its load then
  // reads nothing (a branch around it would split the loads into one round
  // trip per family)
  const uint32_t bs = first ? Buf<T>::kOut : bo;
  V base[NF], ps[NF];
#pragma unroll
  for (int i = 0; i < NF; ++i) {
    base[i] = ldv<V, CPL>(Buf<T>(F.psi + (size_t)fam[i] * n, n), bo);
    ps[i] = ldv<V, CPL>(Buf<T>(F.pst + (size_t)fam[i] * n, n), bs);
  }
#pragma unroll
  for (int i = 0; i < NF; ++i) ps[i] = first ? base[i] : ps[i];
#pragma unroll
  for (int i = 0; i < NF; ++i) {
    const V kn = ca * v[i] + cb * ps[i];
    v[i] = v[i] * ck + ps[i];
    if (last)
      stv<CPW>(Buf<T>(F.psi + (size_t)fam[i] * n, n), fp, base[i] + dt * kn);
    else
      stv<CPW>(Buf<T>(F.pst + (size_t)fam[i] * n, n), fp, base[i] + beta_next * kn);
  }
}

__device__ __forceinline__ int wrap(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

}  // namespace sparc
