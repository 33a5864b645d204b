// api.hip -- host side of the C ABI declared in include/sparc_hip.h.
//
// A context owns every device buffer the reference's STEP / FILTER_VARS read
// or write (module arrays of all_modules.f90:96-146).  State-sized buffers
// form a pool of three: the RK base a and two stage-value buffers that the
// stages ping-pong between (a filter pass uses one of them as its output and
// swaps roles with a instead of copying back).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sparc_hip.h"
#include "sparc_dev.h"

namespace sparc {
template <typename T>
hipError_t launch_stage(bool mhd, bool bg3d, bool zp, int vw, int tx, int ty, const StageArgs<T>& A,
                        hipStream_t s);
template <typename T>
hipError_t launch_zpass(const StageArgs<T>& A, hipStream_t s);
template <typename T>
hipError_t launch_zpass_batch(const StageArgs<T>* runs, const int* item_start, int nrun, int nblk,
                              int* work, hipStream_t s);
long zpass_items(int nx, int ny, int nchunk);
int zpass_batch_blocks(bool f32, int ncu);
bool stage_tile_supported(int tx, int ty);
template <typename T>
hipError_t launch_filter(int axis, int g, const T* in, T* out, int nx, int ny, int nz, long fs,
                         int nfield, int periodic, T s, int bare, hipStream_t st, const int* halt,
                         bool inplace = false);
template <typename T>
hipError_t launch_filter_xy(int g, const void* in, void* out, int nx, int ny, int nz, long fs,
                            int nfield, int periodic, T s, hipStream_t st, const int* halt, bool cbf);
template <typename T>
hipError_t launch_filter_zc(int g, const void* in, void* out, int nx, int ny, int nz, long fs,
                            int nfield, T s, int bare, hipStream_t st, const int* halt, bool inplace,
                            bool cbf);
hipError_t launch_deriv(int axis, const double* in, double* out, int nx, int ny, int nz, double sxy,
                        const double* stretch, int bc, hipStream_t st);
hipError_t launch_diag_reduce(const double* partial, int nblk, double* out, const GuardArgs& g,
                              hipStream_t st);
hipError_t launch_zero(void* p, long n, bool f32, hipStream_t st);
template <typename T>
hipError_t launch_stage_batch(bool mhd, bool bg3d, bool zp, int vw, int tx, int ty,
                              const StageArgs<T>* runs, const int* item_start, int nrun, int nblk,
                              int* work, hipStream_t s, bool dma);
int stage_batch_blocks(bool f32, bool mhd, bool bg3d, bool zp, int vw, int tx, int ty, int ncu, bool dma);
hipError_t launch_diag_reduce_batch(const void* jobs, int njobs, hipStream_t st);
hipError_t launch_tick(int* c, int reset, hipStream_t st);
hipError_t launch_copy(const void* in, void* out, long n, bool f32, hipStream_t st, const int* halt);
hipError_t launch_stream(const double* const* in, int nread, double* out, long n, int nblk,
                         hipStream_t st);
hipError_t launch_alfven_max(const double* red, const double* b2, const double* rho0, long n,
                             double dimb2, double T, double den0, double* part, int nblk,
                             hipStream_t st);
struct DampPlan;
DampPlan* damp_create(int nx, int ny, int nz, long fs, bool f32, const double* rates, hipStream_t st);
void damp_destroy(DampPlan* d);
hipError_t damp_terms(DampPlan* d, const void* u, hipStream_t st);
const void* damp_term_slab(const DampPlan* d, int f);
}  // namespace sparc

using namespace sparc;

static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(SPARC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

enum BgSlot {
  B_RHO0, B_RHOINV, B_C2RHO0, B_GR0Z, B_GP0Z, B_SPONGE, B_P0,
  B_BOX, B_BOY, B_BOZ, B_CBOX, B_CBOY, B_CBOZ, B_GR0X, B_GR0Y, B_GP0X, B_GP0Y, B_RED,
  B_COUNT
};

struct FaceDev {
  int active = 0, n = 0, off = 0;
  long fsize = 0;
  void *a = nullptr, *b = nullptr, *k = nullptr, *psi = nullptr, *pst = nullptr;
};

struct TimeRec {
  int kind;
  hipEvent_t e0, e1;
};

}  // namespace

struct sparc_ctx {
  sparc_cfg cfg{};
  std::vector<double> stretch_h;
  int dev = 0;
  hipStream_t st = nullptr;
  bool f32 = false, mhd = false, bg3d = false;
  int nv = 5, npsi = 2;
  int nx = 0, ny = 0, nz = 0;
  long N = 0, fs = 0;
  size_t esz = 8;
  void* pool[3] = {nullptr, nullptr, nullptr};
  int ra = 0, rk0 = 1, rk1 = 2;
  void* comp[2] = {nullptr, nullptr};
  int rc = 0;
  // MHD: the z pass writes the six z derivatives here (6 slabs); zp is off
  // only for the fused single-kernel MHD stage (SPARC_MHD_FUSED=1)
  bool zp = false;
  // x points per lane of the stage kernel: 2 for FP32 on even nx (quiet or
  // MHD xy kernel; SPARC_LANE_POINTS=1 keeps one), else 1.  Tile widths
  // count lanes, so a tile spans vw*tx points.
  int vw = 1;
  // MHD xy kernel: stage with direct-to-LDS loads where the tile is 16x16
  // and each row has an even number of 8-byte lanes (nx even in FP64, a
  // multiple of 4 with FP32's two-point lanes); SPARC_MHD_DMA=0 keeps the
  // register staging
  bool dma = false;
  void* dzbuf = nullptr;
  void* bg[B_COUNT] = {};
  void *g = nullptr, *stretch = nullptr, *sx = nullptr, *sy = nullptr, *sz = nullptr;
  void* stretch_lo = nullptr;  // FP32: stretch(k) - fl32(stretch(k))
  double* stretch_d = nullptr;  // FP64 copy for the standalone operators
  FaceDev face[6];
  DampPlan* damp = nullptr;  // FFT damping (DAMPING = 1), null when off
  // PULSE = 0 driver: forcing map (nx*ny) and source_dep (nz), element type
  void *fmap = nullptr, *sdep = nullptr;
  bool forcing_on = false;
  double* partial = nullptr;
  long partial_cap = 0;  // blocks the partial buffer holds
  double* diag = nullptr;
  // failure guard record (GuardWord layout) and the raw diagnostics of the
  // failing step; host copy of the mode
  int* guard = nullptr;
  double* gdiag = nullptr;
  int guard_mode = GUARD_HALT;
  double dimc = 1.0, dimrho = 1.0;
  int tx = 64, ty = 8, kc = 8, ntx = 1, nty = 1, nchunk = 1;
  int zkc = 8, znchunk = 1;  // z-pass chunking
  bool bg_set = false, state_set = false;
  // the uploaded rhoinv is bitwise 1/rho0 (initialize.f90:1012) at every point
  bool rinv_rho0 = false;
  long steps_done = 0;
  double* op_buf = nullptr;  // 2*N doubles for the standalone operators
  bool timing = false;
  std::vector<TimeRec> trecs;
  long tcount[5] = {0, 0, 0, 0, 0};
  double tms[5] = {0, 0, 0, 0, 0};
};

static long round_up(long v, long m) { return (v + m - 1) / m * m; }

static int dalloc(void** p, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(SPARC_E_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " +
                                   hipGetErrorString(e));
  }
  return SPARC_OK;
}

// Upload n doubles converted to the context's element type.
static int upload(sparc_ctx* c, void* dst, const double* src, long n) {
  if (n <= 0) return SPARC_OK;
  if (!c->f32) {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  } else {
    std::vector<float> tmp(n);
    for (long q = 0; q < n; ++q) tmp[q] = (float)src[q];
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  }
  return SPARC_OK;
}

static int download(sparc_ctx* c, double* dst, const void* src, long n) {
  if (n <= 0) return SPARC_OK;
  if (!c->f32) {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  } else {
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (long q = 0; q < n; ++q) dst[q] = (double)tmp[q];
  }
  return SPARC_OK;
}

// n bfloat16 compensation terms from the device
static int download_comp(sparc_ctx* c, uint16_t* dst, const void* src, long n) {
  if (n <= 0) return SPARC_OK;
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(uint16_t) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SPARC_OK;
}

static void free_all(sparc_ctx* c) {
  auto F = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (auto& p : c->pool) F(p);
  for (auto& p : c->comp) F(p);
  F(c->dzbuf);
  for (auto& p : c->bg) F(p);
  F(c->g);
  F(c->stretch);
  F(c->stretch_lo);
  F(c->sx);
  F(c->sy);
  F(c->sz);
  F(c->stretch_d);
  F(c->partial);
  F(c->diag);
  F(c->guard);
  F(c->gdiag);
  F(c->op_buf);
  F(c->fmap);
  F(c->sdep);
  damp_destroy(c->damp);
  c->damp = nullptr;
  for (auto& f : c->face) {
    F(f.a);
    F(f.b);
    F(f.k);
    F(f.psi);
    F(f.pst);
  }
  for (auto& r : c->trecs) {
    (void)hipEventDestroy(r.e0);
    (void)hipEventDestroy(r.e1);
  }
  if (c->st) (void)hipStreamDestroy(c->st);
}

// Tile choice.  Shapes (in lanes) the stage kernel is built for, in order of
// preference per variant (measured on MI355X, 512x512x300 quiet and 256^3
// MHD: quiet FP64 and FP32 run fastest at 32x8, the MHD xy kernel at 32x16
// in FP64, at 32x8 in FP32 with one point per lane and at 16x16 with two).
// The first shape that fits the grid (one-sided closures need >= 6 rows in
// the last tile of a non-periodic axis) is taken.  The z chunk is sized so
// the launch has >= ~2048 blocks (quiet) or ~8192 (MHD xy kernel) without
// dropping below 4 planes per block.
static bool tile_fits(const sparc_ctx* c, int tx, int ty) {
  if (!stage_tile_supported(tx, ty)) return false;
  if (c->cfg.periodic) return true;
  const int rx = c->nx % (tx * c->vw), ry = c->ny % ty;
  return ty >= 6 && !(rx > 0 && rx < 6) && !(ry > 0 && ry < 6);
}

static void set_geometry(sparc_ctx* c, int tx, int ty, int kc);

static void choose_tile(sparc_ctx* c) {
  const sparc_cfg& q = c->cfg;
  int tx = q.tile[0], ty = q.tile[1];
  if (!(tx > 0 && ty > 0 && stage_tile_supported(tx, ty))) {
    static const int quiet64[4][2] = {{32, 8}, {16, 16}, {32, 16}, {64, 4}};
    static const int quiet32[4][2] = {{32, 8}, {64, 4}, {16, 16}, {32, 16}};
    static const int mhd64[4][2] = {{16, 16}, {32, 16}, {32, 8}, {64, 4}};
    static const int mhd32[4][2] = {{32, 8}, {32, 16}, {16, 16}, {64, 4}};
    // two points per lane (FP32, even nx): the MHD xy kernel at 16x16 lanes
    static const int mhd32x2[4][2] = {{16, 16}, {32, 8}, {32, 16}, {64, 4}};
    const int(*pref)[2] = c->mhd ? (c->f32 ? (c->vw == 2 ? mhd32x2 : mhd32) : mhd64)
                                 : (c->f32 ? quiet32 : quiet64);
    tx = 32;
    ty = 8;
    for (int n = 0; n < 4; ++n) {
      // narrow grids: a 16-wide tile wastes less of each block
      if (c->nx <= 16 && pref[n][0] * c->vw > 16) continue;
      if (tile_fits(c, pref[n][0], pref[n][1])) {
        tx = pref[n][0];
        ty = pref[n][1];
        break;
      }
    }
    // quiet: the kernel runs fastest with about 512 blocks of full-depth
    // chunks; a plane of more than 512 tiles of 32x8 takes 32x16 tiles
    // instead (512x512x300 FP64, one stage: 32x8 2.61 ms at 150-plane
    // chunks, 2.53 ms at 300; 32x16 2.22 ms at 300)
    const long t32x8 = (long)((c->nx + 32 * c->vw - 1) / (32 * c->vw)) * ((c->ny + 7) / 8);
    if (!c->mhd && tx == 32 && ty == 8 && t32x8 > 512 && tile_fits(c, 32, 16)) ty = 16;
  }
  set_geometry(c, tx, ty, q.tile[2]);
}

// Tile shape, z chunk (kc <= 0: automatic) and the z pass's chunking.
static void set_geometry(sparc_ctx* c, int tx, int ty, int kc) {
  c->tx = tx;
  c->ty = ty;
  c->ntx = (c->nx + tx * c->vw - 1) / (tx * c->vw);
  c->nty = (c->ny + ty - 1) / ty;
  long tiles = (long)c->ntx * c->nty;
  if (kc <= 0) {
    // the quiet kernel re-reads 10 planes of its z queue per chunk, so it
    // keeps chunks deep (~512 blocks: 512x512x300 FP32 32x8 lanes, one stage
    // 1.41 ms at 2048 blocks, 1.30 at 1024, 1.23 at 512); the MHD xy kernel
    // has no queue and runs fastest with many shallow chunks (4 planes,
    // measured on 256^3: 16x16 FP64 1.85 ms at 8 planes, 1.77 ms at 4)
    // small grids: 2-plane chunks (64x64x128 quiet, one step: FP64 32x8
    // 0.192 ms at 4 planes, 0.184 at 2, 0.198 at 1; FP32 0.171 / 0.119 /
    // 0.120), so a launch still has ~512 blocks
    long want = c->zp ? 16384 : 512;
    kc = (int)std::max<long>(c->zp ? 4 : 2, (c->nz * tiles + want - 1) / want);
    kc = std::min(kc, c->nz);
  }
  c->kc = kc;
  c->nchunk = (c->nz + kc - 1) / kc;
  // z pass: about 512 blocks of 256 columns, at least 32 planes deep so the
  // 10-plane queue warm-up stays a small overhead (256^3, two chunks of 128
  // planes: FP64 0.471 ms, FP32 0.263 ms; one chunk of 256 planes, 256
  // blocks: 0.492 / 0.391 ms; four chunks, 1024 blocks: 0.564 / 0.284 ms)
  const long ncb = ((long)c->nx * c->ny + 255) / 256;
  c->zkc = (int)std::min<long>(c->nz, std::max<long>(32, (c->nz * ncb + 511) / 512));
  c->znchunk = (c->nz + c->zkc - 1) / c->zkc;
}

static int check_tile(sparc_ctx* c) {
  if (!stage_tile_supported(c->tx, c->ty))
    return fail(SPARC_E_INVALID, "unsupported stage tile");
  // one-sided closures need rows 0..10 and n-11..n-1 inside one LDS tile
  if (!tile_fits(c, c->tx, c->ty))
    return fail(SPARC_E_INVALID,
                "non-periodic runs need tile_y >= 6 and the last tile to hold >= 6 cells");
  return SPARC_OK;
}

// The exported functions keep the C linkage of their declarations in
// sparc_hip.h.

int sparc_hip_version(void) { return 1; }

const char* sparc_hip_last_error(void) { return g_last_error.c_str(); }

const char* sparc_hip_strerror(int code) {
  switch (code) {
    case SPARC_OK: return "ok";
    case SPARC_E_INVALID: return "invalid argument or configuration";
    case SPARC_E_HIP: return "HIP runtime error";
    case SPARC_E_NOMEM: return "device out of memory";
    case SPARC_E_STATE: return "call out of order";
    default: return "unknown error";
  }
}

// Staging of the allocation-free derivative (sparc_hip_deriv_noalloc): input
// and output of one field in a file-scope device array, so that entry point
// needs no hipMalloc.  Every device the library runs on holds its own copy
// (64 MiB) until the process ends.  In host code `scratch` names a host-side
// stand-in: its device address comes from hipGetSymbolAddress, once per
// device, when the first context on that device is created.
static constexpr long kNoAllocMaxPoints = 1L << 22;  // 4,194,304 grid points
static __device__ __align__(256) double scratch[2 * kNoAllocMaxPoints];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// With device `dev` current: look up and keep this device's scratch address.
static int scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(SPARC_E_INVALID, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return SPARC_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(SPARC_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = p;
  return SPARC_OK;
}

int sparc_hip_create(const sparc_cfg* cfg, sparc_ctx** out) {
  if (!cfg || !out) return fail(SPARC_E_INVALID, "null argument");
  *out = nullptr;
  const sparc_cfg& q = *cfg;
  if (q.nx < 16 || q.ny < 16 || q.nz < 16) return fail(SPARC_E_INVALID, "grid needs >= 16 cells per axis");
  if (q.option != 1 && q.option != 3) return fail(SPARC_E_INVALID, "option must be 1 (quiet) or 3 (MHD)");
  if (q.precision != SPARC_PREC_FP64 && q.precision != SPARC_PREC_FP32)
    return fail(SPARC_E_INVALID, "precision must be FP64 or FP32");
  if ((q.g_xy_filt != 5 && q.g_xy_filt != 3) || (q.g_z_filt != 5 && q.g_z_filt != 3))
    return fail(SPARC_E_INVALID, "filter half-widths must be 5 or 3");
  if (q.xyfreq < 1 || q.zfreq < 1) return fail(SPARC_E_INVALID, "filter cadences must be >= 1");
  if (q.pulse_dir < 0 || q.pulse_dir > 3) return fail(SPARC_E_INVALID, "pulse_dir must be 0..3");
  if (!q.stretch) return fail(SPARC_E_INVALID, "stretch (nz values) is required");
  if (q.use_hpml && q.periodic) return fail(SPARC_E_INVALID, "USE_HPML requires PERIODIC off");
  const int ext[3] = {q.nx, q.ny, q.nz};
  for (int f = 0; f < 6; ++f) {
    bool on = (f < 4 ? q.use_hpml : q.use_pml) && q.npml[f] > 0;
    if (on && (q.npml[f] < 2 || 2 * q.npml[f] > ext[f / 2]))
      return fail(SPARC_E_INVALID, "PML width must be >= 2 and fit the axis");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(SPARC_E_HIP, "no HIP device available");
  if (q.device < 0 || q.device >= ndev) return fail(SPARC_E_INVALID, "bad device ordinal");

  sparc_ctx* c = new sparc_ctx();
  c->cfg = q;
  c->stretch_h.assign(q.stretch, q.stretch + q.nz);
  c->cfg.stretch = c->stretch_h.data();
  c->dev = q.device;
  c->f32 = q.precision == SPARC_PREC_FP32;
  c->esz = c->f32 ? 4 : 8;
  c->mhd = q.option == 3;
  c->bg3d = c->mhd ? true : (q.bg3d != 0);
  {
    const char* fused = getenv("SPARC_MHD_FUSED");
    c->zp = c->mhd && !(fused && fused[0] == '1');
    const char* lp = getenv("SPARC_LANE_POINTS");
    c->vw = (c->f32 && q.nx % 2 == 0 && (c->zp || !c->mhd) && !(lp && lp[0] == '1')) ? 2 : 1;
    const char* dm = getenv("SPARC_MHD_DMA");
    c->dma = c->zp && (c->f32 ? (c->vw == 2 && q.nx % 4 == 0) : q.nx % 2 == 0) && !(dm && dm[0] == '0');
  }
  c->nv = c->mhd ? 8 : 5;
  c->npsi = c->mhd ? 6 : 2;
  c->nx = q.nx;
  c->ny = q.ny;
  c->nz = q.nz;
  c->N = (long)q.nx * q.ny * q.nz;
  c->fs = round_up(c->N, 256);
  // kernels address one field slab (and one CPML face array) with 32-bit
  // byte offsets through buffer descriptors
  if ((unsigned long long)c->fs * c->esz >= (1ull << 32)) {
    delete c;
    return fail(SPARC_E_INVALID, "one field slab must stay below 4 GiB: split the domain in z");
  }
  choose_tile(c);
  int rc = check_tile(c);
  if (rc) {
    delete c;
    return rc;
  }
  auto bail = [&](int code) {
    free_all(c);
    delete c;
    return code;
  };
  if (hipSetDevice(c->dev) != hipSuccess) return bail(fail(SPARC_E_HIP, "hipSetDevice failed"));
  if ((rc = scratch_setup(c->dev))) return bail(rc);
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SPARC_E_HIP, "hipStreamCreate failed"));

  const size_t sbytes = c->esz * (size_t)c->nv * (size_t)c->fs;
  for (int p = 0; p < 3; ++p)
    if ((rc = dalloc(&c->pool[p], sbytes))) return bail(rc);
  if (c->f32)
    for (int p = 0; p < 2; ++p)
      if ((rc = dalloc(&c->comp[p], sbytes / 2))) return bail(rc);  // bfloat16 terms
  if (c->zp)
    if ((rc = dalloc(&c->dzbuf, c->esz * 6 * (size_t)c->fs))) return bail(rc);
  const long bgn = c->bg3d ? c->fs : q.nz;
  const int nbg = c->mhd ? B_COUNT : B_P0 + 1;
  for (int s = 0; s < nbg; ++s)
    if ((rc = dalloc(&c->bg[s], c->esz * bgn))) return bail(rc);
  if ((rc = dalloc(&c->g, c->esz * q.nz))) return bail(rc);
  if ((rc = dalloc(&c->stretch, c->esz * q.nz))) return bail(rc);
  if ((rc = dalloc(&c->sx, c->esz * q.nx))) return bail(rc);
  if ((rc = dalloc(&c->sy, c->esz * q.ny))) return bail(rc);
  if ((rc = dalloc(&c->sz, c->esz * q.nz))) return bail(rc);
  if ((rc = dalloc((void**)&c->stretch_d, sizeof(double) * q.nz))) return bail(rc);
  const long nblk = (long)c->ntx * c->nty * c->nchunk;
  if ((rc = dalloc((void**)&c->partial, sizeof(double) * 3 * nblk))) return bail(rc);
  c->partial_cap = nblk;
  if ((rc = dalloc((void**)&c->diag, sizeof(double) * 3))) return bail(rc);
  if ((rc = dalloc((void**)&c->guard, sizeof(int) * G_WORDS))) return bail(rc);
  if ((rc = dalloc((void**)&c->gdiag, sizeof(double) * 4))) return bail(rc);
  {
    int g0[G_WORDS] = {};
    g0[G_MODE] = GUARD_HALT;
    if (hipMemcpy(c->guard, g0, sizeof g0, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(SPARC_E_HIP, "guard upload failed"));
  }
  for (int f = 0; f < 6; ++f) {
    FaceDev& F = c->face[f];
    F.active = (f < 4 ? q.use_hpml : q.use_pml) && q.npml[f] > 0;
    if (!F.active) continue;
    const int axis = f / 2;
    F.n = q.npml[f];
    F.off = (f % 2) ? ext[axis] - F.n : 0;
    long d[3] = {q.nx, q.ny, q.nz};
    d[axis] = F.n;
    F.fsize = d[0] * d[1] * d[2];
    if ((rc = dalloc(&F.a, c->esz * F.fsize))) return bail(rc);
    if ((rc = dalloc(&F.b, c->esz * F.fsize))) return bail(rc);
    if ((rc = dalloc(&F.k, c->esz * F.fsize))) return bail(rc);
    if ((rc = dalloc(&F.psi, c->esz * c->npsi * F.fsize))) return bail(rc);
    if ((rc = dalloc(&F.pst, c->esz * c->npsi * F.fsize))) return bail(rc);
    if (launch_zero(F.psi, c->npsi * F.fsize, c->f32, c->st) != hipSuccess ||
        launch_zero(F.pst, c->npsi * F.fsize, c->f32, c->st) != hipSuccess)
      return bail(fail(SPARC_E_HIP, "zero fill failed"));
  }
  for (int p = 0; p < 3; ++p)
    if (launch_zero(c->pool[p], (long)c->nv * c->fs, c->f32, c->st) != hipSuccess)
      return bail(fail(SPARC_E_HIP, "zero fill failed"));
  if (c->f32)
    for (int p = 0; p < 2; ++p)  // bfloat16 terms: half as many 4-byte words
      if (launch_zero(c->comp[p], (long)c->nv * c->fs / 2, true, c->st) != hipSuccess)
        return bail(fail(SPARC_E_HIP, "zero fill failed"));
  if ((rc = upload(c, c->stretch, q.stretch, q.nz))) return bail(rc);
  if (c->f32) {
    std::vector<float> lo(q.nz);
    for (int k = 0; k < q.nz; ++k) lo[k] = (float)(q.stretch[k] - (double)(float)q.stretch[k]);
    if ((rc = dalloc(&c->stretch_lo, sizeof(float) * q.nz))) return bail(rc);
    if (hipMemcpy(c->stretch_lo, lo.data(), sizeof(float) * q.nz, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(SPARC_E_HIP, "stretch upload failed"));
  }
  if (hipMemcpy(c->stretch_d, q.stretch, sizeof(double) * q.nz, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(SPARC_E_HIP, "stretch upload failed"));
  if (hipStreamSynchronize(c->st) != hipSuccess) return bail(fail(SPARC_E_HIP, "sync failed"));
  *out = c;
  return SPARC_OK;
}

int sparc_hip_destroy(sparc_ctx* c) {
  if (!c) return SPARC_OK;
  (void)hipSetDevice(c->dev);
  if (c->st) (void)hipStreamSynchronize(c->st);
  free_all(c);
  delete c;
  return SPARC_OK;
}

int sparc_hip_set_background(sparc_ctx* c, const sparc_bg* bg) {
  if (!c || !bg) return fail(SPARC_E_INVALID, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  const long bgn = c->bg3d ? c->N : c->nz;
  const double* src[B_COUNT] = {bg->rho0, bg->rhoinv, bg->c2rho0, bg->gradrho0_z, bg->gradp0_z,
                                bg->spongexyz, bg->p0, bg->box, bg->boy, bg->boz, bg->curlbox,
                                bg->curlboy, bg->curlboz, bg->gradrho0_x, bg->gradrho0_y,
                                bg->gradp0_x, bg->gradp0_y, bg->reduction};
  const int nbg = c->mhd ? B_COUNT : B_P0 + 1;
  for (int s = 0; s < nbg; ++s) {
    if (!src[s]) return fail(SPARC_E_INVALID, "missing background array " + std::to_string(s));
    int rc = upload(c, c->bg[s], src[s], bgn);
    if (rc) return rc;
  }
  if (!bg->g) return fail(SPARC_E_INVALID, "missing g");
  int rc = upload(c, c->g, bg->g, c->nz);
  if (rc) return rc;
  for (int f = 0; f < 6; ++f) {
    FaceDev& F = c->face[f];
    if (!F.active) continue;
    if (!bg->pml_a[f] || !bg->pml_b[f] || !bg->pml_kappa[f])
      return fail(SPARC_E_INVALID, "missing PML coefficients for an active face");
    if ((rc = upload(c, F.a, bg->pml_a[f], F.fsize))) return rc;
    if ((rc = upload(c, F.b, bg->pml_b[f], F.fsize))) return rc;
    std::vector<double> kinv(F.fsize);  // the stage kernel multiplies by 1/kappa
    for (long q = 0; q < F.fsize; ++q) kinv[q] = 1.0 / bg->pml_kappa[f][q];
    if ((rc = upload(c, F.k, kinv.data(), F.fsize))) return rc;
  }
  if (c->cfg.pulse_dir) {
    if (!bg->src_x || !bg->src_y || !bg->src_z) return fail(SPARC_E_INVALID, "missing source profile");
    if ((rc = upload(c, c->sx, bg->src_x, c->nx))) return rc;
    if ((rc = upload(c, c->sy, bg->src_y, c->ny))) return rc;
    if ((rc = upload(c, c->sz, bg->src_z, c->nz))) return rc;
  }
  // The FP64 MHD xy kernel recomputes rhoinv as 1/rho0 when the host's array
  // holds exactly that (SPARC_RHOINV=load keeps reading it).
  {
    const char* ri = getenv("SPARC_RHOINV");
    bool same = !c->f32 && c->mhd && !(ri && std::string(ri) == "load");
    for (long i = 0; same && i < bgn; ++i) same = bg->rhoinv[i] == 1.0 / bg->rho0[i];
    c->rinv_rho0 = same;
  }
  c->dimc = bg->dimc;
  c->dimrho = bg->dimrho;
  c->bg_set = true;
  return SPARC_OK;
}

int sparc_hip_set_damping(sparc_ctx* c, const double* rates) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  HIPCHK(hipSetDevice(c->dev));
  if (c->damp) {
    HIPCHK(hipStreamSynchronize(c->st));
    damp_destroy(c->damp);
    c->damp = nullptr;
  }
  if (!rates) return SPARC_OK;
  c->damp = damp_create(c->nx, c->ny, c->nz, c->fs, c->f32, rates, c->st);
  if (!c->damp) return fail(SPARC_E_HIP, "FFT damping: plans or buffers could not be made");
  return SPARC_OK;
}

int sparc_hip_set_forcing(sparc_ctx* c, const double* map, const double* source_dep) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  HIPCHK(hipSetDevice(c->dev));
  if (!map) {
    c->forcing_on = false;
    return SPARC_OK;
  }
  int rc;
  if (!c->fmap && (rc = dalloc(&c->fmap, c->esz * (size_t)c->nx * c->ny))) return rc;
  if (!c->sdep && (rc = dalloc(&c->sdep, c->esz * (size_t)c->nz))) return rc;
  if (source_dep) {
    if ((rc = upload(c, c->sdep, source_dep, c->nz))) return rc;
  } else if (!c->forcing_on) {
    return fail(SPARC_E_INVALID, "the first forcing map needs source_dep");
  }
  if ((rc = upload(c, c->fmap, map, (long)c->nx * c->ny))) return rc;
  c->forcing_on = true;
  return SPARC_OK;
}

int sparc_hip_put_state(sparc_ctx* c, const double* a, const double* const psi[6]) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  HIPCHK(hipSetDevice(c->dev));
  if (c->f32)
    for (int p = 0; p < 2; ++p)
      if (launch_zero(c->comp[p], (long)c->nv * c->fs / 2, true, c->st) != hipSuccess)
        return fail(SPARC_E_HIP, "zero fill failed");
  if (a) {
    if (!c->f32) {
      for (int f = 0; f < c->nv; ++f) {
        int rc = upload(c, (char*)c->pool[c->ra] + c->esz * (size_t)f * c->fs, a + (long)f * c->N, c->N);
        if (rc) return rc;
      }
    } else {
      // FP32: the value is held as hi - comp (the Kahan pair get_state
      // reads back as a - comp), so an FP64 state -- one written from an
      // FP32 run included -- comes back with its compensation:
      // hi = fl32(x), comp = bf16(fl32(hi - x))
      std::vector<float> hi(c->N);
      std::vector<uint16_t> lo(c->N);
      for (int f = 0; f < c->nv; ++f) {
        const double* x = a + (long)f * c->N;
        for (long q = 0; q < c->N; ++q) {
          hi[q] = (float)x[q];
          lo[q] = f2bf((float)((double)hi[q] - x[q]));
        }
        HIPCHK(hipMemcpyAsync((char*)c->pool[c->ra] + 4 * (size_t)f * c->fs, hi.data(), 4 * c->N,
                              hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync((char*)c->comp[c->rc] + 2 * (size_t)f * c->fs, lo.data(), 2 * c->N,
                              hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
      }
    }
  }
  for (int f = 0; f < 6; ++f) {
    FaceDev& F = c->face[f];
    if (!F.active) continue;
    if (psi && psi[f]) {
      int rc = upload(c, F.psi, psi[f], c->npsi * F.fsize);
      if (rc) return rc;
    }
    if (launch_zero(F.pst, c->npsi * F.fsize, c->f32, c->st) != hipSuccess)
      return fail(SPARC_E_HIP, "zero fill failed");
  }
  for (int p : {c->rk0, c->rk1})
    if (launch_zero(c->pool[p], (long)c->nv * c->fs, c->f32, c->st) != hipSuccess)
      return fail(SPARC_E_HIP, "zero fill failed");
  HIPCHK(hipStreamSynchronize(c->st));
  // a new state starts with a clear guard (the step count runs on)
  const int clear[3] = {0, 0, 0};
  HIPCHK(hipMemcpy(c->guard, clear, sizeof clear, hipMemcpyHostToDevice));
  c->state_set = true;
  return SPARC_OK;
}

int sparc_hip_get_state(sparc_ctx* c, double* a, double* const psi[6]) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  HIPCHK(hipSetDevice(c->dev));
  if (a) {
    for (int f = 0; f < c->nv; ++f) {
      double* dst = a + (long)f * c->N;
      int rc = download(c, dst, (char*)c->pool[c->ra] + c->esz * (size_t)f * c->fs, c->N);
      if (rc) return rc;
      if (c->f32) {  // the compensated value is a - c
        std::vector<uint16_t> cc(c->N);
        rc = download_comp(c, cc.data(), (char*)c->comp[c->rc] + 2 * (size_t)f * c->fs, c->N);
        if (rc) return rc;
        for (long q = 0; q < c->N; ++q) dst[q] -= (double)bf2f(cc[q]);
      }
    }
  }
  for (int f = 0; f < 6; ++f) {
    FaceDev& F = c->face[f];
    if (!F.active || !psi || !psi[f]) continue;
    int rc = download(c, psi[f], F.psi, c->npsi * F.fsize);
    if (rc) return rc;
  }
  return SPARC_OK;
}

static void timed_begin(sparc_ctx* c, int kind, TimeRec* r) {
  if (!c->timing) return;
  r->kind = kind;
  (void)hipEventCreate(&r->e0);
  (void)hipEventCreate(&r->e1);
  (void)hipEventRecord(r->e0, c->st);
}
static void timed_end(sparc_ctx* c, TimeRec* r) {
  if (!c->timing) return;
  (void)hipEventRecord(r->e1, c->st);
  c->trecs.push_back(*r);
}

// Stage arguments common to all five stages of one run.
template <typename T>
static void fill_common(sparc_ctx* c, StageArgs<T>& A) {
  A = StageArgs<T>{};
  A.nx = c->nx;
  A.ny = c->ny;
  A.nz = c->nz;
  A.kc = c->kc;
  A.fs = c->fs;
  A.ntx = c->ntx;
  A.nty = c->nty;
  A.nblk_xy = c->ntx * c->nty;
  A.nchunk = c->nchunk;
  A.zkc = c->zkc;
  A.znchunk = c->znchunk;
  A.dma = c->dma && c->tx == 16 && c->ty == 16;
  A.periodic = c->cfg.periodic ? 1 : 0;
  A.pulse_dir = c->cfg.pulse_dir;
  A.rinv_rho0 = c->rinv_rho0 ? 1 : 0;
  A.dt = (T)c->cfg.deltat;
  A.sxs = (T)c->cfg.stretchx;
  A.sys = (T)c->cfg.stretchy;
  // FP32: the rest of each FP64 scale (0 in FP64)
  A.dt_lo = (T)(c->cfg.deltat - (double)A.dt);
  A.sxs_lo = (T)(c->cfg.stretchx - (double)A.sxs);
  A.sys_lo = (T)(c->cfg.stretchy - (double)A.sys);
  A.stretch_lo = (const T*)c->stretch_lo;
  auto B = [&](int s) { return (const T*)c->bg[s]; };
  A.rho0 = B(B_RHO0);
  A.rhoinv = B(B_RHOINV);
  A.c2rho0 = B(B_C2RHO0);
  A.gr0z = B(B_GR0Z);
  A.gp0z = B(B_GP0Z);
  A.sponge = B(B_SPONGE);
  A.p0 = B(B_P0);
  if (c->mhd) {
    A.box = B(B_BOX);
    A.boy = B(B_BOY);
    A.boz = B(B_BOZ);
    A.cbox = B(B_CBOX);
    A.cboy = B(B_CBOY);
    A.cboz = B(B_CBOZ);
    A.gr0x = B(B_GR0X);
    A.gr0y = B(B_GR0Y);
    A.gp0x = B(B_GP0X);
    A.gp0y = B(B_GP0Y);
    A.red = B(B_RED);
  }
  A.g = (const T*)c->g;
  A.stretch = (const T*)c->stretch;
  A.sx = (const T*)c->sx;
  A.sy = (const T*)c->sy;
  A.sz = (const T*)c->sz;
  for (int f = 0; f < 6; ++f) {
    const FaceDev& F = c->face[f];
    Face<T>& D = A.face[f];
    D.active = F.active;
    D.n = F.n;
    D.off = F.off;
    D.fsize = F.fsize;
    D.a = (const T*)F.a;
    D.b = (const T*)F.b;
    D.kinv = (const T*)F.k;
    D.psi = (T*)F.psi;
    D.pst = (T*)F.pst;
  }
  A.partial = c->partial;
  A.halt = c->guard;
  for (int f = 0; f < 8; ++f) {
    A.comp[f] = c->f32 ? (uint16_t*)c->comp[c->rc] + (size_t)f * c->fs : nullptr;
    A.a[f] = (T*)c->pool[c->ra] + (size_t)f * c->fs;
  }
  for (int f = 0; f < 6; ++f) A.dz[f] = c->dzbuf ? (T*)c->dzbuf + (size_t)f * c->fs : nullptr;
  for (int f = 0; f < 3; ++f) A.damp[f] = c->damp ? (const T*)damp_term_slab(c->damp, f) : nullptr;
  A.fmap = c->forcing_on ? (const T*)c->fmap : nullptr;
  A.sdep = (const T*)c->sdep;
}

// Stage s of LDDRK-5: u_1 = a; stage s reads u_s and writes
// u_{s+1} = a + beta_{s+1} K_s into the other stage buffer; stage 5 updates a.
template <typename T>
static void set_stage(sparc_ctx* c, StageArgs<T>& A, int s) {
  const int uin = s == 1 ? c->ra : (s % 2 == 0 ? c->rk0 : c->rk1);
  const int uout = s % 2 == 1 ? c->rk0 : c->rk1;
  A.first = s == 1;
  A.last = s == 5;
  A.beta_next = A.last ? (T)0 : (T)c->cfg.betas[s];
  A.beta_lo = A.last ? (T)0 : (T)(c->cfg.betas[s] - (double)A.beta_next);
  for (int f = 0; f < 8; ++f) {
    A.u[f] = (const T*)c->pool[uin] + (size_t)f * c->fs;
    A.un[f] = A.last ? A.a[f] : (T*)c->pool[uout] + (size_t)f * c->fs;
  }
}

static GuardArgs guard_args(const sparc_ctx* c) {
  GuardArgs g;
  g.guard = c->guard;
  g.gdiag = c->gdiag;
  g.npts = (double)c->nx * (double)c->ny * (double)c->nz;
  g.ke_scale = (c->dimc * c->dimc) * c->dimrho;
  return g;
}

template <typename T>
static int do_step(sparc_ctx* c, double amp) {
  StageArgs<T> A;
  fill_common<T>(c, A);
  A.src_amp = (T)amp;
  for (int s = 1; s <= 5; ++s) {
    set_stage<T>(c, A, s);
    TimeRec r;
    if (c->damp) {
      // damp_velocity of this stage's vx, vy, vz (RHS.f90:166-172)
      const hipError_t e = damp_terms(c->damp, A.u[0], c->st);
      if (e != hipSuccess) return fail(SPARC_E_HIP, std::string("damping transforms: ") + hipGetErrorString(e));
    }
    if (c->zp) {
      timed_begin(c, 4, &r);
      hipError_t e = launch_zpass<T>(A, c->st);
      timed_end(c, &r);
      if (e != hipSuccess) return fail(SPARC_E_HIP, std::string("z pass launch: ") + hipGetErrorString(e));
    }
    timed_begin(c, A.first ? 0 : (A.last ? 2 : 1), &r);
    hipError_t e = launch_stage<T>(c->mhd, c->bg3d, c->zp, c->vw, c->tx, c->ty, A, c->st);
    timed_end(c, &r);
    if (e != hipSuccess) return fail(SPARC_E_HIP, std::string("stage launch: ") + hipGetErrorString(e));
  }
  const int nblk = c->ntx * c->nty * c->nchunk;
  if (launch_diag_reduce(c->partial, nblk, c->diag, guard_args(c), c->st) != hipSuccess)
    return fail(SPARC_E_HIP, "diag reduce launch failed");
  c->steps_done++;
  return SPARC_OK;
}

int sparc_hip_step(sparc_ctx* c, double amp) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  if (!c->bg_set || !c->state_set) return fail(SPARC_E_STATE, "set_background and put_state must precede step");
  HIPCHK(hipSetDevice(c->dev));
  return c->f32 ? do_step<float>(c, amp) : do_step<double>(c, amp);
}

// One filter pass out of place, then the buffer roles swap.  axis 2: z;
// axis -1: the fused y-then-x pass.  The FP32 compensation terms (bfloat16)
// go through the same operator in FP32.
template <typename T>
static int filter_pass(sparc_ctx* c, int axis, int g, T s) {
  auto launch = [&](const void* in, void* out, bool cbf) -> hipError_t {
    if (axis < 0)
      return launch_filter_xy<T>(g, in, out, c->nx, c->ny, c->nz, c->fs, c->nv, c->cfg.periodic, s,
                                 c->st, c->guard, cbf);
    return launch_filter_zc<T>(g, in, out, c->nx, c->ny, c->nz, c->fs, c->nv, s, 0, c->st, c->guard,
                               false, cbf);
  };
  TimeRec r;
  timed_begin(c, 3, &r);
  hipError_t e = launch(c->pool[c->ra], c->pool[c->rk0], false);
  timed_end(c, &r);
  if (e != hipSuccess) return fail(SPARC_E_HIP, "filter launch failed");
  std::swap(c->ra, c->rk0);
  if (c->f32) {  // the compensation term is filtered with the same linear operator
    e = launch(c->comp[c->rc], c->comp[1 - c->rc], true);
    if (e != hipSuccess) return fail(SPARC_E_HIP, "filter launch failed");
    c->rc = 1 - c->rc;
  }
  return SPARC_OK;
}

template <typename T>
static int do_filter(sparc_ctx* c, int time) {
  int rc;
  if (time % c->cfg.xyfreq == 0) {
    if ((rc = filter_pass<T>(c, -1, c->cfg.g_xy_filt, (T)c->cfg.s_xy))) return rc;
  }
  if (time % c->cfg.zfreq == 0) {
    if ((rc = filter_pass<T>(c, 2, c->cfg.g_z_filt, (T)c->cfg.s_z))) return rc;
  }
  return SPARC_OK;
}

int sparc_hip_filter(sparc_ctx* c, int time) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  HIPCHK(hipSetDevice(c->dev));
  return c->f32 ? do_filter<float>(c, time) : do_filter<double>(c, time);
}

int sparc_hip_advance(sparc_ctx* c, int n0, int n1, const double* amp) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  if (n1 < n0) return SPARC_OK;
  if (c->cfg.pulse_dir && !amp) return fail(SPARC_E_INVALID, "source amplitudes required");
  for (int nn = n0; nn <= n1; ++nn) {
    int rc = sparc_hip_step(c, amp ? amp[nn - n0] : 0.0);
    if (rc) return rc;
    if (nn != 0 && (rc = sparc_hip_filter(c, nn))) return rc;
  }
  return SPARC_OK;
}

int sparc_hip_diag(sparc_ctx* c, double out[3]) {
  if (!c || !out) return fail(SPARC_E_INVALID, "null argument");
  if (c->steps_done == 0) return fail(SPARC_E_STATE, "no step taken yet");
  HIPCHK(hipSetDevice(c->dev));
  double h[3];
  HIPCHK(hipMemcpyAsync(h, c->diag, sizeof h, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  const double nrm = std::pow(h[0] / ((double)c->nx * (double)c->ny * (double)c->nz), 0.5);
  out[0] = nrm * (c->dimc * c->dimc) * c->dimrho;
  out[1] = h[1];
  out[2] = h[2];
  return SPARC_OK;
}

int sparc_hip_set_guard(sparc_ctx* c, int mode) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  if (mode != GUARD_OFF && mode != GUARD_RECORD && mode != GUARD_HALT)
    return fail(SPARC_E_INVALID, "guard mode must be 0 (off), 1 (record) or 2 (halt)");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  int g[G_WORDS];
  HIPCHK(hipMemcpy(g, c->guard, sizeof g, hipMemcpyDeviceToHost));
  g[G_HALT] = g[G_TRIPPED] = g[G_AT] = 0;
  g[G_MODE] = mode;
  HIPCHK(hipMemcpy(c->guard, g, sizeof g, hipMemcpyHostToDevice));
  c->guard_mode = mode;
  return SPARC_OK;
}

int sparc_hip_guard(sparc_ctx* c, int* tripped, long* at_step, long* steps, double diag[3]) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  int g[G_WORDS];
  double d[4];
  HIPCHK(hipMemcpy(g, c->guard, sizeof g, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(d, c->gdiag, sizeof d, hipMemcpyDeviceToHost));
  if (tripped) *tripped = g[G_TRIPPED];
  if (at_step) *at_step = g[G_TRIPPED] ? g[G_AT] : 0;
  if (steps) *steps = g[G_COUNT];
  if (diag) {
    if (g[G_TRIPPED]) {
      const double n = (double)c->nx * (double)c->ny * (double)c->nz;
      diag[0] = std::pow(d[0] / n, 0.5) * (c->dimc * c->dimc) * c->dimrho;
      diag[1] = d[1];
      diag[2] = d[2];
    } else {
      diag[0] = diag[1] = diag[2] = 0.0;
    }
  }
  return SPARC_OK;
}

static int op_common(sparc_ctx* c, const double* in, double* out) {
  if (!c || !in || !out) return fail(SPARC_E_INVALID, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  if (!c->op_buf) {
    int rc = dalloc((void**)&c->op_buf, sizeof(double) * 2 * c->N);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpyAsync(c->op_buf, in, sizeof(double) * c->N, hipMemcpyHostToDevice, c->st));
  return SPARC_OK;
}

static int op_finish(sparc_ctx* c, double* out) {
  HIPCHK(hipMemcpyAsync(out, c->op_buf + c->N, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SPARC_OK;
}

static int op_deriv(sparc_ctx* c, int axis, const double* in, double* out, int bc) {
  if (axis < 2 && bc != 1 && bc != 5) return fail(SPARC_E_INVALID, "bc must be 1 or 5");
  if (axis == 2 && bc != 1) return fail(SPARC_E_INVALID, "z derivatives take bc = 1 (z is never periodic)");
  int rc = op_common(c, in, out);
  if (rc) return rc;
  const double sxy = axis == 0 ? c->cfg.stretchx : c->cfg.stretchy;
  if (launch_deriv(axis, c->op_buf, c->op_buf + c->N, c->nx, c->ny, c->nz, sxy, c->stretch_d, bc,
                   c->st) != hipSuccess)
    return fail(SPARC_E_HIP, "derivative launch failed");
  return op_finish(c, out);
}

// ddx/ddy/ddz without any allocation: the same kernel as op_deriv, staged
// through this device's scratch (input at its start, output after
// kNoAllocMaxPoints).  Every element of the output slot is written, so no
// clearing is needed.  Calls on one device share the array: calls that
// overlap on one device (from several threads or contexts) race.
int sparc_hip_deriv_noalloc(sparc_ctx* c, int axis, const double* in, double* out, int bc) {
  if (!c || !in || !out) return fail(SPARC_E_INVALID, "null argument");
  if (axis < 0 || axis > 2) return fail(SPARC_E_INVALID, "axis must be 0 (x), 1 (y) or 2 (z)");
  if (axis < 2 && bc != 1 && bc != 5) return fail(SPARC_E_INVALID, "bc must be 1 or 5");
  if (axis == 2 && bc != 1) return fail(SPARC_E_INVALID, "z derivatives take bc = 1 (z is never periodic)");
  if (c->N > kNoAllocMaxPoints)
    return fail(SPARC_E_INVALID, "grid of " + std::to_string(c->N) + " points exceeds the " +
                                     std::to_string(kNoAllocMaxPoints) +
                                     "-point staging of sparc_hip_deriv_noalloc; use sparc_hip_ddx/ddy/ddz");
  if (c->dev < 0 || c->dev >= kMaxDevices) return fail(SPARC_E_INVALID, "device ordinal beyond the scratch table");
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch_ptr[c->dev];
  }
  if (!p) return fail(SPARC_E_STATE, "scratch address of the device was not set up");
  HIPCHK(hipSetDevice(c->dev));
  double* s_in = (double*)p;
  double* s_out = s_in + kNoAllocMaxPoints;
  HIPCHK(hipMemcpyAsync(s_in, in, sizeof(double) * c->N, hipMemcpyHostToDevice, c->st));
  const double sxy = axis == 0 ? c->cfg.stretchx : c->cfg.stretchy;
  if (launch_deriv(axis, s_in, s_out, c->nx, c->ny, c->nz, sxy, c->stretch_d, bc, c->st) != hipSuccess)
    return fail(SPARC_E_HIP, "derivative launch failed");
  HIPCHK(hipMemcpyAsync(out, s_out, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SPARC_OK;
}

int sparc_hip_ddx(sparc_ctx* c, const double* in, double* out, int bc) { return op_deriv(c, 0, in, out, bc); }
int sparc_hip_ddy(sparc_ctx* c, const double* in, double* out, int bc) { return op_deriv(c, 1, in, out, bc); }
int sparc_hip_ddz(sparc_ctx* c, const double* in, double* out, int bc) { return op_deriv(c, 2, in, out, bc); }

static int op_filt(sparc_ctx* c, int axis, const double* in, double* out) {
  int rc = op_common(c, in, out);
  if (rc) return rc;
  // the bare filtx/filty/filtz operators (FP64): no edge smoothing (that
  // belongs to FILTER_VARS) and the result is filt itself
  const int g = axis == 2 ? c->cfg.g_z_filt : c->cfg.g_xy_filt;
  if (launch_filter<double>(axis, g, c->op_buf, c->op_buf + c->N, c->nx, c->ny, c->nz, c->N, 1,
                            c->cfg.periodic, 0.0, 1, c->st, nullptr) != hipSuccess)
    return fail(SPARC_E_HIP, "filter launch failed");
  return op_finish(c, out);
}

int sparc_hip_filtx(sparc_ctx* c, const double* in, double* out) { return op_filt(c, 0, in, out); }
int sparc_hip_filty(sparc_ctx* c, const double* in, double* out) { return op_filt(c, 1, in, out); }
int sparc_hip_filtz(sparc_ctx* c, const double* in, double* out) { return op_filt(c, 2, in, out); }

int sparc_hip_sync(sparc_ctx* c) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  return SPARC_OK;
}

int sparc_hip_timing(sparc_ctx* c, int enable) {
  if (!c) return fail(SPARC_E_INVALID, "null context");
  c->timing = enable != 0;
  return SPARC_OK;
}

int sparc_hip_kernel_time(sparc_ctx* c, int kind, long* count, double* total_ms) {
  if (!c || kind < 0 || kind > 4 || !count || !total_ms) return fail(SPARC_E_INVALID, "bad argument");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  for (auto& r : c->trecs) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.e0, r.e1));
    c->tcount[r.kind]++;
    c->tms[r.kind] += ms;
    (void)hipEventDestroy(r.e0);
    (void)hipEventDestroy(r.e1);
  }
  c->trecs.clear();
  *count = c->tcount[kind];
  *total_ms = c->tms[kind];
  return SPARC_OK;
}

// Re-tile a context between steps (the partial-sum buffer grows if needed).
static int retile(sparc_ctx* c, int tx, int ty, int kc) {
  if (!tile_fits(c, tx, ty)) return fail(SPARC_E_INVALID, "tile shape unsupported or does not fit the grid");
  set_geometry(c, tx, ty, kc);
  const long nblk = (long)c->ntx * c->nty * c->nchunk;
  if (nblk > c->partial_cap) {
    if (c->partial) (void)hipFree(c->partial);
    c->partial = nullptr;
    c->partial_cap = 0;
    int rc = dalloc((void**)&c->partial, sizeof(double) * 3 * nblk);
    if (rc) return rc;
    c->partial_cap = nblk;
  }
  return SPARC_OK;
}

int sparc_hip_set_tile(sparc_ctx* c, const int tile[3]) {
  if (!c || !tile) return fail(SPARC_E_INVALID, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  return retile(c, tile[0], tile[1], tile[2]);
}

// Average time of one middle RK stage (stage 2: reads the first stage buffer,
// writes the second, the CPML stage values and, for MHD, the z-pass slabs --
// all overwritten by the next step, so the run's state is untouched).
template <typename T>
static int time_middle_stage(sparc_ctx* c, int reps, double* ms) {
  StageArgs<T> A;
  fill_common<T>(c, A);
  set_stage<T>(c, A, 2);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  hipError_t e = hipSuccess;
  for (int q = 0; q <= reps && e == hipSuccess; ++q) {
    if (q == 1) e = hipEventRecord(e0, c->st);
    if (e == hipSuccess && c->zp) e = launch_zpass<T>(A, c->st);
    if (e == hipSuccess) e = launch_stage<T>(c->mhd, c->bg3d, c->zp, c->vw, c->tx, c->ty, A, c->st);
  }
  if (e == hipSuccess) e = hipEventRecord(e1, c->st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(SPARC_E_HIP, std::string("autotune launch: ") + hipGetErrorString(e));
  *ms = t / reps;
  return SPARC_OK;
}

int sparc_hip_autotune(sparc_ctx* c, int reps, int tile_out[3], double* ms_out) {
  if (!c || !tile_out || !ms_out || reps < 1) return fail(SPARC_E_INVALID, "bad argument");
  if (!c->bg_set || !c->state_set) return fail(SPARC_E_STATE, "set_background and put_state must precede autotune");
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->st));
  static const int shapes[5][2] = {{16, 16}, {32, 8}, {32, 16}, {64, 4}, {64, 8}};
  int best[3] = {c->tx, c->ty, c->kc};
  double best_ms = 1e30;
  for (const auto& sh : shapes) {
    if (!tile_fits(c, sh[0], sh[1])) continue;
    int rc = retile(c, sh[0], sh[1], 0);
    if (rc) return rc;
    const int kauto = c->kc;
    // the automatic z chunk, a half and a quarter of it (more blocks: small
    // grids run fastest at 1-2 planes) and twice it
    const int kcs[4] = {kauto, std::max(1, kauto / 2), std::max(1, kauto / 4), std::min(c->nz, kauto * 2)};
    for (int q = 0; q < 4; ++q) {
      const int kc = kcs[q];
      if (q > 0 && (kc == kauto || kc == kcs[q - 1])) continue;
      if ((rc = retile(c, sh[0], sh[1], kc))) return rc;
      double ms = 0.0;
      rc = c->f32 ? time_middle_stage<float>(c, reps, &ms) : time_middle_stage<double>(c, reps, &ms);
      if (rc) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best[0] = sh[0];
        best[1] = sh[1];
        best[2] = kc;
      }
    }
  }
  int rc = retile(c, best[0], best[1], best[2]);
  if (rc) return rc;
  tile_out[0] = best[0];
  tile_out[1] = best[1];
  tile_out[2] = best[2];
  *ms_out = best_ms;
  return SPARC_OK;
}

int sparc_hip_tile(sparc_ctx* c, int tile[3], long* nblocks) {
  if (!c || !tile || !nblocks) return fail(SPARC_E_INVALID, "null argument");
  tile[0] = c->tx;
  tile[1] = c->ty;
  tile[2] = c->kc;
  *nblocks = (long)c->ntx * c->nty * c->nchunk;
  return SPARC_OK;
}


// ===========================================================================
// Batched form: many independent runs (contexts) advanced together.  Runs are
// grouped by kernel variant and tile shape; each RK stage is one persistent
// launch per group over the work items of all its runs.  Stage arguments live
// in device memory, the source amplitude comes from a per-run device table
// indexed by a device step counter, so one step is the same launch sequence
// every time and can be captured in a HIP graph.
// ===========================================================================
namespace {
struct BatchGroup {
  bool mhd = false, bg3d = false, zp = false, dma = false;
  int vw = 1, tx = 0, ty = 0, nblk = 0, znblk = 0;
  std::vector<int> members;
  std::vector<int> zkc;        // z-pass chunk depth per member in this batch
  void* args[5] = {};         // StageArgs<T>[members] per stage (device)
  int* item_start = nullptr;   // prefix of stage work items (device)
  int* zitem_start = nullptr;  // prefix of z-pass work items (device)
};
}  // namespace

struct sparc_batch {
  int dev = 0;
  hipStream_t st = nullptr;
  bool f32 = false, use_graph = false;
  std::vector<sparc_ctx*> runs;
  std::vector<BatchGroup> groups;
  int* tstep = nullptr;
  // work counters of the persistent launches (8 per-XCD item counters and a
  // count of the blocks that left; each launch leaves them zero)
  int* work = nullptr;
  std::vector<void*> amp;  // per run, amp_cap entries (device)
  int amp_cap = 0;
  void* jobs = nullptr;    // DiagJob[nrun] (device)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

static void batch_free(sparc_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto& g : b->groups) {
    for (auto& p : g.args)
      if (p) (void)hipFree(p);
    if (g.item_start) (void)hipFree(g.item_start);
    if (g.zitem_start) (void)hipFree(g.zitem_start);
  }
  for (auto& p : b->amp)
    if (p) (void)hipFree(p);
  if (b->tstep) (void)hipFree(b->tstep);
  if (b->work) (void)hipFree(b->work);
  if (b->jobs) (void)hipFree(b->jobs);
  if (b->st) (void)hipStreamDestroy(b->st);
}

template <typename T>
static int batch_upload_args(sparc_batch* b) {
  for (auto& g : b->groups) {
    for (int s = 1; s <= 5; ++s) {
      std::vector<StageArgs<T>> h(g.members.size());
      for (size_t q = 0; q < g.members.size(); ++q) {
        const int r = g.members[q];
        sparc_ctx* c = b->runs[r];
        fill_common<T>(c, h[q]);
        set_stage<T>(c, h[q], s);
        if (g.zp) {
          h[q].zkc = g.zkc[q];
          h[q].znchunk = (c->nz + g.zkc[q] - 1) / g.zkc[q];
        }
        h[q].amp_tab = c->cfg.pulse_dir ? (const T*)b->amp[r] : nullptr;
        h[q].tstep = b->tstep;
      }
      HIPCHK(hipMemcpy(g.args[s - 1], h.data(), sizeof(StageArgs<T>) * h.size(), hipMemcpyHostToDevice));
    }
  }
  return SPARC_OK;
}

template <typename T>
static int batch_enqueue_step(sparc_batch* b) {
  for (int s = 1; s <= 5; ++s)
    for (auto& g : b->groups) {
      const StageArgs<T>* args = (const StageArgs<T>*)g.args[s - 1];
      const int nm = (int)g.members.size();
      hipError_t e = hipSuccess;
      if (g.zp) e = launch_zpass_batch<T>(args, g.zitem_start, nm, g.znblk, b->work, b->st);
      if (e == hipSuccess)
        e = launch_stage_batch<T>(g.mhd, g.bg3d, g.zp, g.vw, g.tx, g.ty, args, g.item_start, nm, g.nblk,
                                  b->work, b->st, g.dma);
      if (e != hipSuccess) return fail(SPARC_E_HIP, std::string("batch stage launch: ") + hipGetErrorString(e));
    }
  if (launch_diag_reduce_batch(b->jobs, (int)b->runs.size(), b->st) != hipSuccess ||
      launch_tick(b->tstep, 0, b->st) != hipSuccess)
    return fail(SPARC_E_HIP, "batch diag/tick launch failed");
  return SPARC_OK;
}

// FILTER_VARS for one run on the batch stream, ending in the run's own
// buffers (the device stage arguments keep pointing at them): the fused xy
// pass into the spare buffer, then the z pass back out of place when both
// cadences meet; the xy pass alone is copied back; the z pass alone runs in
// place (one chunk per column).
template <typename T>
static int batch_filter(sparc_batch* b, sparc_ctx* c, int time) {
  const long n = (long)c->nv * c->fs;
  const bool xy = time % c->cfg.xyfreq == 0, z = time % c->cfg.zfreq == 0;
  // cbf: the FP32 compensation terms (bfloat16, copied as n/2 4-byte words)
  auto run = [&](void* a, void* t, bool cbf) -> hipError_t {
    hipError_t e = hipSuccess;
    if (xy)
      e = launch_filter_xy<T>(c->cfg.g_xy_filt, a, t, c->nx, c->ny, c->nz, c->fs, c->nv,
                              c->cfg.periodic, (T)c->cfg.s_xy, b->st, c->guard, cbf);
    if (e == hipSuccess && z)
      e = launch_filter_zc<T>(c->cfg.g_z_filt, xy ? t : a, a, c->nx, c->ny, c->nz, c->fs, c->nv,
                              (T)c->cfg.s_z, 0, b->st, c->guard, !xy, cbf);
    else if (e == hipSuccess && xy)
      e = cbf ? launch_copy(t, a, n / 2, true, b->st, c->guard) : launch_copy(t, a, n, c->f32, b->st, c->guard);
    return e;
  };
  hipError_t e = run(c->pool[c->ra], c->pool[c->rk0], false);
  if (e == hipSuccess && c->f32) e = run(c->comp[c->rc], c->comp[1 - c->rc], true);
  if (e != hipSuccess) return fail(SPARC_E_HIP, std::string("batch filter: ") + hipGetErrorString(e));
  return SPARC_OK;
}

// The batched MHD xy kernel stages with direct-to-LDS loads in FP64 only:
// the FP32 persistent kernel then needs 256 registers and spills (the
// 64-run population: 2.75 -> 2.56 GPt/s; FP64 batch 1.24 -> 1.29 GPt/s).
static bool batch_dma(const sparc_ctx* c) { return c->dma && !c->f32 && c->tx == 16 && c->ty == 16; }

int sparc_hip_batch_create(sparc_ctx* const* runs, int nrun, int use_graph, sparc_batch** out) {
  if (!runs || nrun < 1 || !out) return fail(SPARC_E_INVALID, "batch needs at least one run");
  *out = nullptr;
  for (int r = 0; r < nrun; ++r) {
    if (!runs[r]) return fail(SPARC_E_INVALID, "null run in batch");
    if (!runs[r]->bg_set || !runs[r]->state_set)
      return fail(SPARC_E_STATE, "every run needs set_background and put_state first");
    if (runs[r]->dev != runs[0]->dev || runs[r]->f32 != runs[0]->f32)
      return fail(SPARC_E_INVALID, "batched runs must share device and precision");
    for (int q = 0; q < r; ++q)
      if (runs[q] == runs[r]) return fail(SPARC_E_INVALID, "a run appears twice in the batch");
    if (runs[r]->damp)
      return fail(SPARC_E_INVALID, "runs with FFT damping are advanced alone, not batched");
    if (runs[r]->forcing_on)
      return fail(SPARC_E_INVALID, "runs driven by a forcing map are advanced alone, not batched");
  }
  sparc_batch* b = new sparc_batch();
  b->dev = runs[0]->dev;
  b->f32 = runs[0]->f32;
  b->use_graph = use_graph != 0;
  b->runs.assign(runs, runs + nrun);
  auto bail = [&](int code) {
    batch_free(b);
    delete b;
    return code;
  };
  if (hipSetDevice(b->dev) != hipSuccess) return bail(fail(SPARC_E_HIP, "hipSetDevice failed"));
  for (auto* c : b->runs)
    if (hipStreamSynchronize(c->st) != hipSuccess) return bail(fail(SPARC_E_HIP, "run stream sync failed"));
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SPARC_E_HIP, "hipStreamCreate failed"));
  int ncu = 256;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->dev);
  for (int r = 0; r < nrun; ++r) {
    sparc_ctx* c = b->runs[r];
    BatchGroup* g = nullptr;
    for (auto& h : b->groups)
      if (h.mhd == c->mhd && h.bg3d == c->bg3d && h.zp == c->zp && h.vw == c->vw && h.tx == c->tx &&
          h.ty == c->ty && h.dma == batch_dma(c))
        g = &h;
    if (!g) {
      b->groups.emplace_back();
      g = &b->groups.back();
      g->mhd = c->mhd;
      g->bg3d = c->bg3d;
      g->zp = c->zp;
      g->vw = c->vw;
      g->tx = c->tx;
      g->ty = c->ty;
      g->dma = batch_dma(c);
    }
    g->members.push_back(r);
  }
  const size_t asz = b->f32 ? sizeof(StageArgs<float>) : sizeof(StageArgs<double>);
  int rc;
  for (auto& g : b->groups) {
    std::vector<int> start(g.members.size() + 1, 0);
    for (size_t q = 0; q < g.members.size(); ++q) {
      const sparc_ctx* c = b->runs[g.members[q]];
      start[q + 1] = start[q] + c->ntx * c->nty * c->nchunk;
    }
    for (int s = 0; s < 5; ++s)
      if ((rc = dalloc(&g.args[s], asz * g.members.size()))) return bail(rc);
    if ((rc = dalloc((void**)&g.item_start, sizeof(int) * start.size()))) return bail(rc);
    if (hipMemcpy(g.item_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(SPARC_E_HIP, "item table upload failed"));
    g.nblk = std::min(stage_batch_blocks(b->f32, g.mhd, g.bg3d, g.zp, g.vw, g.tx, g.ty, ncu, g.dma),
                      (start.back() + 7) / 8 * 8);
    if (g.zp) {
      // z-pass chunks sized for the whole group rather than per run: about
      // twice the resident grid in items, each chunk >= 48 planes deep (the
      // 10-plane queue warm-up is re-read per chunk)
      g.znblk = zpass_batch_blocks(b->f32, ncu);
      long ncols = 0;
      for (int r : g.members) ncols += zpass_items(b->runs[r]->nx, b->runs[r]->ny, 1);
      const long per_run = std::max<long>(1, (2L * g.znblk + ncols - 1) / ncols);
      std::vector<int> zs(g.members.size() + 1, 0);
      g.zkc.assign(g.members.size(), 0);
      for (size_t q = 0; q < g.members.size(); ++q) {
        const sparc_ctx* c = b->runs[g.members[q]];
        const int kc = (int)std::min<long>(c->nz, std::max<long>(48, (c->nz + per_run - 1) / per_run));
        g.zkc[q] = kc;
        zs[q + 1] = zs[q] + (int)zpass_items(c->nx, c->ny, (c->nz + kc - 1) / kc);
      }
      if ((rc = dalloc((void**)&g.zitem_start, sizeof(int) * zs.size()))) return bail(rc);
      if (hipMemcpy(g.zitem_start, zs.data(), sizeof(int) * zs.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SPARC_E_HIP, "item table upload failed"));
      g.znblk = std::min(zpass_batch_blocks(b->f32, ncu), (zs.back() + 7) / 8 * 8);
    }
  }
  if ((rc = dalloc((void**)&b->tstep, sizeof(int)))) return bail(rc);
  if ((rc = dalloc((void**)&b->work, 16 * sizeof(int)))) return bail(rc);
  if (launch_zero(b->work, 16, true, b->st) != hipSuccess || hipStreamSynchronize(b->st) != hipSuccess)
    return bail(fail(SPARC_E_HIP, "work counter clear failed"));
  std::vector<DiagJob> jobs(nrun);
  for (int r = 0; r < nrun; ++r) {
    sparc_ctx* c = b->runs[r];
    jobs[r].partial = c->partial;
    jobs[r].out = c->diag;
    jobs[r].nblk = c->ntx * c->nty * c->nchunk;
    jobs[r].pad_ = 0;
    jobs[r].g = guard_args(c);
  }
  if ((rc = dalloc(&b->jobs, sizeof(DiagJob) * nrun))) return bail(rc);
  if (hipMemcpy(b->jobs, jobs.data(), sizeof(DiagJob) * nrun, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(SPARC_E_HIP, "diag job upload failed"));
  b->amp.assign(nrun, nullptr);
  *out = b;
  return SPARC_OK;
}

int sparc_hip_batch_destroy(sparc_batch* b) {
  if (!b) return SPARC_OK;
  (void)hipSetDevice(b->dev);
  if (b->st) (void)hipStreamSynchronize(b->st);
  batch_free(b);
  delete b;
  return SPARC_OK;
}

template <typename T>
static int batch_advance(sparc_batch* b, int nsteps, const int* time0, const double* amps, int flags) {
  const int nrun = (int)b->runs.size();
  int rc;
  if (nsteps > b->amp_cap) {  // grow the amplitude tables (device pointers change)
    for (auto& p : b->amp)
      if (p) {
        (void)hipFree(p);
        p = nullptr;
      }
    for (int r = 0; r < nrun; ++r)
      if ((rc = dalloc(&b->amp[r], sizeof(T) * nsteps))) return rc;
    b->amp_cap = nsteps;
  }
  for (int r = 0; r < nrun; ++r) {
    if (!b->runs[r]->cfg.pulse_dir) continue;
    std::vector<T> h(nsteps);
    for (int n = 0; n < nsteps; ++n) h[n] = (T)amps[(long)r * nsteps + n];
    HIPCHK(hipMemcpy(b->amp[r], h.data(), sizeof(T) * nsteps, hipMemcpyHostToDevice));
  }
  if ((rc = batch_upload_args<T>(b))) return rc;
  if (launch_tick(b->tstep, 1, b->st) != hipSuccess) return fail(SPARC_E_HIP, "tick launch failed");
  if (b->use_graph && !b->gexec) {
    HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    rc = batch_enqueue_step<T>(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->st, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SPARC_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  }
  for (int n = 0; n < nsteps; ++n) {
    if (b->gexec) {
      HIPCHK(hipGraphLaunch(b->gexec, b->st));
    } else if ((rc = batch_enqueue_step<T>(b))) {
      return rc;
    }
    // SPARC_DEFER_FILTER: the last step's filters wait for
    // sparc_hip_batch_filter, so the caller can read the state the reference
    // writes out, after STEP and before FILTER_VARS (driver.f90:198-200)
    if (n == nsteps - 1 && (flags & SPARC_DEFER_FILTER)) break;
    for (int r = 0; r < nrun; ++r) {
      const int t = time0[r] + n;
      sparc_ctx* c = b->runs[r];
      if (t != 0 && (t % c->cfg.xyfreq == 0 || t % c->cfg.zfreq == 0))
        if ((rc = batch_filter<T>(b, c, t))) return rc;
    }
  }
  HIPCHK(hipStreamSynchronize(b->st));
  for (auto* c : b->runs) c->steps_done += nsteps;
  return SPARC_OK;
}

int sparc_hip_batch_advance(sparc_batch* b, int nsteps, const int* time0, const double* amps,
                            int flags) {
  if (!b || nsteps < 0 || (nsteps > 0 && (!time0 || !amps))) return fail(SPARC_E_INVALID, "bad argument");
  if (flags & ~SPARC_DEFER_FILTER) return fail(SPARC_E_INVALID, "unknown batch advance flag");
  if (nsteps == 0) return SPARC_OK;
  HIPCHK(hipSetDevice(b->dev));
  return b->f32 ? batch_advance<float>(b, nsteps, time0, amps, flags)
                : batch_advance<double>(b, nsteps, time0, amps, flags);
}

template <typename T>
static int batch_filter_all(sparc_batch* b, const int* time) {
  int rc;
  for (size_t r = 0; r < b->runs.size(); ++r) {
    sparc_ctx* c = b->runs[r];
    const int t = time[r];
    if (t != 0 && (t % c->cfg.xyfreq == 0 || t % c->cfg.zfreq == 0))
      if ((rc = batch_filter<T>(b, c, t))) return rc;
  }
  HIPCHK(hipStreamSynchronize(b->st));
  return SPARC_OK;
}

int sparc_hip_batch_filter(sparc_batch* b, const int* time) {
  if (!b || !time) return fail(SPARC_E_INVALID, "null argument");
  HIPCHK(hipSetDevice(b->dev));
  return b->f32 ? batch_filter_all<float>(b, time) : batch_filter_all<double>(b, time);
}

// Average time of one middle RK stage of the whole batch (every group's z
// pass and stage launch with the stage-2 arguments).  Like the single-run
// autotune timing it writes only buffers the next step overwrites.
template <typename T>
static int batch_time_stage_t(sparc_batch* b, int reps, double* ms) {
  int rc = batch_upload_args<T>(b);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  hipError_t e = hipSuccess;
  for (int q = 0; q <= reps && e == hipSuccess; ++q) {
    if (q == 1) e = hipEventRecord(e0, b->st);
    for (auto& g : b->groups) {
      if (e != hipSuccess) break;
      const StageArgs<T>* args = (const StageArgs<T>*)g.args[1];
      const int nm = (int)g.members.size();
      if (g.zp) e = launch_zpass_batch<T>(args, g.zitem_start, nm, g.znblk, b->work, b->st);
      if (e == hipSuccess)
        e = launch_stage_batch<T>(g.mhd, g.bg3d, g.zp, g.vw, g.tx, g.ty, args, g.item_start, nm, g.nblk,
                                  b->work, b->st, g.dma);
    }
  }
  if (e == hipSuccess) e = hipEventRecord(e1, b->st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(SPARC_E_HIP, std::string("batch timing: ") + hipGetErrorString(e));
  *ms = t / reps;
  return SPARC_OK;
}

int sparc_hip_batch_time_stage(sparc_batch* b, int reps, double* ms_out) {
  if (!b || !ms_out || reps < 1) return fail(SPARC_E_INVALID, "bad argument");
  HIPCHK(hipSetDevice(b->dev));
  return b->f32 ? batch_time_stage_t<float>(b, reps, ms_out) : batch_time_stage_t<double>(b, reps, ms_out);
}

int sparc_hip_batch_info(sparc_batch* b, int* ngroups, long* nlaunch_blocks) {
  if (!b || !ngroups || !nlaunch_blocks) return fail(SPARC_E_INVALID, "null argument");
  *ngroups = (int)b->groups.size();
  long n = 0;
  for (auto& g : b->groups) n += g.nblk;
  *nlaunch_blocks = n;
  return SPARC_OK;
}

// ===========================================================================
// Set-up helper and read-out
// ===========================================================================

// Decimal search of the Alfven limiter (initialize.f90:962-986) with the
// maximum over the grid evaluated on the device.  red holds the ratio
// |B0|^2/(rho0 cq^2) of initialize.f90:963, b2 = |B0|^2, all nondimensional.
int sparc_hip_alfven_search(sparc_ctx* c, const double* red, const double* b2, const double* rho0,
                            double dimb, double dimrho, double maxva, double* temp_out) {
  if (!c || !red || !b2 || !rho0 || !temp_out) return fail(SPARC_E_INVALID, "null argument");
  HIPCHK(hipSetDevice(c->dev));
  const long n = c->N;
  double rmax = -INFINITY;
  for (long i = 0; i < n; ++i)
    if (red[i] > rmax) rmax = red[i];
  if (!(rmax > 0)) {
    *temp_out = 1.0e10;
    return SPARC_OK;
  }
  double* d = nullptr;
  const int nblk = 1024;
  int rc = dalloc((void**)&d, sizeof(double) * (3 * n + nblk));
  if (rc) return rc;
  std::vector<double> part(nblk);
  auto cleanup = [&](int code) {
    (void)hipFree(d);
    return code;
  };
  if (hipMemcpy(d, red, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + n, b2, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + 2 * n, rho0, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(SPARC_E_HIP, "alfven search upload failed"));
  const double PI_REF = 3.14159265358979;
  const double den0 = 4.0 * PI_REF * dimrho;
  auto va_of = [&](double T, double* va) -> int {
    if (launch_alfven_max(d, d + n, d + 2 * n, n, dimb * dimb, T, den0, d + 3 * n, nblk, c->st) != hipSuccess ||
        hipMemcpyAsync(part.data(), d + 3 * n, sizeof(double) * nblk, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
      return fail(SPARC_E_HIP, "alfven search launch failed");
    double m = -1.0e300;
    for (double v : part)
      if (v > m) m = v;
    *va = m;
    return SPARC_OK;
  };
  double va = 0.0, dtemp = 1.0e9, temp = -dtemp;
  while ((dtemp > 1.0e-5) && (temp < 1.0e10 - 1.1)) {
    while ((va < maxva) && (temp < 1.0e10)) {
      temp = temp + dtemp;
      if ((rc = va_of(temp, &va))) return cleanup(rc);
    }
    temp = temp - dtemp;
    dtemp = dtemp / 10.0;
    va = 0.0;
  }
  *temp_out = temp;
  return cleanup(SPARC_OK);
}

// Planes k0..k1 (inclusive) of field f of the current state, FP64, x fastest
// -- the chunk output of WRITE_OUT_CHUNK (all_modules.f90:1566-1799) for a
// full-width slab, e.g. the photospheric vz slice (zcb = zct).
int sparc_hip_get_slab(sparc_ctx* c, int field, int k0, int k1, double* out) {
  if (!c || !out || field < 0 || field >= c->nv || k0 < 0 || k1 < k0 || k1 >= c->nz)
    return fail(SPARC_E_INVALID, "bad slab request");
  HIPCHK(hipSetDevice(c->dev));
  const long nxy = (long)c->nx * c->ny;
  const long n = nxy * (k1 - k0 + 1);
  const size_t off = (size_t)field * c->fs + (size_t)k0 * nxy;
  int rc = download(c, out, (const char*)c->pool[c->ra] + c->esz * off, n);
  if (rc) return rc;
  if (c->f32) {
    std::vector<uint16_t> cc(n);
    rc = download_comp(c, cc.data(), (const char*)c->comp[c->rc] + 2 * off, n);
    if (rc) return rc;
    for (long q = 0; q < n; ++q) out[q] -= (double)bf2f(cc[q]);
  }
  return SPARC_OK;
}

// ===========================================================================
// Measured bandwidth ceiling (no counterpart in the reference)
// ===========================================================================
int sparc_hip_stream_bench(int device, long n, int nread, int reps, double* gbps) {
  if (!gbps || n < 2 || n % 2 || nread < 1 || nread > 48 || reps < 1)
    return fail(SPARC_E_INVALID, "bad argument");
  HIPCHK(hipSetDevice(device));
  std::vector<double*> buf(nread + 1, nullptr);
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SPARC_OK;
  auto cleanup = [&]() {
    for (auto* p : buf)
      if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  };
  for (auto& p : buf)
    if ((rc = dalloc((void**)&p, sizeof(double) * n))) {
      cleanup();
      return rc;
    }
  int ncu = 256;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  for (auto* p : buf)
    if (e == hipSuccess) e = launch_zero(p, n, false, st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const int nblk = ncu * 8;
  for (int q = 0; q <= reps && e == hipSuccess; ++q) {
    if (q == 1) e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = launch_stream(buf.data(), nread, buf[nread], n, nblk, st);
  }
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  cleanup();
  if (e != hipSuccess) return fail(SPARC_E_HIP, std::string("stream probe: ") + hipGetErrorString(e));
  *gbps = (double)(nread + 1) * 8.0 * (double)n / (ms * 1e-3 / reps) / 1e9;
  return SPARC_OK;
}

// Device memory (free, total) in bytes: the served form's HBM budget.
int sparc_hip_device_memory(int device, long* free_bytes, long* total_bytes) {
  if (!free_bytes || !total_bytes) return fail(SPARC_E_INVALID, "null argument");
  HIPCHK(hipSetDevice(device));
  size_t f = 0, t = 0;
  HIPCHK(hipMemGetInfo(&f, &t));
  *free_bytes = (long)f;
  *total_bytes = (long)t;
  return SPARC_OK;
}
