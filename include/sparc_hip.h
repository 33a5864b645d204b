/*
 * sparc_hip.h -- C ABI of the MI355X (gfx950) SPARC2 time-step library.
 *
 * The reference's hot path is a set of argument-less Fortran module
 * procedures working on module globals:
 *   STEP               step.f90:20-63          (called at driver.f90:157)
 *   MP_QUIET_3D        RHS.f90:18-175          (called at step.f90:40)
 *   MP_MHD_3D          RHS.f90:363-668         (called at step.f90:44)
 *   ddx / ddy / ddz    derivatives.f90:10,226,451
 *   filtx/filty/filtz  derivatives.f90:119,345,565
 *   FILTER_VARS        derivatives.f90:659-703 (called at driver.f90:200)
 *   norm + guards      all_modules.f90:684-706, driver.f90:161-183
 * This header replaces them with a context object that owns the device
 * copies of everything those routines read or write.  All host arrays are
 * FP64 in the reference's Fortran order (x fastest) and are borrowed only
 * for the duration of a call.  Every entry point returns SPARC_OK (0) or a
 * negative error code; sparc_hip_strerror() names the code and
 * sparc_hip_last_error() gives the message of the most recent failure.
 * Calls on one context must be serialised by the caller.
 */
#ifndef SPARC_HIP_H
#define SPARC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SPARC_OK 0
#define SPARC_E_INVALID (-1) /* bad argument or configuration */
#define SPARC_E_HIP (-2)     /* HIP runtime error (no device, launch failure, ...) */
#define SPARC_E_NOMEM (-3)   /* device allocation failed */
#define SPARC_E_STATE (-4)   /* call out of order (e.g. step before background) */

#define SPARC_PREC_FP64 0 /* plain FP64 path */
#define SPARC_PREC_FP32 1 /* FP32 fields/stencils/CPML, compensated RK update */

/* PML faces, the order of psi arguments below. */
#define SPARC_XBOT 0
#define SPARC_XTOP 1
#define SPARC_YBOT 2
#define SPARC_YTOP 3
#define SPARC_ZBOT 4
#define SPARC_ZTOP 5

typedef struct sparc_ctx sparc_ctx;

/* Run geometry and switches.  Values are the ones the reference derives in
 * READ_SPARC_PARAMETERS / INITANDALLOCATE / INITIALIZE_STEP. */
typedef struct sparc_cfg {
  int nx, ny, nz;        /* global grid (nx, ny, nz >= 16) */
  int option;            /* 1 = MP_QUIET_3D (quiet/thermal), 3 = MP_MHD_3D;
                            initialize.f90:1343-1363 */
  int precision;         /* SPARC_PREC_FP64 or SPARC_PREC_FP32 */
  int periodic;          /* PERIODIC: x/y periodic */
  int use_pml, use_hpml; /* USE_PML (z faces), USE_HPML (x/y faces) */
  int npml[6];           /* nxpmlbot, nxpmltop, nypmlbot, nypmltop, nzpmlbot, nzpmltop */
  int g_xy_filt, g_z_filt; /* filter half-widths, 5 or 3 */
  double s_xy, s_z;      /* filter strengths (used when g = 3) */
  int xyfreq, zfreq;     /* filter cadences in steps, initialize.f90:1911-1912 */
  int pulse_dir;         /* 0 no source, 1 vx, 2 vy, 3 vz */
  int bg3d;              /* 1: background arrays are nx*ny*nz; 0: nz-long z profiles
                            (quiet backgrounds, all_modules.f90:353-384) */
  double deltat;         /* nondimensional dt, initialize.f90:733 */
  double betas[5];       /* RK stage weights x deltat, initialize.f90:1366-1381 */
  double stretchx, stretchy; /* initialize.f90:503-517 */
  const double* stretch; /* nz values, initialize.f90:728-730 */
  int device;            /* HIP device ordinal */
  int tile[3];           /* stage tile (x, y, z-chunk); 0 = library choice */
} sparc_cfg;

/* Backgrounds after INITIALISE_RHS (initialize.f90:883-1228), nondimensional.
 * The first group is nz-long when cfg.bg3d == 0, else nx*ny*nz.  When rhoinv
 * is bitwise 1.0/rho0 at every point (initialize.f90:1012 makes it so), the
 * FP64 MHD stage forms it by division instead of reading it from HBM; the
 * environment variable SPARC_RHOINV=load keeps the read. */
typedef struct sparc_bg {
  const double *rho0, *rhoinv, *c2rho0, *gradrho0_z, *gradp0_z, *spongexyz, *p0;
  const double* g; /* nz, PML-scaled gravity */
  /* MHD only (nx*ny*nz): */
  const double *box, *boy, *boz, *curlbox, *curlboy, *curlboz;
  const double *gradrho0_x, *gradrho0_y, *gradp0_x, *gradp0_y, *reduction;
  /* CPML coefficients az/bzpml/kappaz & co. per face (face-shaped, x fastest):
   * x faces (npml, ny, nz), y faces (nx, npml, nz), z faces (nx, ny, npml).
   * NULL for inactive faces. */
  const double *pml_a[6], *pml_b[6], *pml_kappa[6];
  /* Separable pulse profile: forcing_p(i,j,k) = amp(t) * src_x[i] * src_y[j] *
   * src_z[k] (driver.f90:140-145); NULL when pulse_dir == 0. */
  const double *src_x, *src_y, *src_z;
  double dimc, dimrho; /* scales of the KE norm (driver.f90:161) */
} sparc_bg;

/* Context lifetime. */
int sparc_hip_create(const sparc_cfg* cfg, sparc_ctx** out);
int sparc_hip_destroy(sparc_ctx* ctx);

/* Upload the post-set-up backgrounds (replaces the module arrays STEP reads). */
int sparc_hip_set_background(sparc_ctx* ctx, const sparc_bg* bg);

/* FFT damping of the velocities (DAMPING = 1): rates is damping_rates(nx/2+1,
 * ny), Fortran order, as INIT_DAMPING sets it (initialize.f90:1276-1336; the
 * rates carry the 1/(nx*ny) of the unnormalised transform pair).  Every RK
 * stage then subtracts damp_velocity(v) -- per z plane the c2r transform of
 * the r2c transform of vx, vy, vz times the rates (all_modules.f90:1802-1888)
 * -- from their right-hand sides after the boundary zeroing (RHS.f90:166-172,
 * 659-665).  NULL switches damping off.  Runs with damping are not batched. */
int sparc_hip_set_damping(sparc_ctx* ctx, const double* rates);

/* PULSE = 0 driver (driver.f90:73-80, 130-133): map is this step's
 * forcing(nx, ny), Fortran order, nondimensional -- LAGRANGE_INTERP of the
 * forcing cube times source_damp (all_modules.f90:1892-1950) -- and
 * source_dep the depth profile (nz values, all_modules.f90:289-301; NULL keeps
 * the last one).  Every RK stage then adds forcing*source_dep to RHSv_z
 * (RHS.f90:146-147, 640-641) in place of the pulse.  Set before each step
 * (the call waits for the context's queued work); map NULL returns to the
 * pulse.  Runs driven this way are not batched. */
int sparc_hip_set_forcing(sparc_ctx* ctx, const double* map, const double* source_dep);

/* State in / out: a is nv*nx*ny*nz (field-major, the reference's `a`,
 * all_modules.f90:104); psi[f] is npsi*face_size for each active face (the
 * PML bases zpmlvars/xpmlvars/ypmlvars), NULL for inactive faces or to leave
 * them untouched (put) / skip them (get).  put_state clears the stage buffers.
 * In FP32 a state value is held as a pair (v, c) meaning v - c (the Kahan
 * compensation of the RK update, stored as bfloat16): get_state returns
 * v - c in FP64 and put_state splits x into v = fl32(x), c = bf16(fl32(v - x)),
 * so a state written from an FP32 run and read back continues as the
 * uninterrupted run (to ~2^-33 of the value). */
int sparc_hip_put_state(sparc_ctx* ctx, const double* a, const double* const psi[6]);
int sparc_hip_get_state(sparc_ctx* ctx, double* a, double* const psi[6]);

/* STEP (step.f90:20-63) with the source amplitude of this step,
 * amp = pulse_amp/0.56458272/dimc * sin(..) * exp(..) at t = time*timestep
 * (driver.f90:140-142).  Asynchronous on the context's stream. */
int sparc_hip_step(sparc_ctx* ctx, double src_amp);

/* FILTER_VARS (derivatives.f90:659-703) for time index `time`. */
int sparc_hip_filter(sparc_ctx* ctx, int time);

/* The driver loop body for nn = n0..n1 (driver.f90:123-200): STEP with
 * src_amp[nn-n0], then FILTER_VARS when nn != 0.  Asynchronous. */
int sparc_hip_advance(sparc_ctx* ctx, int n0, int n1, const double* src_amp);

/* Diagnostics of the most recent step: out[0] = norm_KE in cgs
 * (driver.f90:161), out[1] = min(rho+rho0), out[2] = min(p+p0) over the
 * stage-5 state (driver.f90:180).  Synchronises. */
int sparc_hip_diag(sparc_ctx* ctx, double out[3]);

/* The reference's failure checks of every step (driver.f90:175-196:
 * norm_KE > 1e3, min(rho+rho0) <= 0, min(p+p0) <= 0), evaluated on the
 * device by each step's diagnostic reduction, with no host synchronisation.
 * A NaN norm or minimum also fails.  Modes: SPARC_GUARD_HALT (the default)
 * records the first failure and stops the run there: every later STEP and
 * FILTER_VARS of the context leaves the state alone, so it stays that of the
 * failing step before its filter -- what the reference writes to
 * full_XXXX.h5 before it stops (all_modules.f90:1090-1100).
 * SPARC_GUARD_RECORD records the first failure and keeps stepping;
 * SPARC_GUARD_OFF checks nothing.  set_guard also clears a recorded failure,
 * and so does put_state. */
#define SPARC_GUARD_OFF 0
#define SPARC_GUARD_RECORD 1
#define SPARC_GUARD_HALT 2
int sparc_hip_set_guard(sparc_ctx* ctx, int mode);
/* Synchronises; tripped = 1 after a failure, at_step = its step count (the
 * n-th step the context completed, counting from 1), steps = steps the
 * context has completed, diag = norm_KE (cgs), min(rho+rho0), min(p+p0) of
 * the failing step.  Any output pointer may be NULL. */
int sparc_hip_guard(sparc_ctx* ctx, int* tripped, long* at_step, long* steps, double diag[3]);

/* Drop-in operators on one host FP64 array of nx*ny*nz (the context supplies
 * the metric).  bc: 1 (one-sided edges) or 5 (periodic) for x/y; 1 for z. */
int sparc_hip_ddx(sparc_ctx* ctx, const double* in, double* out, int bc);
int sparc_hip_ddy(sparc_ctx* ctx, const double* in, double* out, int bc);
int sparc_hip_ddz(sparc_ctx* ctx, const double* in, double* out, int bc);
int sparc_hip_filtx(sparc_ctx* ctx, const double* in, double* out);
int sparc_hip_filty(sparc_ctx* ctx, const double* in, double* out);
int sparc_hip_filtz(sparc_ctx* ctx, const double* in, double* out);

/* ddx / ddy / ddz (axis 0, 1, 2) without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  It runs the kernel of sparc_hip_ddx/ddy/ddz with
 * the host input and output staged through a static device array the
 * library holds on every device it runs on (64 MiB each, until the process
 * ends) instead of the context's operator buffer, which those entry points
 * allocate on their first call.  Limit: nx*ny*nz <= 4194304 points; larger
 * grids are refused (SPARC_E_INVALID).  All calls on one device share that
 * array: calls that overlap on one device race.  Synchronises. */
int sparc_hip_deriv_noalloc(sparc_ctx* ctx, int axis, const double* in, double* out, int bc);

/* Wait for all work queued on the context. */
int sparc_hip_sync(sparc_ctx* ctx);

/* Kernel timing with HIP events on the context's stream: enable != 0 starts
 * recording every stage launch; sparc_hip_kernel_time returns the number of
 * timed launches and their total milliseconds for kind 0 = RK stage 1,
 * 1 = RK stages 2-4, 2 = RK stage 5 (with the final update), 3 = filter passes,
 * 4 = MHD z passes (the z derivatives each MHD stage launch reads). */
int sparc_hip_timing(sparc_ctx* ctx, int enable);
int sparc_hip_kernel_time(sparc_ctx* ctx, int kind, long* count, double* total_ms);

/* Tile (x, y, z-chunk) the stage kernel uses and its launch size. */
int sparc_hip_tile(sparc_ctx* ctx, int tile[3], long* nblocks);

/* Tile autotuner (no counterpart in the reference; it serves the batched and
 * served forms, whose runs differ in grid size).  Times `reps` launches of a
 * middle RK stage for every supported tile shape that fits the grid, at the
 * automatic z chunk, half and twice it, and keeps the fastest: tile_out = (tx, ty,
 * kc), ms_out = its time per stage.  A middle stage writes only buffers the
 * next step overwrites, so the run's state is unchanged.  Call between steps
 * after set_background and put_state.  sparc_hip_set_tile applies a known
 * choice (kc <= 0: automatic); it fails for a shape that does not fit. */
int sparc_hip_autotune(sparc_ctx* ctx, int reps, int tile_out[3], double* ms_out);
int sparc_hip_set_tile(sparc_ctx* ctx, const int tile[3]);

/* Set-up helper: the Alfven-limiter decimal search of initialize.f90:962-986
 * (the search temperature T; the host then sets reduction = T/(T + red)).
 * red = |B0|^2/(rho0 cq^2), b2 = |B0|^2, rho0: nx*ny*nz nondimensional arrays;
 * maxva in cm/s.  Synchronises. */
int sparc_hip_alfven_search(sparc_ctx* ctx, const double* red, const double* b2,
                            const double* rho0, double dimb, double dimrho, double maxva,
                            double* temp_out);

/* Planes k0..k1 (0-based, inclusive) of state field `field` into out
 * (nx*ny*(k1-k0+1) FP64, x fastest): the full-width chunk output of
 * WRITE_OUT_CHUNK (all_modules.f90:1566-1799).  Synchronises. */
int sparc_hip_get_slab(sparc_ctx* ctx, int field, int k0, int k1, double* out);

/* Batched form: advance many independent runs (contexts on one device, one
 * precision) together.  Each RK stage is one persistent launch per kernel
 * variant and tile shape over the work items of all runs; with use_graph the
 * launch sequence of one step is captured once in a HIP graph and replayed.
 * No counterpart in the reference (one run per MPI job there); it serves
 * callers that run many forward simulations at once.
 * batch_advance: nsteps driver steps; time0[r] is run r's time index of the
 * first of them, amps[r*nsteps + n] its source amplitude at step n.  Filters
 * follow each run's own cadence.  With flags = SPARC_DEFER_FILTER the
 * filters of the last step are left out: the runs then hold the state the
 * reference writes to its output files (after STEP, before FILTER_VARS,
 * driver.f90:198-200), and sparc_hip_batch_filter(b, time) applies them,
 * time[r] being run r's time index of that step.  Each run's failure guard
 * works as for a single run.  Both synchronise before returning. */
#define SPARC_DEFER_FILTER 1
typedef struct sparc_batch sparc_batch;
int sparc_hip_batch_create(sparc_ctx* const* runs, int nrun, int use_graph, sparc_batch** out);
int sparc_hip_batch_advance(sparc_batch* b, int nsteps, const int* time0, const double* amps,
                            int flags);
int sparc_hip_batch_filter(sparc_batch* b, const int* time);
int sparc_hip_batch_info(sparc_batch* b, int* ngroups, long* nlaunch_blocks);
int sparc_hip_batch_destroy(sparc_batch* b);
/* Time per middle RK stage of the whole batch (reps launches of every
 * group), for tuning the tile shape of a group of runs; leaves every run's
 * state unchanged.  No counterpart in the reference. */
int sparc_hip_batch_time_stage(sparc_batch* b, int reps, double* ms_out);

/* Bandwidth probe (no counterpart in the reference): reps launches of a
 * kernel reading nread arrays (1..48) of n doubles and writing their sum to
 * one more (16-byte non-temporal accesses), on device `device`; gbps = bytes
 * moved per second / 1e9.  It allocates its own buffers.  The stage
 * kernels' roofline fractions are read against it at their read-to-write
 * ratios as well as against the data-sheet HBM peak. */
int sparc_hip_stream_bench(int device, long n, int nread, int reps, double* gbps);

/* Free and total device memory in bytes (hipMemGetInfo): the served form
 * sizes its HBM budget with it.  No counterpart in the reference. */
int sparc_hip_device_memory(int device, long* free_bytes, long* total_bytes);

const char* sparc_hip_strerror(int code);
const char* sparc_hip_last_error(void);
int sparc_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SPARC_HIP_H */
