/*
 * sparc_oracle.h -- CPU FP64 restatement of the SPARC2 time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: it is linked by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product library (sparc2_amd/csrc) never includes,
 * links or calls it.
 *
 * Layout: every 3-D array is Fortran order (x fastest): idx = i + nx*(j + ny*k),
 * 0-based.  Multi-field arrays are field-major: field f starts at f*nx*ny*nz.
 * Field order (all_modules.f90:104, initialize.f90:780-826):
 *   0 rho, 1 vx, 2 vy, 3 vz, 4 p, 5 bx, 6 by, 7 bz.
 * PML faces: 0 xbot, 1 xtop, 2 ybot, 3 ytop, 4 zbot, 5 ztop.
 *   x faces (npml, ny, nz), y faces (nx, npml, nz), z faces (nx, ny, npml).
 */
#ifndef SPARC_ORACLE_H
#define SPARC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cfg {
  int nx, ny, nz;
  int magnetic;          /* control-file MAGNETIC: 0 quiet, 1 thermal, 2 MHD */
  int periodic, use_pml, use_hpml, stabilise;
  int npml[6];           /* nxpmlbot, nxpmltop, nypmlbot, nypmltop, nzpmlbot, nzpmltop */
  double pmlc[6][4];     /* per face: RC, n, kappa, sponge */
  double xlength, ylength, timestep;
  int g_xy_filt, g_z_filt;           /* 5 (11-point) or 3 (7-point) */
  double s_xy, s_z, t_xy, t_z;
  int pulse_dir;                     /* 1 vx, 2 vy, 3 vz, 0 no source */
  double pulse_t0, pulse_t1, pulse_st, pulse_px, pulse_py, pulse_pz;
  double pulse_sr, pulse_sz, pulse_amp;
  double sp_str, sp_length, sp_decay;
  double maxva;
} orc_cfg;

typedef struct orc_ctx orc_ctx;

/* READ_IN_HDF5 scaling + INITANDALLOCATE + INITIALISE_RHS + INITIALIZE_STEP.
 * Background inputs are cgs, Fortran order; zaxis in Mm; grav in cgs.
 * bx0/by0/bz0 may be NULL unless magnetic == 2.  Returns NULL on bad config. */
orc_ctx* orc_create(const orc_cfg* cfg, const double* rho0, const double* p0,
                    const double* cs0, const double* bx0, const double* by0,
                    const double* bz0, const double* zaxis_mm, const double* grav);
void orc_destroy(orc_ctx* c);
const char* orc_last_error(void);

int orc_nv(const orc_ctx* c);
int orc_npsi(const orc_ctx* c);
long orc_face_size(const orc_ctx* c, int face);

/* Named scalars / arrays of the set-up state (see name tables in the .c file). */
double orc_scalar(const orc_ctx* c, const char* name);
long orc_array_len(const orc_ctx* c, const char* name);
int orc_get_array(const orc_ctx* c, const char* name, double* out);

/* State access: a is nv*N; psi face arrays are npsi*face_size. */
void orc_put_state(orc_ctx* c, const double* a);
void orc_get_state(const orc_ctx* c, double* a);
void orc_get_temp_step(const orc_ctx* c, double* u);
void orc_get_scr(const orc_ctx* c, double* k);
int orc_put_psi(orc_ctx* c, int face, const double* psi);
int orc_get_psi(const orc_ctx* c, int face, double* psi);

/* One driver iteration pieces (driver.f90:123-200). */
void orc_forcing(orc_ctx* c, int time);      /* forcing_p at t = time*timestep */
void orc_step(orc_ctx* c);                    /* STEP */
void orc_filter_vars(orc_ctx* c, int time);   /* FILTER_VARS */
void orc_diag(const orc_ctx* c, double out[3]); /* norm_KE (cgs), min(rho+rho0), min(p+p0) */
/* Driver loop nn = n0..n1 (source, step, diag, filter if nn != 0).
 * hist (may be NULL) receives 3 doubles per iteration. */
void orc_run(orc_ctx* c, int n0, int n1, double* hist);

/* DAMPING: on != 0 computes damping_rates (INIT_DAMPING, initialize.f90:
 * 1276-1336) and makes every RHS subtract the damping term of vx, vy, vz
 * (RHS.f90:166-172, 659-665).  orc_damp_term: damp_velocity of one 3-D
 * field (all_modules.f90:1802-1888) with the current rates. */
int orc_set_damping(orc_ctx* c, int on);
/* PULSE = 0: this step's forcing(nx, ny) (LAGRANGE_INTERP x source_damp) and
 * source_dep(nz) (NULL keeps it); the RHS adds forcing*source_dep to RHSv_z
 * (RHS.f90:146-147, 640-641).  map NULL returns to the pulse. */
int orc_set_forcing_map(orc_ctx* c, const double* map, const double* sdep);
void orc_damp_term(const orc_ctx* c, const double* v, double* term);

/* RHS of the current temp_step / psi stage values into scr (one call of
 * MP_QUIET_3D or MP_MHD_3D with the stage psi = base + beta*K). */
void orc_rhs_stage(orc_ctx* c, int stage);

/* Standalone operators on one 3-D array (derivatives.f90). bc: 1 or 5. */
int orc_ddx(const orc_ctx* c, const double* in, double* out, int bc);
int orc_ddy(const orc_ctx* c, const double* in, double* out, int bc);
int orc_ddz(const orc_ctx* c, const double* in, double* out, int bc);
void orc_filtx(const orc_ctx* c, const double* in, double* out);
void orc_filty(const orc_ctx* c, const double* in, double* out);
void orc_filtz(const orc_ctx* c, const double* in, double* out);

/* Coefficient tables: dp is 6x11 (rows 0..4 = one-sided rows 1..5, row 5 =
 * centred), filter is 2g+1 long for g = 5 or 3. */
void orc_coeffs(double dp[66], double f11[11], double f7[7]);

void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
