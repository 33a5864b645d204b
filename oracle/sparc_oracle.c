/*
 * sparc_oracle.c -- CPU FP64 restatement of the SPARC2 time-step path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle).  Used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the product
 * library never links or calls it.
 *
 * Every routine below restates the reference Fortran (cited file:line) with
 * the domain-decomposed edge logic rewritten in global terms.  The reference
 * decides edge behaviour per MPI block (IAM_XBOT/IAM_XTOP/...); with a single
 * block in an axis both flags are true and the reference takes only the
 * "bottom" branch (RHS.f90:51-61, derivatives.f90:62,504).  The oracle
 * emulates at least two blocks per axis, i.e. bottom and top edges each get
 * their own branch, which is what every multi-block reference run computes.
 *
 * Arithmetic follows the reference's evaluation order where it is visible in
 * the source (left-to-right products, SUM in index order); the file is built
 * with -ffp-contract=off so no FMA contraction is introduced.
 *
 * Parity status of each function is recorded in DESIGN.md; the pinning tests
 * live in tests/test_oracle_*.py.
 */
#include "sparc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* all_modules.f90:56 -- the truncated constants used throughout. */
#define RSUN 69598946770.0
#define PI 3.14159265358979

static char g_err[256];
const char* orc_last_error(void) { return g_err; }
static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* Coefficient tables: initialize.f90:1927-2034 (initexplicitders).          */
/* dp[r][m]: r = 0..4 are the one-sided rows dpz(1..5,:), r = 5 the centred  */
/* row dpz(6,:).  The literal corrections (+1e-12 etc.) are kept as written. */
/* ------------------------------------------------------------------------ */
static double DP[6][11];
static double F11[11], F7[7];
static int coeffs_ready = 0;

static void init_coeffs(void) {
  if (coeffs_ready) return;
  /* filter 1 (7-point), initialize.f90:1932-1938 */
  const double c7[7] = {-0.015625000, 0.093750000, -0.234375000, 0.312500000,
                        -0.234375000, 0.093750000, -0.015625000};
  for (int m = 0; m < 7; ++m) F7[m] = c7[m];
  /* filter 2 (11-point), initialize.f90:1941-1951 (0.5 is an exact literal) */
  F11[0] = 0.5 * 0.00195312500000;
  F11[1] = -0.5 * 0.01953125000000;
  F11[2] = 0.5 * 0.08789062500000;
  F11[3] = -0.5 * 0.23437500000000;
  F11[4] = 0.5 * 0.41015625000000;
  F11[5] = 0.75390625000000;
  F11[6] = 0.5 * 0.41015625000000;
  F11[7] = -0.5 * 0.23437500000000;
  F11[8] = 0.5 * 0.08789062500000;
  F11[9] = -0.5 * 0.01953125000000;
  F11[10] = 0.5 * 0.00195312500000;
  /* centred 11-point DRP row, initialize.f90:1959-1969 */
  const double c[11] = {-0.002484594688, +0.020779405824, -0.090320001280,
                        +0.286511173973, -0.872756993962, 0.0,
                        +0.872756993962, -0.286511173973, +0.090320001280,
                        -0.020779405824, +0.002484594688};
  for (int m = 0; m < 11; ++m) DP[5][m] = c[m];
  /* one-sided rows, initialize.f90:1971-2029 */
  const double r1[11] = {-2.391602219538, 5.832490322294,
                         -7.650218001182 + 0.000000000001, 7.907810563576,
                         -5.922599052629, 3.071037015445, -1.014956769726,
                         0.170022256519, 0.002819958377, -0.004791009708,
                         -0.000013063429};
  const double r2[11] = {-0.180022054228, -1.237550583044, 2.484731692990,
                         -1.810320814061, 1.112990048440, -0.481086916514,
                         0.126598690230, -0.015510730165,
                         0.000024609059 - 0.000003,
                         0.000156447571 - 0.000000000001, -0.000007390277};
  const double r3[11] = {0.057982271137, -0.536135360383,
                         -0.264089548967 + 0.000000000002,
                         0.917445877606 - 0.000000000002, -0.169688364841,
                         -0.029716326170, 0.029681617641, -0.005222483773,
                         -0.000118806260, -0.000118806260, -0.000020069730};
  const double r4[11] = {-0.013277273810, 0.115976072920, -0.617479187931,
                         -0.274113948206 + 0.000000000002,
                         1.086208764655 - 0.000000000002, -0.402951626982,
                         0.131066986242, -0.028154858354, 0.002596328316,
                         0.000128743150, 0.0};
  const double r5[11] = {0.016756572303, -0.117478455239, 0.411034935097,
                         -1.130286765151, 0.341435872100 - 0.000000000001,
                         0.556396830543, -0.082525734207, 0.003565834658,
                         0.001173034777, -0.000071772671 + 0.000000000064,
                         -0.000000352273};
  for (int m = 0; m < 11; ++m) {
    DP[0][m] = r1[m];
    DP[1][m] = r2[m];
    DP[2][m] = r3[m];
    DP[3][m] = r4[m];
    DP[4][m] = r5[m];
  }
  coeffs_ready = 1;
}

void orc_coeffs(double dp[66], double f11[11], double f7[7]) {
  init_coeffs();
  for (int r = 0; r < 6; ++r)
    for (int m = 0; m < 11; ++m) dp[r * 11 + m] = DP[r][m];
  for (int m = 0; m < 11; ++m) f11[m] = F11[m];
  for (int m = 0; m < 7; ++m) f7[m] = F7[m];
}

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */
enum { NWORK = 24 };

struct orc_ctx {
  orc_cfg cfg;
  int nx, ny, nz;
  long N;
  int nv, npsi, mhd, quiet;
  int xyfreq, zfreq;
  double dimc, dimrho, diml, dimb;
  double stretchx, stretchy, deltat, betas[5];
  double *xaxis, *yaxis, *zaxis, *g, *cq, *stretch, *unstretch, *spongez;
  double *rho0, *p0, *cs, *c2, *rhoinv, *c2rho0, *gradrho0_z, *gradp0_z;
  double *spongexyz;
  double *box, *boy, *boz, *curlbox, *curlboy, *curlboz;
  double *gradrho0_x, *gradrho0_y, *gradp0_x, *gradp0_y, *reduction;
  int fdim[6][3], foff[6][3];
  long fsize[6];
  int factive[6];
  double *pa[6], *pb[6], *pk[6];
  double *pvars[6], *ppsi[6], *pscr[6];
  double *a, *u, *scr, *forcing;
  double *w[NWORK];
  double* damp_rates; /* (ny, nx/2+1) when DAMPING is on, else NULL */
  /* PULSE = 0: this step's forcing (nx*ny, LAGRANGE_INTERP x source_damp)
   * and source_dep (nz); fmap NULL: the pulse */
  double *fmap, *sdep;
};

#define IX(c, i, j, k) ((long)(i) + (long)(c)->nx * ((long)(j) + (long)(c)->ny * (long)(k)))

static double* dalloc(long n) {
  double* p = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  return p;
}

int orc_nv(const orc_ctx* c) { return c->nv; }
int orc_npsi(const orc_ctx* c) { return c->npsi; }
long orc_face_size(const orc_ctx* c, int f) {
  return (f >= 0 && f < 6) ? c->fsize[f] : -1;
}

/* ------------------------------------------------------------------------ */
/* Derivatives: derivatives.f90 ddx 10-116, ddy 226-341, ddz 451-562.        */
/* ------------------------------------------------------------------------ */

/* bc = 5: periodic wrap (the Cartesian communicator is periodic in x, y;
 *         initialize.f90:32-36), centred everywhere.
 * bc = 1: bottom rows one-sided dpxy(i,:) over f(1:11) (derivatives.f90:66-75),
 *         top rows the mirrored one-sided form with a sign flip
 *         (derivatives.f90:92-101). */
static void deriv_line(const double* f, long stride, int n, double s, int bc,
                       int i, double* out) {
  double acc = 0.0;
  if (bc == 5) {
    for (int m = 0; m < 11; ++m) {
      int ii = i + m - 5;
      ii = (ii % n + n) % n;
      acc += f[(long)ii * stride] * DP[5][m];
    }
    *out = s * acc;
  } else {
    if (i < 5) {
      for (int m = 0; m < 11; ++m) acc += f[(long)m * stride] * DP[i][m];
      *out = s * acc;
    } else if (i >= n - 5) {
      int r = n - 1 - i;
      for (int m = 0; m < 11; ++m) acc += f[(long)(n - 1 - m) * stride] * DP[r][m];
      *out = -s * acc;
    } else {
      for (int m = 0; m < 11; ++m) acc += f[(long)(i + m - 5) * stride] * DP[5][m];
      *out = s * acc;
    }
  }
}

int orc_ddx(const orc_ctx* c, const double* in, double* out, int bc) {
  if (bc != 1 && bc != 5) return -1;
  init_coeffs();
#pragma omp parallel for
  for (int k = 0; k < c->nz; ++k)
    for (int j = 0; j < c->ny; ++j) {
      const double* row = in + IX(c, 0, j, k);
      for (int i = 0; i < c->nx; ++i)
        deriv_line(row, 1, c->nx, c->stretchx, bc, i, &out[IX(c, i, j, k)]);
    }
  return 0;
}

int orc_ddy(const orc_ctx* c, const double* in, double* out, int bc) {
  if (bc != 1 && bc != 5) return -1;
  init_coeffs();
#pragma omp parallel for
  for (int k = 0; k < c->nz; ++k)
    for (int j = 0; j < c->ny; ++j)
      for (int i = 0; i < c->nx; ++i)
        deriv_line(in + IX(c, i, 0, k), c->nx, c->ny, c->stretchy, bc, j,
                   &out[IX(c, i, j, k)]);
  return 0;
}

/* z is never periodic.  bc = 1 (the only value the RHS and the set-up use):
 * the bottom block applies the centred row over zero ghost cells
 * (derivatives.f90:465,509-518: temp = 0 and no neighbour below); the top
 * block applies the mirrored one-sided rows (derivatives.f90:536-545). */
int orc_ddz(const orc_ctx* c, const double* in, double* out, int bc) {
  if (bc != 1) return -1;
  init_coeffs();
  const long sz = (long)c->nx * c->ny;
  const int nz = c->nz;
#pragma omp parallel for
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < c->ny; ++j)
      for (int i = 0; i < c->nx; ++i) {
        const double* col = in + IX(c, i, j, 0);
        double acc = 0.0;
        if (k >= nz - 5) {
          int r = nz - 1 - k;
          for (int m = 0; m < 11; ++m) acc += col[(long)(nz - 1 - m) * sz] * DP[r][m];
          out[IX(c, i, j, k)] = -c->stretch[k] * acc;
        } else {
          for (int m = 0; m < 11; ++m) {
            int kk = k + m - 5;
            double v = (kk >= 0) ? col[(long)kk * sz] : 0.0;
            acc += v * DP[5][m];
          }
          out[IX(c, i, j, k)] = c->stretch[k] * acc;
        }
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Filters: derivatives.f90 filtx 119-223, filty 345-447, filtz 565-642.     */
/* ------------------------------------------------------------------------ */
static const double* filt_coeffs(int g) { return g == 5 ? F11 : F7; }

/* x/y: periodic wrap; otherwise the g edge points are copied (g = 5) or zeroed
 * (g = 3) at both ends (derivatives.f90:190-192,215-217,414-416,439-441). */
static void filt_line(const double* f, long stride, int n, int g, int periodic,
                      int i, double* out) {
  const double* w = filt_coeffs(g);
  double acc = 0.0;
  if (periodic) {
    for (int m = 0; m <= 2 * g; ++m) {
      int ii = i + m - g;
      ii = (ii % n + n) % n;
      acc += f[(long)ii * stride] * w[m];
    }
    *out = acc;
  } else if (i < g || i >= n - g) {
    *out = (g == 5) ? f[(long)i * stride] : 0.0;
  } else {
    for (int m = 0; m <= 2 * g; ++m) acc += f[(long)(i + m - g) * stride] * w[m];
    *out = acc;
  }
}

void orc_filtx(const orc_ctx* c, const double* in, double* out) {
  init_coeffs();
#pragma omp parallel for
  for (int k = 0; k < c->nz; ++k)
    for (int j = 0; j < c->ny; ++j)
      for (int i = 0; i < c->nx; ++i)
        filt_line(in + IX(c, 0, j, k), 1, c->nx, c->cfg.g_xy_filt, c->cfg.periodic, i,
                  &out[IX(c, i, j, k)]);
}

void orc_filty(const orc_ctx* c, const double* in, double* out) {
  init_coeffs();
#pragma omp parallel for
  for (int k = 0; k < c->nz; ++k)
    for (int j = 0; j < c->ny; ++j)
      for (int i = 0; i < c->nx; ++i)
        filt_line(in + IX(c, i, 0, k), c->nx, c->ny, c->cfg.g_xy_filt, c->cfg.periodic,
                  j, &out[IX(c, i, j, k)]);
}

/* z: the g bottom and top layers are copied for both filter widths
 * (derivatives.f90:624,635); interior centred. */
void orc_filtz(const orc_ctx* c, const double* in, double* out) {
  init_coeffs();
  const int g = c->cfg.g_z_filt, nz = c->nz;
  const double* w = filt_coeffs(g);
  const long sz = (long)c->nx * c->ny;
#pragma omp parallel for
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < c->ny; ++j)
      for (int i = 0; i < c->nx; ++i) {
        const double* col = in + IX(c, i, j, 0);
        if (k < g || k >= nz - g) {
          out[IX(c, i, j, k)] = col[(long)k * sz];
        } else {
          double acc = 0.0;
          for (int m = 0; m <= 2 * g; ++m) acc += col[(long)(k + m - g) * sz] * w[m];
          out[IX(c, i, j, k)] = acc;
        }
      }
}

/* ------------------------------------------------------------------------ */
/* Set-up                                                                    */
/* ------------------------------------------------------------------------ */

static void face_geometry(orc_ctx* c) {
  const int* np = c->cfg.npml;
  for (int f = 0; f < 6; ++f) {
    int d[3] = {c->nx, c->ny, c->nz}, o[3] = {0, 0, 0};
    int ax = f / 2, top = f % 2;
    d[ax] = np[f];
    o[ax] = top ? ((ax == 0 ? c->nx : ax == 1 ? c->ny : c->nz) - np[f]) : 0;
    for (int q = 0; q < 3; ++q) {
      c->fdim[f][q] = d[q];
      c->foff[f][q] = o[q];
    }
    c->fsize[f] = (long)d[0] * d[1] * d[2];
    c->factive[f] = (ax < 2 ? c->cfg.use_hpml : c->cfg.use_pml) && np[f] > 0;
  }
}

static double maxval_nan_skip(const double* v, long n) {
  double m = -INFINITY;
  for (long i = 0; i < n; ++i)
    if (v[i] > m) m = v[i];
  return m;
}

/* Decimal search for the Alfven limiter, initialize.f90:962-986. */
static void reduction_search(orc_ctx* c) {
  const long N = c->N;
  double* red = c->reduction;
  double* tmp = dalloc(N);
  for (int k = 0; k < c->nz; ++k)
    for (int j = 0; j < c->ny; ++j)
      for (int i = 0; i < c->nx; ++i) {
        long p = IX(c, i, j, k);
        red[p] = (c->boz[p] * c->boz[p] + c->box[p] * c->box[p] + c->boy[p] * c->boy[p]) /
                 (c->rho0[p] * (c->cq[k] * c->cq[k]));
      }
  double va = 0.0, dtemp = 1.0e9, temp = -dtemp;
  if (maxval_nan_skip(red, N) > 0) {
    const double maxva = c->cfg.maxva;
    while ((dtemp > 1.0e-5) && (temp < 1.0e10 - 1.1)) {
      while ((va < maxva) && (temp < 1.0e10)) {
        temp = temp + dtemp;
        const double t = temp;
#pragma omp parallel for
        for (long p = 0; p < N; ++p) {
          double b2 = c->box[p] * c->box[p] + c->boy[p] * c->boy[p] + c->boz[p] * c->boz[p];
          tmp[p] = sqrt(c->dimb * c->dimb * t / (t + red[p]) * b2 /
                        (4.0 * PI * c->dimrho * c->rho0[p]));
        }
        va = maxval_nan_skip(tmp, N);
      }
      temp = temp - dtemp;
      dtemp = dtemp / 10.0;
      va = 0.0;
    }
  } else {
    temp = 1.0e10;
  }
  for (long p = 0; p < N; ++p) red[p] = temp / (temp + red[p]);
  free(tmp);
}

orc_ctx* orc_create(const orc_cfg* cfg, const double* rho0_in, const double* p0_in,
                    const double* cs_in, const double* bx_in, const double* by_in,
                    const double* bz_in, const double* zaxis_mm, const double* grav) {
  init_coeffs();
  if (cfg->nx < 16 || cfg->ny < 16 || cfg->nz < 16) {
    set_err("grid too small (need >= 16 cells per axis)");
    return NULL;
  }
  if ((cfg->g_xy_filt != 5 && cfg->g_xy_filt != 3) ||
      (cfg->g_z_filt != 5 && cfg->g_z_filt != 3)) {
    set_err("filter half-width must be 5 or 3");
    return NULL;
  }
  if (cfg->magnetic == 2 && (!bx_in || !by_in || !bz_in)) {
    set_err("magnetic run needs bx0, by0, bz0");
    return NULL;
  }
  if (cfg->use_hpml && cfg->periodic) {
    set_err("USE_HPML forces PERIODIC off (initialize.f90:1610-1620)");
    return NULL;
  }
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  c->cfg = *cfg;
  if (!cfg->use_hpml)
    c->cfg.npml[0] = c->cfg.npml[1] = c->cfg.npml[2] = c->cfg.npml[3] = 0;
  if (!cfg->use_pml) c->cfg.npml[4] = c->cfg.npml[5] = 0;
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  const long N = (long)nx * ny * nz;
  c->N = N;
  c->mhd = cfg->magnetic == 2;
  c->quiet = cfg->magnetic == 0;
  c->nv = c->mhd ? 8 : 5;
  c->npsi = c->mhd ? 6 : 2;

  /* t_xy/t_z clamp and cadences: initialize.f90:1762-1768,1911-1912 */
  double txy = cfg->t_xy < cfg->timestep ? cfg->timestep : cfg->t_xy;
  double tz = cfg->t_z < cfg->timestep ? cfg->timestep : cfg->t_z;
  c->xyfreq = (int)floor(txy / cfg->timestep + 1.0e-5);
  c->zfreq = (int)floor(tz / cfg->timestep + 1.0e-5);

  c->xaxis = dalloc(nx);
  c->yaxis = dalloc(ny);
  c->zaxis = dalloc(nz);
  c->g = dalloc(nz);
  c->cq = dalloc(nz);
  c->stretch = dalloc(nz);
  c->unstretch = dalloc(nz);
  c->spongez = dalloc(nz);
  c->rho0 = dalloc(N);
  c->p0 = dalloc(N);
  c->cs = dalloc(N);
  c->c2 = dalloc(N);
  c->rhoinv = dalloc(N);
  c->c2rho0 = dalloc(N);
  c->gradrho0_z = dalloc(N);
  c->gradp0_z = dalloc(N);
  c->spongexyz = dalloc(N);
  if (c->mhd) {
    c->box = dalloc(N);
    c->boy = dalloc(N);
    c->boz = dalloc(N);
    c->curlbox = dalloc(N);
    c->curlboy = dalloc(N);
    c->curlboz = dalloc(N);
    c->gradrho0_x = dalloc(N);
    c->gradrho0_y = dalloc(N);
    c->gradp0_x = dalloc(N);
    c->gradp0_y = dalloc(N);
    c->reduction = dalloc(N);
  }
  c->a = dalloc(c->nv * N);
  c->u = dalloc(c->nv * N);
  c->scr = dalloc(c->nv * N);
  c->forcing = dalloc(N);
  for (int q = 0; q < NWORK; ++q) c->w[q] = dalloc(N);
  face_geometry(c);
  for (int f = 0; f < 6; ++f) {
    long fs = c->fsize[f];
    c->pa[f] = dalloc(fs);
    c->pb[f] = dalloc(fs);
    c->pk[f] = dalloc(fs);
    c->pvars[f] = dalloc(c->npsi * fs);
    c->ppsi[f] = dalloc(c->npsi * fs);
    c->pscr[f] = dalloc(c->npsi * fs);
  }

  /* Axes, initialize.f90:470-492 */
  for (int i = 0; i < nx; ++i)
    c->xaxis[i] = cfg->periodic ? (double)i / (double)nx : (double)i / (double)(nx - 1);
  for (int j = 0; j < ny; ++j)
    c->yaxis[j] = cfg->periodic ? (double)j / (double)ny : (double)j / (double)(ny - 1);
  /* stretchx/y, initialize.f90:503-517 */
  c->stretchx = cfg->periodic ? RSUN / cfg->xlength * nx : RSUN / cfg->xlength * (nx - 1);
  c->stretchy = cfg->periodic ? RSUN / cfg->ylength * ny : RSUN / cfg->ylength * (ny - 1);

  /* READ_IN_HDF5 after the read, all_modules.f90:284-412 */
  memcpy(c->rho0, rho0_in, sizeof(double) * N);
  memcpy(c->p0, p0_in, sizeof(double) * N);
  memcpy(c->cs, cs_in, sizeof(double) * N);
  if (c->mhd) {
    memcpy(c->box, bx_in, sizeof(double) * N);
    memcpy(c->boy, by_in, sizeof(double) * N);
    memcpy(c->boz, bz_in, sizeof(double) * N);
  }
  for (int k = 0; k < nz; ++k) {
    c->g[k] = grav[k];
    c->zaxis[k] = 1.0 + zaxis_mm[k] * 1.0e8 / RSUN;
  }
  c->dimc = c->cs[0];
  c->dimrho = c->rho0[0];
  c->diml = RSUN;
  c->dimb = pow(4.0 * PI * (c->dimc * c->dimc) * c->dimrho, 0.5);
  for (int k = 0; k < nz; ++k) c->cq[k] = c->cs[IX(c, 0, 0, k)];
  if (c->quiet) { /* horizontal copy of column (1,1,:), all_modules.f90:353-384 */
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          long p = IX(c, i, j, k), p1 = IX(c, 0, 0, k);
          c->p0[p] = c->p0[p1];
          c->rho0[p] = c->rho0[p1];
          c->cs[p] = c->cs[p1];
        }
  }
  const double dimc = c->dimc, dimrho = c->dimrho, diml = c->diml;
  for (int k = 0; k < nz; ++k) {
    c->g[k] = c->g[k] * diml / (dimc * dimc);
    c->cq[k] = c->cq[k] / dimc;
  }
  for (long p = 0; p < N; ++p) {
    c->p0[p] = c->p0[p] / (dimrho * (dimc * dimc));
    c->rho0[p] = c->rho0[p] / dimrho;
    c->cs[p] = c->cs[p] / dimc;
    if (c->mhd) {
      c->box[p] /= c->dimb;
      c->boy[p] /= c->dimb;
      c->boz[p] /= c->dimb;
    }
    c->c2[p] = c->cs[p] * c->cs[p];
  }

  /* unstretch / stretch, initialize.f90:674-730, with the bottom and top
   * blocks distinct. */
  for (int k = 0; k < nz; ++k) {
    double us;
    if (k < 5) {
      double acc = 0.0;
      for (int m = 0; m < 11; ++m) acc += c->zaxis[m] * DP[k][m];
      us = acc * (nz - 1);
    } else if (k >= nz - 5) {
      int r = nz - 1 - k;
      double acc = 0.0;
      for (int m = 0; m < 11; ++m) acc += c->zaxis[nz - 1 - m] * DP[r][m];
      us = -acc * (nz - 1);
    } else {
      const double* z = c->zaxis;
      us = ((-z[k - 5] + z[k + 5]) * DP[5][10] + (-z[k - 4] + z[k + 4]) * DP[5][9] +
            (-z[k - 3] + z[k + 3]) * DP[5][8] + (-z[k - 2] + z[k + 2]) * DP[5][7] +
            (-z[k - 1] + z[k + 1]) * DP[5][6]) *
           (nz - 1);
    }
    c->unstretch[k] = us;
    c->stretch[k] = (nz - 1) * 1.0 / us;
  }
  c->deltat = cfg->timestep * dimc / diml; /* initialize.f90:733 */

  /* ---- INITIALISE_RHS, initialize.f90:883-1228 ---- */
  const int bc = cfg->periodic ? 5 : 1;
  const double* z = c->zaxis;
  /* spongez: only defined without PML (initialize.f90:905-910); with PML the
   * reference leaves it uninitialised -- the intended value 0 is used. */
  for (int k = 0; k < nz; ++k)
    c->spongez[k] = cfg->use_pml ? 0.0
                                 : 400.0 / (1.0 + exp((1.002 - z[k]) / 0.0001)) +
                                       40.0 / (1.0 + exp((z[k] - (z[0] + 0.005)) / 0.0005));
  {
    const double start_sponge = cfg->sp_length / cfg->xlength;
    const double finish_sponge = 1.0 - cfg->sp_length / cfg->xlength;
    double* spx = dalloc(nx);
    double* spy = dalloc(ny);
    if (!cfg->periodic && !cfg->use_hpml) {
      for (int i = 0; i < nx; ++i) {
        double x = c->xaxis[i];
        spx[i] = cfg->sp_str * (1.0 - 1.0 / ((1.0 + exp((start_sponge - x) / cfg->sp_decay)) *
                                             (1.0 + exp((x - finish_sponge) / cfg->sp_decay))));
      }
      for (int j = 0; j < ny; ++j) {
        double y = c->yaxis[j];
        spy[j] = cfg->sp_str * (1.0 - 1.0 / ((1.0 + exp((start_sponge - y) / cfg->sp_decay)) *
                                             (1.0 + exp((y - finish_sponge) / cfg->sp_decay))));
      }
    }
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          c->spongexyz[IX(c, i, j, k)] =
              cfg->periodic ? c->spongez[k] : spx[i] + spy[j] + c->spongez[k];
    free(spx);
    free(spy);
  }

  if (c->mhd) {
    reduction_search(c);
    double* t = c->w[0];
    /* curlB0, initialize.f90:993-1006 */
    orc_ddx(c, c->boy, c->curlboz, bc);
    orc_ddy(c, c->box, t, bc);
    for (long p = 0; p < N; ++p) c->curlboz[p] = c->curlboz[p] - t[p];
    orc_ddy(c, c->boz, t, bc);
    orc_ddz(c, c->boy, c->curlbox, 1);
    for (long p = 0; p < N; ++p) c->curlbox[p] = t[p] - c->curlbox[p];
    orc_ddz(c, c->box, c->curlboy, 1);
    orc_ddx(c, c->boz, t, bc);
    for (long p = 0; p < N; ++p) c->curlboy[p] = c->curlboy[p] - t[p];
  }

  for (long p = 0; p < N; ++p) {
    c->rhoinv[p] = 1.0 / c->rho0[p];
    c->c2rho0[p] = c->c2[p] * c->rho0[p];
  }
  orc_ddz(c, c->rho0, c->gradrho0_z, 1);
  orc_ddz(c, c->p0, c->gradp0_z, 1);

  if (cfg->stabilise) { /* initialize.f90:1018-1034; the band limits are
                           default-real literals */
    const double lo = (double)0.9998f, hi = (double)1.0001f;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          long p = IX(c, i, j, k);
          double gr = c->gradrho0_z[p];
          if (z[k] > lo && z[k] < 1.0) {
            double v1 = c->c2[p] * gr + 1.e-4 * RSUN / (dimrho * (dimc * dimc));
            double v2 = -c->rho0[p] * c->g[k];
            c->gradp0_z[p] = v1 > v2 ? v1 : v2;
          } else if (z[k] <= lo || z[k] >= hi) {
            double v1 = c->c2[p] * gr;
            double v2 = -c->rho0[p] * c->g[k];
            c->gradp0_z[p] = v1 > v2 ? v1 : v2;
          }
        }
  }

  if (cfg->use_pml) {
    double* tin = c->w[1];
    for (long p = 0; p < N; ++p) {
      if (c->mhd) {
        double b2 = c->box[p] * c->box[p] + c->boy[p] * c->boy[p] + c->boz[p] * c->boz[p];
        tin[p] = sqrt(c->c2[p] + c->reduction[p] * b2 / c->rho0[p]);
      } else {
        tin[p] = sqrt(c->c2[p]);
      }
    }
    const double pml_f0 = PI * 0.005 * diml / dimc;
    /* ZBOT, initialize.f90:1048-1074 */
    if (c->factive[4]) {
      const int nzb = c->cfg.npml[4];
      const double* pc = c->cfg.pmlc[4];
      const double L = z[nzb - 1] - z[0];
      for (int kk = 0; kk < nzb; ++kk) {
        const double s = (z[nzb - 1] - z[kk]) / L;
        const double alpha = pml_f0 * (z[kk] - z[0]) / L;
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            long fp = (long)i + (long)nx * ((long)j + (long)ny * kk);
            long gp = IX(c, i, j, kk);
            double damp = pc[0] * tin[IX(c, i, j, nzb - 1)] * (pc[1] + 1.0) / (2.0 * L) * pow(s, pc[1]);
            double kap = 1.0 + (pc[2] - 1.0) * pow(s, pc[1]);
            c->pk[4][fp] = kap;
            c->pb[4][fp] = -alpha - damp / kap;
            c->pa[4][fp] = -damp / (kap * kap);
            c->gradrho0_z[gp] = c->gradrho0_z[gp] * alpha / kap / (damp / kap + alpha);
            c->gradp0_z[gp] = c->gradp0_z[gp] * alpha / kap / (damp / kap + alpha);
            c->spongexyz[gp] = c->spongexyz[gp] + pc[3] * tin[IX(c, i, j, nzb - 1)] * pow(s, pc[1] + 1.0);
          }
        double d11 = pc[0] * tin[IX(c, 0, 0, nzb - 1)] * (pc[1] + 1.0) / (2.0 * L) * pow(s, pc[1]);
        double k11 = 1.0 + (pc[2] - 1.0) * pow(s, pc[1]);
        c->g[kk] = c->g[kk] * alpha / k11 / (d11 / k11 + alpha);
      }
    }
    /* ZTOP, initialize.f90:1075-1104 */
    if (c->factive[5]) {
      const int nzt = c->cfg.npml[5];
      const double* pc = c->cfg.pmlc[5];
      const int k0 = nz - nzt;
      const double L = z[nz - 1] - z[k0];
      for (int kk = 0; kk < nzt; ++kk) {
        const int kg = k0 + kk;
        const double s = -(z[k0] - z[kg]) / L;
        const double alpha = pml_f0 * (z[nz - 1] - z[kg]) / L;
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            long fp = (long)i + (long)nx * ((long)j + (long)ny * kk);
            long gp = IX(c, i, j, kg);
            double damp = pc[0] * tin[IX(c, i, j, k0)] * (pc[1] + 1.0) / (2.0 * L) * pow(s, pc[1]);
            double kap = 1.0 + (pc[2] - 1.0) * pow(s, pc[1]);
            c->pk[5][fp] = kap;
            c->pb[5][fp] = -alpha - damp / kap;
            c->pa[5][fp] = -damp / (kap * kap);
            c->gradrho0_z[gp] = c->gradrho0_z[gp] * alpha / kap / (damp / kap + alpha);
            c->gradp0_z[gp] = c->gradp0_z[gp] * alpha / kap / (damp / kap + alpha);
            c->spongexyz[gp] = c->spongexyz[gp] + pc[3] * tin[IX(c, i, j, k0)] * pow(s, pc[1] + 1.0);
          }
        double d11 = pc[0] * tin[IX(c, 0, 0, k0)] * (pc[1] + 1.0) / (2.0 * L) * pow(s, pc[1]);
        double k11 = 1.0 + (pc[2] - 1.0) * pow(s, pc[1]);
        c->g[kg] = c->g[kg] * alpha / k11 / (d11 / k11 + alpha);
      }
    }
    if (cfg->use_hpml) {
      /* XBOT / XTOP, initialize.f90:1108-1152 */
      for (int f = 0; f < 2; ++f) {
        if (!c->factive[f]) continue;
        const int npl = c->cfg.npml[f];
        const double* pc = c->cfg.pmlc[f];
        const int i0 = f ? nx - npl : 0;
        const int iin = f ? i0 : npl - 1; /* tempin sample plane */
        const double L = f ? c->xaxis[nx - 1] - c->xaxis[i0] : c->xaxis[npl - 1] - c->xaxis[0];
        for (int ii = 0; ii < npl; ++ii) {
          const int ig = i0 + ii;
          const double s = f ? -(c->xaxis[i0] - c->xaxis[ig]) / L : (c->xaxis[npl - 1] - c->xaxis[ig]) / L;
          const double alpha = f ? pml_f0 * (c->xaxis[nx - 1] - c->xaxis[ig]) / L
                                 : pml_f0 * (c->xaxis[ig] - c->xaxis[0]) / L;
          for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
              long fp = (long)ii + (long)npl * ((long)j + (long)ny * k);
              long gp = IX(c, ig, j, k);
              double tv = tin[IX(c, iin, j, k)];
              double damp = pc[0] * tv * (pc[1] + 1.0) / (2.0 * L) * pow(s, pc[1]);
              double kap = 1.0 + (pc[2] - 1.0) * pow(s, pc[1]);
              c->pk[f][fp] = kap;
              c->pb[f][fp] = -alpha - damp / kap;
              c->pa[f][fp] = -damp / (kap * kap);
              c->spongexyz[gp] = c->spongexyz[gp] + pc[3] * tv * pow(s, pc[1] + 1.0);
            }
        }
      }
      /* YBOT / YTOP, initialize.f90:1154-1199 */
      for (int f = 2; f < 4; ++f) {
        if (!c->factive[f]) continue;
        const int npl = c->cfg.npml[f];
        const double* pc = c->cfg.pmlc[f];
        const int j0 = (f == 3) ? ny - npl : 0;
        const int jin = (f == 3) ? j0 : npl - 1;
        const double L = (f == 3) ? c->yaxis[ny - 1] - c->yaxis[j0] : c->yaxis[npl - 1] - c->yaxis[0];
        for (int jj = 0; jj < npl; ++jj) {
          const int jg = j0 + jj;
          const double s = (f == 3) ? (-(c->yaxis[j0] - c->yaxis[jg])) / L
                                    : (c->yaxis[npl - 1] - c->yaxis[jg]) / L;
          const double alpha = (f == 3) ? pml_f0 * (c->yaxis[ny - 1] - c->yaxis[jg]) / L
                                        : pml_f0 * (c->yaxis[jg] - c->yaxis[0]) / L;
          for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) {
              long fp = (long)i + (long)nx * ((long)jj + (long)npl * k);
              long gp = IX(c, i, jg, k);
              double tv = tin[IX(c, i, jin, k)];
              double damp = pc[0] * tv * (pc[1] + 1.0) / (2.0 * L) * pow(s, pc[1]);
              double kap = 1.0 + (pc[2] - 1.0) * pow(s, pc[1]);
              c->pk[f][fp] = kap;
              c->pb[f][fp] = -alpha - damp / kap;
              c->pa[f][fp] = -damp / (kap * kap);
              c->spongexyz[gp] = c->spongexyz[gp] + pc[3] * tv * pow(s, pc[1] + 1.0);
            }
        }
      }
    }
  }

  if (c->mhd) { /* initialize.f90:1219-1228 */
    orc_ddx(c, c->rho0, c->gradrho0_x, bc);
    orc_ddx(c, c->p0, c->gradp0_x, bc);
    orc_ddy(c, c->rho0, c->gradrho0_y, bc);
    orc_ddy(c, c->p0, c->gradp0_y, bc);
  }

  /* INITIALIZE_STEP, initialize.f90:1366-1381 */
  {
    double opt[5] = {1.0, 0.5, 0.1665579725791151184, 0.03950410250586729987,
                     0.007810706393337838236};
    double b[5];
    b[4] = opt[1];
    b[3] = opt[2] / b[4];
    b[2] = opt[3] / (b[4] * b[3]);
    b[1] = opt[4] / (b[4] * b[3] * b[2]);
    b[0] = 0.0;
    for (int s = 0; s < 5; ++s) c->betas[s] = b[s] * c->deltat;
  }
  return c;
}

void orc_destroy(orc_ctx* c) {
  if (c) {
    free(c->fmap);
    free(c->sdep);
  }
  if (c) free(c->damp_rates);
  if (!c) return;
  double* ptrs[] = {c->xaxis, c->yaxis, c->zaxis, c->g, c->cq, c->stretch, c->unstretch,
                    c->spongez, c->rho0, c->p0, c->cs, c->c2, c->rhoinv, c->c2rho0,
                    c->gradrho0_z, c->gradp0_z, c->spongexyz, c->box, c->boy, c->boz,
                    c->curlbox, c->curlboy, c->curlboz, c->gradrho0_x, c->gradrho0_y,
                    c->gradp0_x, c->gradp0_y, c->reduction, c->a, c->u, c->scr, c->forcing};
  for (size_t q = 0; q < sizeof ptrs / sizeof ptrs[0]; ++q) free(ptrs[q]);
  for (int q = 0; q < NWORK; ++q) free(c->w[q]);
  for (int f = 0; f < 6; ++f) {
    free(c->pa[f]);
    free(c->pb[f]);
    free(c->pk[f]);
    free(c->pvars[f]);
    free(c->ppsi[f]);
    free(c->pscr[f]);
  }
  free(c);
}

/* ------------------------------------------------------------------------ */
/* Named access for the tests                                                */
/* ------------------------------------------------------------------------ */
double orc_scalar(const orc_ctx* c, const char* n) {
  if (!strcmp(n, "dimc")) return c->dimc;
  if (!strcmp(n, "dimrho")) return c->dimrho;
  if (!strcmp(n, "diml")) return c->diml;
  if (!strcmp(n, "dimb")) return c->dimb;
  if (!strcmp(n, "stretchx")) return c->stretchx;
  if (!strcmp(n, "stretchy")) return c->stretchy;
  if (!strcmp(n, "deltat")) return c->deltat;
  if (!strcmp(n, "xyfreq")) return c->xyfreq;
  if (!strcmp(n, "zfreq")) return c->zfreq;
  if (!strncmp(n, "beta", 4)) {
    int s = atoi(n + 4);
    if (s >= 1 && s <= 5) return c->betas[s - 1];
  }
  return NAN;
}

static const double* named(const orc_ctx* c, const char* n, long* len) {
  const long N = c->N;
  struct {
    const char* name;
    const double* p;
    long n;
  } t[] = {
      {"xaxis", c->xaxis, c->nx},         {"yaxis", c->yaxis, c->ny},
      {"zaxis", c->zaxis, c->nz},         {"g", c->g, c->nz},
      {"cq", c->cq, c->nz},               {"stretch", c->stretch, c->nz},
      {"unstretch", c->unstretch, c->nz}, {"rho0", c->rho0, N},
      {"p0", c->p0, N},                   {"c2", c->c2, N},
      {"rhoinv", c->rhoinv, N},           {"c2rho0", c->c2rho0, N},
      {"gradrho0_z", c->gradrho0_z, N},   {"gradp0_z", c->gradp0_z, N},
      {"spongexyz", c->spongexyz, N},     {"box", c->box, N},
      {"boy", c->boy, N},                 {"boz", c->boz, N},
      {"curlbox", c->curlbox, N},         {"curlboy", c->curlboy, N},
      {"curlboz", c->curlboz, N},         {"gradrho0_x", c->gradrho0_x, N},
      {"gradrho0_y", c->gradrho0_y, N},   {"gradp0_x", c->gradp0_x, N},
      {"gradp0_y", c->gradp0_y, N},       {"reduction", c->reduction, N},
      {"forcing", c->forcing, N},
      {"damping_rates", c->damp_rates, (long)c->ny * (c->nx / 2 + 1)},
  };
  for (size_t q = 0; q < sizeof t / sizeof t[0]; ++q)
    if (!strcmp(n, t[q].name)) {
      *len = t[q].p ? t[q].n : -1;
      return t[q].p;
    }
  /* PML coefficient arrays: pml_a0 .. pml_k5 */
  if (!strncmp(n, "pml_", 4) && strlen(n) == 6) {
    int f = n[5] - '0';
    if (f >= 0 && f < 6) {
      *len = c->fsize[f];
      if (n[4] == 'a') return c->pa[f];
      if (n[4] == 'b') return c->pb[f];
      if (n[4] == 'k') return c->pk[f];
    }
  }
  *len = -1;
  return NULL;
}

long orc_array_len(const orc_ctx* c, const char* n) {
  long len;
  named(c, n, &len);
  return len;
}

int orc_get_array(const orc_ctx* c, const char* n, double* out) {
  long len;
  const double* p = named(c, n, &len);
  if (!p || len < 0) return -1;
  memcpy(out, p, sizeof(double) * len);
  return 0;
}

void orc_put_state(orc_ctx* c, const double* a) { memcpy(c->a, a, sizeof(double) * c->nv * c->N); }
void orc_get_state(const orc_ctx* c, double* a) { memcpy(a, c->a, sizeof(double) * c->nv * c->N); }
void orc_get_temp_step(const orc_ctx* c, double* u) { memcpy(u, c->u, sizeof(double) * c->nv * c->N); }
void orc_get_scr(const orc_ctx* c, double* k) { memcpy(k, c->scr, sizeof(double) * c->nv * c->N); }
int orc_put_psi(orc_ctx* c, int f, const double* psi) {
  if (f < 0 || f > 5) return -1;
  memcpy(c->pvars[f], psi, sizeof(double) * c->npsi * c->fsize[f]);
  return 0;
}
int orc_get_psi(const orc_ctx* c, int f, double* psi) {
  if (f < 0 || f > 5) return -1;
  memcpy(psi, c->pvars[f], sizeof(double) * c->npsi * c->fsize[f]);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Source: driver.f90:133-149 (PULSE = 1, not PSR).                           */
/* ------------------------------------------------------------------------ */
void orc_forcing(orc_ctx* c, int time) {
  const orc_cfg* q = &c->cfg;
  const double t = time * q->timestep;
  const double diml = c->diml;
  const double st = sin(2.0 * PI * (t - q->pulse_t1 + 2.0 * q->pulse_t0) / q->pulse_t0);
  const double et = exp(-1.0 * pow(t - q->pulse_t1, 2) / (2.0 * (q->pulse_st * q->pulse_st)));
  const double srd = q->pulse_sr / diml, szd = q->pulse_sz / diml;
#pragma omp parallel for
  for (int k = 0; k < c->nz; ++k)
    for (int j = 0; j < c->ny; ++j)
      for (int i = 0; i < c->nx; ++i) {
        double X = c->xaxis[i] * q->xlength / RSUN - q->pulse_px / diml;
        double Y = c->yaxis[j] * q->ylength / RSUN - q->pulse_py / diml;
        double Z = c->zaxis[k] - (q->pulse_pz / diml + 1.0);
        c->forcing[IX(c, i, j, k)] = q->pulse_amp / 0.56458272 / c->dimc * st * et *
                                     exp(-1.0 * (X * X + Y * Y) / (2.0 * (srd * srd))) *
                                     exp(-1.0 * (Z * Z) / (2.0 * (szd * szd)));
      }
}

/* ------------------------------------------------------------------------ */
/* CPML on one derivative family of one face: RHS.f90:50-93,424-546.          */
/* RHSpsi = a*d + b*psi, then d = d/kappa + psi (Appendix order: psi RHS from */
/* the raw derivative first).                                                */
/* ------------------------------------------------------------------------ */
static void pml_apply(orc_ctx* c, int f, int q, double* d) {
  if (!c->factive[f]) return;
  const int fx = c->fdim[f][0], fy = c->fdim[f][1], fz = c->fdim[f][2];
  const int ox = c->foff[f][0], oy = c->foff[f][1], oz = c->foff[f][2];
  const long fs = c->fsize[f];
  double* psi = c->ppsi[f] + q * fs;
  double* rps = c->pscr[f] + q * fs;
#pragma omp parallel for
  for (int kk = 0; kk < fz; ++kk)
    for (int jj = 0; jj < fy; ++jj)
      for (int ii = 0; ii < fx; ++ii) {
        long fp = (long)ii + (long)fx * ((long)jj + (long)fy * kk);
        long gp = IX(c, ox + ii, oy + jj, oz + kk);
        rps[fp] = c->pa[f][fp] * d[gp] + c->pb[f][fp] * psi[fp];
        d[gp] = d[gp] / c->pk[f][fp] + psi[fp];
      }
}

/* RHS zeroing: RHS.f90:154-164, 647-657. */
/* ------------------------------------------------------------------------ */
/* FFT damping (DAMPING = 1).  INIT_DAMPING (initialize.f90:1276-1336) sets  */
/* damping_rates(nx/2+1, ny); damp_velocity (all_modules.f90:1802-1888)      */
/* returns, per z plane, the unnormalised c2r transform of the r2c transform */
/* of the field times the rates; MP_QUIET_3D / MP_MHD_3D subtract it from    */
/* the velocity RHS after the boundary zeroing (RHS.f90:166-172, 659-665).   */
/* Here the transforms are direct DFT sums (separable, FP64): the r2c half   */
/* spectrum extended by Hermitian symmetry, the c2r the real part of the     */
/* full inverse sum.                                                         */
/* ------------------------------------------------------------------------ */
int orc_set_damping(orc_ctx* c, int on) {
  free(c->damp_rates);
  c->damp_rates = NULL;
  if (!on) return 0;
  const int nx = c->nx, ny = c->ny, mx = nx / 2 + 1;
  c->damp_rates = dalloc((long)mx * ny);
  if (!c->damp_rates) return -1;
  const double constx = 2.0 * PI * RSUN / c->cfg.xlength;
  const double consty = 2.0 * PI * RSUN / c->cfg.ylength;
  const double kayref = 902.0;
  for (int jj = 1; jj <= ny; ++jj) {
    /* rows jj <= ny/2+1 use jj-1, the rest ny-jj+1 (initialize.f90:1300-1310) */
    const double kyv = jj <= ny / 2 + 1 ? (double)(jj - 1) : (double)(ny - jj + 1);
    for (int ii = 1; ii <= mx; ++ii) {
      const double kxv = (double)(ii - 1);
      const double kay = pow(constx * constx * kxv * kxv + consty * consty * kyv * kyv, 0.5);
      c->damp_rates[(long)(jj - 1) * mx + (ii - 1)] =
          c->diml / c->dimc / nx / ny * 1.0e-6 * (500.0 * pow(kay / kayref, 2.2)) * 20.0;
    }
  }
  return 0;
}

void orc_damp_term(const orc_ctx* c, const double* v, double* term) {
  const int nx = c->nx, ny = c->ny, mx = nx / 2 + 1;
  const long nxy = (long)nx * ny;
  /* twiddles of the transforms (the exact 2*pi, not the reference's pi) */
  const double two_pi = 6.283185307179586476925286766559;
  double *cx = dalloc(nx), *sx = dalloc(nx), *cy = dalloc(ny), *sy = dalloc(ny);
  for (int q = 0; q < nx; ++q) {
    cx[q] = cos(two_pi * q / nx);
    sx[q] = sin(two_pi * q / nx);
  }
  for (int q = 0; q < ny; ++q) {
    cy[q] = cos(two_pi * q / ny);
    sy[q] = sin(two_pi * q / ny);
  }
  const double* R = c->damp_rates;
#pragma omp parallel for
  for (int k = 0; k < c->nz; ++k) {
    const double* p = v + (long)k * nxy;
    double* o = term + (long)k * nxy;
    double *ar = dalloc(nxy), *ai = dalloc(nxy), *br = dalloc(nxy), *bi = dalloc(nxy);
    /* forward along x: A(y, kx) */
    for (int y = 0; y < ny; ++y)
      for (int kx = 0; kx < nx; ++kx) {
        double sr = 0.0, si = 0.0;
        for (int x = 0; x < nx; ++x) {
          const int q = (int)(((long)kx * x) % nx);
          sr += p[(long)y * nx + x] * cx[q];
          si -= p[(long)y * nx + x] * sx[q];
        }
        ar[(long)y * nx + kx] = sr;
        ai[(long)y * nx + kx] = si;
      }
    /* forward along y and the rates: B(ky, kx) = rate * sum_y A e^{-i..} */
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < nx; ++kx) {
        double sr = 0.0, si = 0.0;
        for (int y = 0; y < ny; ++y) {
          const int q = (int)(((long)ky * y) % ny);
          const double xr = ar[(long)y * nx + kx], xi = ai[(long)y * nx + kx];
          sr += xr * cy[q] + xi * sy[q];
          si += xi * cy[q] - xr * sy[q];
        }
        /* half spectrum kx <= nx/2; above it the Hermitian partner (nx-kx, ny-ky) */
        const double r = kx < mx ? R[(long)ky * mx + kx] : R[(long)((ny - ky) % ny) * mx + (nx - kx)];
        br[(long)ky * nx + kx] = sr * r;
        bi[(long)ky * nx + kx] = si * r;
      }
    /* inverse along y: A(y, kx) = sum_ky B e^{+i..} */
    for (int y = 0; y < ny; ++y)
      for (int kx = 0; kx < nx; ++kx) {
        double sr = 0.0, si = 0.0;
        for (int ky = 0; ky < ny; ++ky) {
          const int q = (int)(((long)ky * y) % ny);
          const double xr = br[(long)ky * nx + kx], xi = bi[(long)ky * nx + kx];
          sr += xr * cy[q] - xi * sy[q];
          si += xi * cy[q] + xr * sy[q];
        }
        ar[(long)y * nx + kx] = sr;
        ai[(long)y * nx + kx] = si;
      }
    /* inverse along x, real part */
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        double sr = 0.0;
        for (int kx = 0; kx < nx; ++kx) {
          const int q = (int)(((long)kx * x) % nx);
          sr += ar[(long)y * nx + kx] * cx[q] - ai[(long)y * nx + kx] * sx[q];
        }
        o[(long)y * nx + x] = sr;
      }
    free(ar);
    free(ai);
    free(br);
    free(bi);
  }
  free(cx);
  free(sx);
  free(cy);
  free(sy);
}

/* PULSE = 0 (RHS.f90:146-147, 640-641): RHSv_z += forcing * source_dep(kk),
 * forcing = LAGRANGE_INTERP x source_damp of this step (driver.f90:130-132),
 * given by the caller.  map NULL returns to the pulse. */
int orc_set_forcing_map(orc_ctx* c, const double* map, const double* sdep) {
  if (!map) {
    free(c->fmap);
    c->fmap = NULL;
    return 0;
  }
  if (!c->fmap) c->fmap = dalloc((long)c->nx * c->ny);
  if (!c->sdep) c->sdep = dalloc(c->nz);
  if (!c->fmap || !c->sdep) return -1;
  memcpy(c->fmap, map, sizeof(double) * c->nx * c->ny);
  if (sdep) memcpy(c->sdep, sdep, sizeof(double) * c->nz);
  return 0;
}

/* RHSv -= damp_velocity(v) for the three velocities of the stage value. */
static void damp_rhs(orc_ctx* c) {
  if (!c->damp_rates) return;
  const long N = c->N;
  double* t = dalloc(N);
  for (int f = 1; f <= 3; ++f) {
    orc_damp_term(c, c->u + f * N, t);
    double* s = c->scr + f * N;
    for (long i = 0; i < N; ++i) s[i] = s[i] - t[i];
  }
  free(t);
}

static void zero_planes(orc_ctx* c) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int v = 0; v < c->nv; ++v) {
    double* s = c->scr + (long)v * c->N;
    if (!c->cfg.periodic) {
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
          s[IX(c, nx - 1, j, k)] = 0.0;
          s[IX(c, 0, j, k)] = 0.0;
        }
      for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i) {
          s[IX(c, i, ny - 1, k)] = 0.0;
          s[IX(c, i, 0, k)] = 0.0;
        }
    }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        s[IX(c, i, j, nz - 1)] = 0.0;
        s[IX(c, i, j, 0)] = 0.0;
      }
  }
}

/* MP_QUIET_3D, RHS.f90:18-175 (the FLOWS branch excluded). */
static void rhs_quiet(orc_ctx* c) {
  const long N = c->N;
  const int bh = c->cfg.periodic ? 5 : 1;
  const double *rho = c->u, *vx = c->u + N, *vy = c->u + 2 * N, *vz = c->u + 3 * N,
               *p = c->u + 4 * N;
  double *dvxdx = c->w[0], *dvydy = c->w[1], *dvzdz = c->w[2];
  double *gpx = c->w[3], *gpy = c->w[4], *gpz = c->w[5];
  orc_ddx(c, vx, dvxdx, bh);
  orc_ddy(c, vy, dvydy, bh);
  orc_ddz(c, vz, dvzdz, 1);
  orc_ddx(c, p, gpx, bh);
  orc_ddy(c, p, gpy, bh);
  orc_ddz(c, p, gpz, 1);
  for (int f = 4; f < 6; ++f) {
    pml_apply(c, f, 1, gpz);
    pml_apply(c, f, 0, dvzdz);
  }
  for (int f = 0; f < 2; ++f) {
    pml_apply(c, f, 1, gpx);
    pml_apply(c, f, 0, dvxdx);
  }
  for (int f = 2; f < 4; ++f) {
    pml_apply(c, f, 1, gpy);
    pml_apply(c, f, 0, dvydy);
  }
  double *s_rho = c->scr, *s_vx = c->scr + N, *s_vy = c->scr + 2 * N, *s_vz = c->scr + 3 * N,
         *s_p = c->scr + 4 * N;
  const int pd = c->cfg.pulse_dir;
#pragma omp parallel for
  for (int k = 0; k < c->nz; ++k)
    for (long q = 0; q < (long)c->nx * c->ny; ++q) {
      long i = (long)k * c->nx * c->ny + q;
      double div = dvxdx[i] + dvydy[i] + dvzdz[i];
      double sp = c->spongexyz[i];
      s_rho[i] = -c->gradrho0_z[i] * vz[i] - c->rho0[i] * div - sp * rho[i];
      s_vx[i] = -c->rhoinv[i] * gpx[i] - sp * vx[i];
      s_vy[i] = -c->rhoinv[i] * gpy[i] - sp * vy[i];
      s_vz[i] = -c->rhoinv[i] * (gpz[i] + rho[i] * c->g[k]) - vz[i] * sp;
      if (pd == 1) s_vx[i] = s_vx[i] + c->forcing[i];
      if (pd == 2) s_vy[i] = s_vy[i] + c->forcing[i];
      if (pd == 3) s_vz[i] = s_vz[i] + c->forcing[i];
      if (c->fmap) s_vz[i] = s_vz[i] + c->fmap[q] * c->sdep[k];
      s_p[i] = -c->c2rho0[i] * div - vz[i] * c->gradp0_z[i] - sp * p[i];
    }
  zero_planes(c);
  damp_rhs(c);
}

/* MP_MHD_3D, RHS.f90:363-668 (the FLOWS branch excluded). */
static void rhs_mhd(orc_ctx* c) {
  const long N = c->N;
  const int bh = c->cfg.periodic ? 5 : 1;
  const double *rho = c->u, *vx = c->u + N, *vy = c->u + 2 * N, *vz = c->u + 3 * N,
               *p = c->u + 4 * N, *bx = c->u + 5 * N, *by = c->u + 6 * N, *bz = c->u + 7 * N;
  double *f1 = c->w[0], *f2 = c->w[1], *f3 = c->w[2];
  double *dvxdx = c->w[3], *dvydy = c->w[4], *dvzdz = c->w[5];
  double *dxf3 = c->w[6], *gpy = c->w[7], *dzf1 = c->w[8];
  double *gpx = c->w[9], *dyf3 = c->w[10], *dzf2 = c->w[11];
  double *dxf2 = c->w[12], *dybz = c->w[13], *dzbx = c->w[14];
  double *dxby = c->w[15], *dybx = c->w[16], *gpz = c->w[17];
  double *dxbz = c->w[18], *dyf1 = c->w[19];
  double* dzby = c->w[20];
  const double *box = c->box, *boy = c->boy, *boz = c->boz;
#pragma omp parallel for
  for (long i = 0; i < N; ++i) {
    f1[i] = -vz[i] * boy[i] + vy[i] * boz[i];
    f2[i] = -vx[i] * boz[i] + vz[i] * box[i];
    f3[i] = -vy[i] * box[i] + vx[i] * boy[i];
  }
  orc_ddx(c, vx, dvxdx, bh);
  orc_ddy(c, vy, dvydy, bh);
  orc_ddz(c, vz, dvzdz, 1);
  orc_ddx(c, f3, dxf3, bh);
  orc_ddy(c, p, gpy, bh);
  orc_ddz(c, f1, dzf1, 1);
  orc_ddx(c, p, gpx, bh);
  orc_ddy(c, f3, dyf3, bh);
  orc_ddz(c, f2, dzf2, 1);
  orc_ddx(c, f2, dxf2, bh);
  orc_ddy(c, bz, dybz, bh);
  orc_ddz(c, bx, dzbx, 1);
  orc_ddx(c, by, dxby, bh);
  orc_ddy(c, bx, dybx, bh);
  orc_ddz(c, p, gpz, 1);
  orc_ddx(c, bz, dxbz, bh);
  orc_ddy(c, f1, dyf1, bh);
  orc_ddz(c, by, dzby, 1);
  /* CPML families, pointer map initialize.f90:742-746,755-772,835-871 */
  for (int f = 4; f < 6; ++f) {
    pml_apply(c, f, 1, gpz);
    pml_apply(c, f, 0, dvzdz);
  }
  for (int f = 0; f < 2; ++f) {
    pml_apply(c, f, 1, gpx);
    pml_apply(c, f, 0, dvxdx);
  }
  for (int f = 2; f < 4; ++f) {
    pml_apply(c, f, 1, gpy);
    pml_apply(c, f, 0, dvydy);
  }
  for (int f = 4; f < 6; ++f) {
    pml_apply(c, f, 4, dzf2);
    pml_apply(c, f, 5, dzf1);
    pml_apply(c, f, 2, dzbx);
    pml_apply(c, f, 3, dzby);
  }
  for (int f = 0; f < 2; ++f) {
    pml_apply(c, f, 4, dxf3);
    pml_apply(c, f, 5, dxf2);
    pml_apply(c, f, 2, dxby);
    pml_apply(c, f, 3, dxbz);
  }
  for (int f = 2; f < 4; ++f) {
    pml_apply(c, f, 4, dyf3);
    pml_apply(c, f, 5, dyf1);
    pml_apply(c, f, 2, dybx);
    pml_apply(c, f, 3, dybz);
  }
  double* s = c->scr;
  const int pd = c->cfg.pulse_dir;
#pragma omp parallel for
  for (int k = 0; k < c->nz; ++k)
    for (long q = 0; q < (long)c->nx * c->ny; ++q) {
      long i = (long)k * c->nx * c->ny + q;
      const double red = c->reduction[i], sp = c->spongexyz[i], ri = c->rhoinv[i];
      double rbx = red * (dyf3[i] - dzf2[i]);
      double rby = red * (dzf1[i] - dxf3[i]);
      double rbz = red * (dxf2[i] - dyf1[i]);
      double div = dvxdx[i] + dvydy[i] + dvzdz[i];
      double cbx = dybz[i] - dzby[i];
      double cby = -dxbz[i] + dzbx[i];
      double cbz = dxby[i] - dybx[i];
      double rc = -c->gradrho0_x[i] * vx[i] - c->gradrho0_y[i] * vy[i] -
                  c->gradrho0_z[i] * vz[i] - c->rho0[i] * div - sp * rho[i];
      double rp = -c->c2rho0[i] * div - vx[i] * c->gradp0_x[i] - vy[i] * c->gradp0_y[i] -
                  vz[i] * c->gradp0_z[i] - sp * p[i];
      /* CROSS(curlb, B0), derivatives.f90:646-655 */
      double c1 = cby * boz[i] - cbz * boy[i];
      double c2 = box[i] * cbz - cbx * boz[i];
      double c3 = cbx * boy[i] - cby * box[i];
      double rvx = -ri * (gpx[i] - (c1 * red)) - sp * vx[i];
      double rvy = -ri * (gpy[i] - (c2 * red)) - sp * vy[i];
      double rvz = -ri * (gpz[i] - (c3 * red)) - sp * vz[i];
      /* CROSS(curlB0, b) */
      double e1 = c->curlboy[i] * bz[i] - c->curlboz[i] * by[i];
      double e2 = bx[i] * c->curlboz[i] - c->curlbox[i] * bz[i];
      double e3 = c->curlbox[i] * by[i] - c->curlboy[i] * bx[i];
      rvx = rvx + ri * e1 * red;
      rvy = rvy + ri * e2 * red;
      rvz = rvz + ri * e3 * red;
      rvz = rvz - ri * rho[i] * c->g[k];
      if (pd == 1) rvx = rvx + c->forcing[i];
      if (pd == 2) rvy = rvy + c->forcing[i];
      if (pd == 3) rvz = rvz + c->forcing[i];
      if (c->fmap) rvz = rvz + c->fmap[q] * c->sdep[k];
      s[i] = rc;
      s[N + i] = rvx;
      s[2 * N + i] = rvy;
      s[3 * N + i] = rvz;
      s[4 * N + i] = rp;
      s[5 * N + i] = rbx;
      s[6 * N + i] = rby;
      s[7 * N + i] = rbz;
    }
  zero_planes(c);
  damp_rhs(c);
}

/* Stage values + RHS.  step.f90:27-37 then the RHS dispatch (step.f90:39-46). */
void orc_rhs_stage(orc_ctx* c, int stage) {
  const long M = c->nv * c->N;
  if (stage == 1) {
    memcpy(c->u, c->a, sizeof(double) * M);
  } else {
    const double b = c->betas[stage - 1];
#pragma omp parallel for
    for (long i = 0; i < M; ++i) c->u[i] = c->a[i] + b * c->scr[i];
  }
  for (int f = 0; f < 6; ++f) {
    if (!c->factive[f]) continue;
    const long M2 = c->npsi * c->fsize[f];
    if (stage == 1) {
      memcpy(c->ppsi[f], c->pvars[f], sizeof(double) * M2);
    } else {
      const double b = c->betas[stage - 1];
      for (long i = 0; i < M2; ++i) c->ppsi[f][i] = c->pvars[f][i] + b * c->pscr[f][i];
    }
  }
  if (c->mhd)
    rhs_mhd(c);
  else
    rhs_quiet(c);
}

/* STEP, step.f90:20-63. */
void orc_step(orc_ctx* c) {
  for (int s = 1; s <= 5; ++s) orc_rhs_stage(c, s);
  const long M = c->nv * c->N;
  const double dt = c->deltat;
#pragma omp parallel for
  for (long i = 0; i < M; ++i) c->a[i] = c->a[i] + dt * c->scr[i];
  for (int f = 0; f < 6; ++f) {
    if (!c->factive[f]) continue;
    const long M2 = c->npsi * c->fsize[f];
    for (long i = 0; i < M2; ++i) c->pvars[f][i] = c->pvars[f][i] + dt * c->pscr[f][i];
  }
}

/* FILTER_VARS, derivatives.f90:659-703. */
void orc_filter_vars(orc_ctx* c, int time) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const long N = c->N;
  double* filt = c->w[0];
  double* tmp = c->w[1];
  const int gxy = c->cfg.g_xy_filt, gz = c->cfg.g_z_filt;
  for (int ll = 0; ll < c->nv; ++ll) {
    double* a = c->a + (long)ll * N;
    if (time % c->xyfreq == 0) {
      if (!c->cfg.periodic) { /* y edge smoothing, derivatives.f90:671-675 */
        memcpy(tmp, a, sizeof(double) * N);
        for (int k = 0; k < nz; ++k)
          for (int i = 0; i < nx; ++i) {
            for (int j = ny - gxy; j <= ny - 2; ++j)
              a[IX(c, i, j, k)] = (tmp[IX(c, i, j + 1, k)] + tmp[IX(c, i, j - 1, k)]) / 2.0;
            for (int j = 1; j <= gxy - 1; ++j)
              a[IX(c, i, j, k)] = (tmp[IX(c, i, j - 1, k)] + tmp[IX(c, i, j + 1, k)]) / 2.0;
          }
      }
      orc_filty(c, a, filt);
      if (gxy == 3)
        for (long i = 0; i < N; ++i) a[i] = a[i] - c->cfg.s_xy * filt[i];
      else
        memcpy(a, filt, sizeof(double) * N);
      if (!c->cfg.periodic) { /* x edge smoothing, derivatives.f90:680-684 */
        memcpy(tmp, a, sizeof(double) * N);
        for (int k = 0; k < nz; ++k)
          for (int j = 0; j < ny; ++j) {
            for (int i = nx - gxy; i <= nx - 2; ++i)
              a[IX(c, i, j, k)] = (tmp[IX(c, i + 1, j, k)] + tmp[IX(c, i - 1, j, k)]) / 2.0;
            for (int i = 1; i <= gxy - 1; ++i)
              a[IX(c, i, j, k)] = (tmp[IX(c, i - 1, j, k)] + tmp[IX(c, i + 1, j, k)]) / 2.0;
          }
      }
      orc_filtx(c, a, filt);
      if (gxy == 3)
        for (long i = 0; i < N; ++i) a[i] = a[i] - c->cfg.s_xy * filt[i];
      else
        memcpy(a, filt, sizeof(double) * N);
    }
    if (time % c->zfreq == 0) { /* derivatives.f90:690-697 */
      memcpy(tmp, a, sizeof(double) * N);
      for (int k = nz - gz; k <= nz - 2; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i)
            a[IX(c, i, j, k)] = (tmp[IX(c, i, j, k + 1)] + tmp[IX(c, i, j, k - 1)]) / 2.0;
      memcpy(tmp, a, sizeof(double) * N);
      for (int k = 1; k <= gz - 1; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i)
            a[IX(c, i, j, k)] = (tmp[IX(c, i, j, k - 1)] + tmp[IX(c, i, j, k + 1)]) / 2.0;
      orc_filtz(c, a, filt);
      if (gz == 3)
        for (long i = 0; i < N; ++i) a[i] = a[i] - c->cfg.s_z * filt[i];
      else
        memcpy(a, filt, sizeof(double) * N);
    }
  }
}

/* norm_KE (driver.f90:161 with norm() all_modules.f90:684-706) and the
 * negativity checks on the stage-5 state (driver.f90:180; rho and p alias
 * temp_step, initialize.f90:781,785). */
void orc_diag(const orc_ctx* c, double out[3]) {
  const long N = c->N;
  double sum = 0.0;
  for (long i = 0; i < N; ++i) {
    double v2 = c->a[N + i] * c->a[N + i] + c->a[2 * N + i] * c->a[2 * N + i] +
                c->a[3 * N + i] * c->a[3 * N + i];
    double m = (c->rho0[i] + c->a[i]) * v2 / 2.0;
    sum = sum + m * m;
  }
  double nrm = pow(sum / ((double)c->nx * (double)c->ny * (double)c->nz), 0.5);
  out[0] = nrm * (c->dimc * c->dimc) * c->dimrho;
  double mr = INFINITY, mp = INFINITY;
  for (long i = 0; i < N; ++i) {
    double r = c->u[i] + c->rho0[i];
    double q = c->u[4 * N + i] + c->p0[i];
    if (r < mr) mr = r;
    if (q < mp) mp = q;
  }
  out[1] = mr;
  out[2] = mp;
}

/* Driver loop, driver.f90:123-202 (I/O and the failure dump are host work). */
void orc_run(orc_ctx* c, int n0, int n1, double* hist) {
  for (int nn = n0; nn <= n1; ++nn) {
    if (c->cfg.pulse_dir) orc_forcing(c, nn);
    orc_step(c);
    if (hist) orc_diag(c, hist + 3 * (long)(nn - n0));
    if (nn != 0) orc_filter_vars(c, nn);
  }
}
