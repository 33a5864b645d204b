/* Verifies that the Fortran bind(C) types of sparc_hip.f90 have the layout
 * of the C structs in include/sparc_hip.h.  Prints OK and exits 0 on match. */
#include <stdio.h>
#include <string.h>

#include "sparc_hip.h"

void fill_structs(sparc_cfg* cfg, sparc_bg* bg, const double* stretch);

#define CHECK(c)                                 \
  do {                                           \
    if (!(c)) {                                  \
      printf("layout mismatch: %s\n", #c);       \
      return 1;                                  \
    }                                            \
  } while (0)

int main(void) {
  sparc_cfg cfg;
  sparc_bg bg;
  double stretch[8] = {1, 2, 3, 4, 5, 6, 7, 8};
  memset(&cfg, 0, sizeof cfg);
  memset(&bg, 0, sizeof bg);
  fill_structs(&cfg, &bg, stretch);
  CHECK(cfg.nx == 101 && cfg.ny == 102 && cfg.nz == 103);
  CHECK(cfg.option == 3 && cfg.precision == SPARC_PREC_FP32 && cfg.periodic == 1);
  CHECK(cfg.use_pml == 1 && cfg.use_hpml == 0);
  for (int f = 0; f < 6; ++f) CHECK(cfg.npml[f] == 11 + f);
  CHECK(cfg.g_xy_filt == 5 && cfg.g_z_filt == 3);
  CHECK(cfg.s_xy == 0.25 && cfg.s_z == 0.75);
  CHECK(cfg.xyfreq == 30 && cfg.zfreq == 10 && cfg.pulse_dir == 2 && cfg.bg3d == 1);
  CHECK(cfg.deltat == 1.5e-4);
  for (int s = 0; s < 5; ++s) CHECK(cfg.betas[s] == (double)s);
  CHECK(cfg.stretchx == 7.25 && cfg.stretchy == 8.5);
  CHECK(cfg.stretch == stretch);
  CHECK(cfg.device == 2);
  CHECK(cfg.tile[0] == 64 && cfg.tile[1] == 8 && cfg.tile[2] == 16);
  CHECK(bg.dimc == 4.5e6 && bg.dimrho == 2.5e-4);
  for (int f = 0; f < 6; ++f) CHECK(bg.pml_a[f] == stretch + f);
  CHECK(bg.src_z == stretch + 6);
  printf("OK sizeof(sparc_cfg)=%zu sizeof(sparc_bg)=%zu\n", sizeof cfg, sizeof bg);
  return 0;
}
