/* m3d_eta.h -- TEST INFRASTRUCTURE ONLY: oracle restatement of update_eta_mask
 * (SRC/rt_patch_mod.f90:3671-3802) on dense per-(f, cell) flags. */
#ifndef M3D_ETA_H
#define M3D_ETA_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nf, nx, ny, nz, dir, rx, ry, ghost;  /* ghost: the upstream ghost plane fires (not the domain top) */
  const float *chi, *S, *beta;              /* (z, f, y, x) interior cubes, x fastest */
  const float *gchi, *gS;                   /* (f, y, x) upstream ghost plane */
  uint8_t *src, *mask;                      /* (z, f, y, x) flags: fired here / reaching this cell */
  uint8_t *gsrc;                            /* (f, y, x) flags fired on the ghost plane */
} orc_eta_args;

void orc_update_eta_mask(const orc_eta_args *a);

#ifdef __cplusplus
}
#endif
#endif
