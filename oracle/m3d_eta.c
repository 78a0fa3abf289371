/*
 * m3d_eta.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of update_eta_mask
 * (SRC/rt_patch_mod.f90:3671-3802 of experiments/OpenMULTI3D/ in
 * Bergemann-Group-development/dispatch) on dense per-(f, cell) flags, the checker
 * for the HIP kernel.
 *
 * A pair fires where min(chi, |chi| - q (chi + (1-beta) chi S)) <= 0 (q = 1e-2; on
 * the ghost plane 1-beta = 1).  Pairs that fired and were not yet sources become
 * sources (src only grows) and OR their flag into every cell within rx, ry
 * (periodically) on their own plane and on the next one downwind; the ghost plane's
 * new flags reach the slab's first plane.  The reference stores the flags in rows
 * appended in loop order; the flag sets do not depend on that order (each cell's
 * src changes only by its own fire, mask only grows), which is what the dense form
 * relies on.  Pinned against a line-by-line restatement of the reference's sparse
 * algorithm (tests/test_oracle_eta.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "m3d_eta.h"

static int wrapi(int v, int n) { v %= n; return v < 0 ? v + n : v; }

void orc_update_eta_mask(const orc_eta_args *a)
{
  const int nf = a->nf, nx = a->nx, ny = a->ny, nz = a->nz;
  const size_t nxy = (size_t)nx * ny, npl = nxy * nf, n = npl * nz;
  const float q = 1e-2f;
  uint8_t *nw = (uint8_t *)calloc(n ? n : 1, 1), *gnw = (uint8_t *)calloc(npl ? npl : 1, 1);
  for (size_t i = 0; i < n; i++) {
    const float chi = a->chi[i];
    const float t = fabsf(chi) - q * (chi + (1.0f - a->beta[i]) * chi * a->S[i]);
    const int fire = (chi < t ? chi : t) <= 0.0f;
    nw[i] = (uint8_t)(fire && !a->src[i]);
    if (fire) a->src[i] = 1;
  }
  if (a->ghost)
    for (size_t i = 0; i < npl; i++) {
      const float chi = a->gchi[i];
      const float t = fabsf(chi) - q * (chi + chi * a->gS[i]);
      const int fire = (chi < t ? chi : t) <= 0.0f;
      gnw[i] = (uint8_t)(fire && !a->gsrc[i]);
      if (fire) a->gsrc[i] = 1;
    }
  const int z1 = a->dir > 0 ? 0 : nz - 1;
  for (int z = 0; z < nz; z++)
    for (int f = 0; f < nf; f++)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          int m = 0;
          for (int j = -a->ry; j <= a->ry && !m; j++)
            for (int i = -a->rx; i <= a->rx && !m; i++) {
              const size_t c = (size_t)wrapi(y - j, ny) * nx + wrapi(x - i, nx);
              if (nw[(size_t)z * npl + (size_t)f * nxy + c]) m = 1;
              const int zu = z - a->dir;
              if (zu >= 0 && zu < nz && nw[(size_t)zu * npl + (size_t)f * nxy + c]) m = 1;
              if (a->ghost && z == z1 && gnw[(size_t)f * nxy + c]) m = 1;
            }
          if (m) a->mask[(size_t)z * npl + (size_t)f * nxy + (size_t)y * nx + x] = 1;
        }
  free(nw);
  free(gnw);
}
