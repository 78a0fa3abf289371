/*
 * m3d_rotate.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of rt_patch_t%rotate_array
 * (SRC/rt_patch_mod.f90:3809-3863 of experiments/OpenMULTI3D/ in
 * Bergemann-Group-development/dispatch), the checker for dispatch_amd/csrc/rotate.hip.
 *
 * In the reference's precision: delta = vec(1:2)/(vec(3) ds(1:2)) in real(8) stored
 * as real(4); height = z ds(3) stored as real(4); deltak = (bot2pivot - height) delta
 * (real(4)) + ray_offset (real(8)), stored as real(4); floor, fractions and the four
 * weights in real(4); the weighted sum accumulated in the order a, b, c, d.  Every
 * product and sum rounded apart (the Makefile's -ffp-contract=off).
 *
 * Pinning (tests/test_rotate_cpu.py): a vertical ray leaves the cubes unchanged; an
 * integer shift is np.roll; every plane's sum is conserved; and a plane that is
 * linear in x within the shifted stencil is shifted exactly.
 */
#include <math.h>
#include <stddef.h>
#include "m3d_rotate.h"

void orc_rotate_array(const double vec[3], const double ds[3], double bot2pivot, const double ray_offset[2],
                      int nf, int nx, int ny, int nz, int nk, const float *in, float *out)
{
  const float delta1 = (float)(vec[0] / (vec[2] * ds[0]));
  const float delta2 = (float)(vec[1] / (vec[2] * ds[1]));
  const float b2p = (float)bot2pivot;
  for (int z = 1; z <= nz; z++) {
    const float height = (float)(z * ds[2]);
    const float t = b2p - height;
    const float d1 = (float)((double)(t * delta1) + ray_offset[0]);
    const float d2 = (float)((double)(t * delta2) + ray_offset[1]);
    const int up1 = (int)floorf(d1), up2 = (int)floorf(d2);
    const float f1 = d1 - (float)up1, f2 = d2 - (float)up2;
    const float a = (1.0f - f1) * (1.0f - f2), b = f1 * (1.0f - f2), c = (1.0f - f1) * f2, d = f1 * f2;
    for (int k = 0; k < nk; k++) {
      const size_t pl = ((size_t)k * nz + (z - 1)) * ny * nx * nf;
      for (int y = 1; y <= ny; y++) {
        const int y1 = ((y + up2 - 1) % ny + ny) % ny, y2 = ((y + up2) % ny + ny) % ny;
        for (int x = 1; x <= nx; x++) {
          const int x1 = ((x + up1 - 1) % nx + nx) % nx, x2 = ((x + up1) % nx + nx) % nx;
          for (int f = 0; f < nf; f++) {
            const float *p = in + pl + f;
            float s = a * p[((size_t)y1 * nx + x1) * nf];
            s = b * p[((size_t)y1 * nx + x2) * nf] + s;
            s = c * p[((size_t)y2 * nx + x1) * nf] + s;
            s = d * p[((size_t)y2 * nx + x2) * nf] + s;
            out[pl + ((size_t)(y - 1) * nx + (x - 1)) * nf + f] = s;
          }
        }
      }
    }
  }
}
