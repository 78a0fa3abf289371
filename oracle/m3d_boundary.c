/*
 * m3d_boundary.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the boundary
 * planes and the LTE background factors of experiments/OpenMULTI3D/ in
 * Bergemann-Group-development/dispatch, the checker for
 * dispatch_amd/csrc/background.hip.
 *
 *   lte_background_opacities' beta and Bnu   SRC/rt_patch_mod.f90:2004-2023
 *   set_upstream_boundary, lte_ghost          SRC/rt_patch_mod.f90:3352-3373
 *   set_upstream_boundary, PCHIP bottom ghost SRC/rt_patch_mod.f90:3287-3300, with
 *       pchip_mono4 / pchip_slopes4 of HOP/interpol_mod.f90:193-365 at one plane
 *       below the first three (n = 3, newx = 0)
 *
 * In the reference's precision: h_kT and hv_kT are real(4) results of real(8)
 * products, beta = max(expf(-hv_kT), tiny(1.0)), Bnu = beta/(1 - beta) in real(4);
 * the ghost opacity ratio r500 in real(4).  For the PCHIP ghost the nodes are the
 * real(4) values i*ds(3), i = 1..3, and the point 0: its interval is the first
 * (t = -1, p = 5), the slope at node 2 is never formed (the slopes are needed only
 * from imin = 1 to imax = 1) and stays 0, the slope at node 1 is 2 dy1/dx1, so
 * the value is (1-p) y1 + p y2 + dx1 ((t^3 - 2t^2 + t) m1 + (t^3 - t^2) m2), every
 * product and sum rounded apart (the Makefile's -ffp-contract=off).
 *
 * Pinning: the PCHIP ghost against the reference's compiled pchip_mono
 * (tests/golden/rt_pchip_ghost.npz), bit for bit, and against its closed form
 * 4 y1 - 3 y2; beta and Bnu against exp(-h nu / kT) and 1/(exp(h nu/kT) - 1) in
 * FP64 (tests/test_oracle_boundary.py).
 */
#include <float.h>
#include <math.h>
#include "m3d_boundary.h"

void orc_lte_background(int nf, int ncell, const double *nu, const double *temp, float *beta, float *Bnu)
{
  for (int c = 0; c < ncell; c++) {
    const float h_kT = (float)(ORC_H_K / (double)(float)temp[c]);
    for (int f = 0; f < nf; f++) {
      const float hv = (float)(nu[f] * (double)h_kT);
      float b = expf(-hv);
      if (!(b >= FLT_MIN)) b = FLT_MIN;
      const size_t o = (size_t)f * ncell + c;
      beta[o] = b;
      Bnu[o] = b / (1.0f - b);
    }
  }
}

void orc_lte_ghost(int nf, int nxy, const double *nu, const float *chi1, const float *g500, const float *chi500_1,
                   const float *T_ghost, float *gchi, float *gS, float *gI)
{
  for (int c = 0; c < nxy; c++) {
    const float r500 = g500[c] / chi500_1[c];
    const float h_kT = (float)(ORC_H_K / (double)T_ghost[c]);
    for (int f = 0; f < nf; f++) {
      const size_t o = (size_t)f * nxy + c;
      const float hv = (float)(nu[f] * (double)h_kT);
      float b = expf(-hv);
      if (!(b >= FLT_MIN)) b = FLT_MIN;
      gchi[o] = chi1[o] * r500;
      gS[o] = b / (1.0f - b);
      gI[o] = gS[o];
    }
  }
}

void orc_pchip_ghost(int n, double ds3, const float *y1, const float *y2, const float *y3, float *out)
{
  const float x1 = (float)(1 * ds3), x2 = (float)(2 * ds3);
  const float dx1 = x2 - x1;
  (void)y3;  /* the third node only bounds the slopes that are never needed */
  for (int j = 0; j < n; j++) {
    const float dy1 = y2[j] - y1[j];
    const float m2 = 0.0f;
    const float m1 = 2.0f * dy1 / dx1 - m2;
    const float t1 = (0.0f - x1) / dx1;
    const float t2 = t1 * t1, t3 = t2 * t1;
    const float p = 3.0f * t2 - 2.0f * t3;
    out[j] = (1.0f - p) * y1[j] + p * y2[j] + dx1 * ((t3 - 2.0f * t2 + t1) * m1 + (t3 - t2) * m2);
  }
}
