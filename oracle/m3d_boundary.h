/* TEST INFRASTRUCTURE ONLY: the boundary-plane and LTE background restatements
 * (m3d_boundary.c). */
#ifndef M3D_BOUNDARY_H
#define M3D_BOUNDARY_H
#include <stddef.h>
/* h/k of HOP/const_mod.f90:39,43,71 (hh / bk) [K s] */
#define ORC_H_K (6.626176e-27 / 1.380662e-16)
void orc_lte_background(int nf, int ncell, const double *nu, const double *temp, float *beta, float *Bnu);
void orc_lte_ghost(int nf, int nxy, const double *nu, const float *chi1, const float *g500, const float *chi500_1,
                   const float *T_ghost, float *gchi, float *gS, float *gI);
void orc_pchip_ghost(int n, double ds3, const float *y1, const float *y2, const float *y3, float *out);
#endif
