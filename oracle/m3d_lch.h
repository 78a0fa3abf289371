/* m3d_lch.h -- TEST INFRASTRUCTURE ONLY: oracle restatement of long_char_hermite
 * (SRC/rt_solvers_mod.f90:1052-1504), the cubic-Bezier column solver. */
#ifndef M3D_LCH_H
#define M3D_LCH_H
#include "m3d_oracle.h"
#ifdef __cplusplus
extern "C" {
#endif

/* the boundary below the first plane (SRC/rt_solvers_mod.f90:1100-1130) */
#define ORC_LCH_TOP 0      /* zeroed top ghost: herm_first */
#define ORC_LCH_BOTTOM 1   /* LTE ghost below the domain, the model's Pe/T there */
#define ORC_LCH_UPSTREAM 2 /* a neighbour's last plane (comp_up) */

/* bezier3_weights (SRC/rt_solvers_mod.f90:1301-1323): E = e^-t and the weights of
 * S_up, the two control points and S_target */
void orc_bezier3_weights(float t, float *E, float *w0, float *w1, float *w2, float *w3);
void orc_bezier3_weights_n(int n, const float *t, float *E, float *w0, float *w1, float *w2, float *w3);

/* chi, S, I: work arrays (nz+2, ny, nx, nf) with the ghost planes, f fastest;
 * beta (nz, ny, nx, nf); pe4, temp, dpe_dz, dlnt_dz (nz, ny, nx), gradients per cm
 * along +z; hknu (nf); ghost thermodynamics gpe, gdpe, gtemp, gdlnt (ny, nx) for
 * ORC_LCH_BOTTOM / ORC_LCH_UPSTREAM (unused for ORC_LCH_TOP); Psi (nz, ny, nx, nf)
 * or NULL.  Needs nz >= 2. */
void orc_long_char_hermite_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                               const float *beta, const float *pe4, const float *temp, const float *dpe_dz,
                               const float *dlnt_dz, const float *hknu, int ghost_kind, const float *gpe,
                               const float *gdpe, const float *gtemp, const float *gdlnt, float *Psi);

/* long_char_linear (SRC/rt_solvers_mod.f90:666-757): work arrays as above; Psi, Irem,
 * escape (nz, ny, nx, nf) or NULL (Irem, escape filled with p->split_rates). */
void orc_long_char_linear_f32(const orc_sc_params *p, const float *chi, const float *S, float *I, float *Psi,
                              float *Irem, float *escape);

/* long_char_hermite2 (SRC/rt_solvers_mod.f90:1536-1895): as orc_long_char_hermite_f32
 * without beta, T and their gradients (its source treatment does not use them);
 * gpe, gdpe (ny, nx) the ghost's Pe for ORC_LCH_BOTTOM / ORC_LCH_UPSTREAM. */
void orc_long_char_hermite2_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                                const float *pe4, const float *dpe_dz, int ghost_kind, const float *gpe,
                                const float *gdpe, float *Psi);

#ifdef __cplusplus
}
#endif
#endif
