/* m3d_hermite.h -- TEST INFRASTRUCTURE ONLY: oracle restatement of
 * short_char_hermite (SRC/rt_solvers_mod.f90:1917-2248). */
#ifndef M3D_HERMITE_H
#define M3D_HERMITE_H
#include "m3d_oracle.h"
#ifdef __cplusplus
extern "C" {
#endif

/* chi, S, I: work arrays (nz+2, ny, nx, nf) with the ghost planes, f fastest;
 * beta (nz, ny, nx, nf); pe4, temp, dpe_d[xyz], dlnt_d[xyz] (nz, ny, nx); hknu (nf);
 * Psi, Irem, escape (nz, ny, nx, nf) or NULL. */
void orc_short_char_hermite_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                                const float *beta, const float *pe4, const float *temp, const float *dpe_dx,
                                const float *dpe_dy, const float *dpe_dz, const float *dlnt_dx,
                                const float *dlnt_dy, const float *dlnt_dz, const float *hknu, float *Psi,
                                float *Irem, float *escape);

#ifdef __cplusplus
}
#endif
#endif
