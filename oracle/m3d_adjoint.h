/* m3d_adjoint.h -- TEST INFRASTRUCTURE ONLY: the reverse (adjoint) of the oracle's
 * short_char2 restatement (SRC/rt_solvers_mod.f90:224-507), FP64. */
#ifndef M3D_ADJOINT_H
#define M3D_ADJOINT_H
#include "m3d_oracle.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Cotangents of short_char2's S step (no eta flags, no split rates) in FP64 with exact
 * weights.  chi, S, I: work arrays (nz+2, ny, nx, nf) with the ghost planes, I holding
 * the forward solution of orc_short_char2_f64 (and the upstream ghost plane);
 * chi500, chi500_up (nz, ny, nx), g500 (ny, nx) as the forward.  Ibar: on entry the
 * direct cotangent of every plane's I (planes 1..nz), on return the total one, and
 * the upstream ghost plane's cotangent added.  chibar, Sbar: accumulated (+=) on
 * planes 1..nz and the upstream ghost plane.  chi500 is held fixed. */
void orc_short_char2_adjoint_f64(const orc_sc_params *p, const double *chi, const double *S, const double *I,
                                 const double *chi500, const double *chi500_up, const double *g500,
                                 double *Ibar, double *chibar, double *Sbar);

/* the exact linear-step weights and their derivatives d/dt */
void orc_lin_weights_deriv_f64(double t, double *E, double *W, double *Psi, double *dE, double *dW,
                               double *dPsi);

#ifdef __cplusplus
}
#endif
#endif
