/* m3d_rates.h -- TEST INFRASTRUCTURE ONLY: oracle restatement of the line-rate
 * integrals of bound_bound_rates (SRC/rt_patch_mod.f90:2852-3142). */
#ifndef M3D_RATES_H
#define M3D_RATES_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nline, nf, ncell, nk;
  int precond, own_norm, largest, kahan;
  const int32_t *iblue, *ired, *lev_i, *lev_j;  /* (nline), 1-based */
  const double *Aij, *Bij;                      /* (nline) */
  const float *lower_off;                       /* (nline) line E_low - level E_i [eV] */
  const float *q;                               /* (nline, nf) profile abscissa */
  const double *wnu, *wq;                       /* (nline, nf) quadrature weights */
  float wmu, wtot;
  const double *g;                              /* (nk) */
  const double *b, *nstar;                      /* (nk, ncell) */
  const double *pop_ratio;                      /* (ncell) */
  const float *inv_df, *vproj, *e_kT;           /* (ncell) */
  const float *adamp;                           /* (nline, ncell) */
  const float *Inu, *Psi, *Bnu, *beta;          /* (f, cell) at f*sf + cell*sc */
  size_t sf, sc;
} orc_bb_args;

void orc_bb_rates(const orc_bb_args *a, float *gamma_ij, float *gamma_ji);

#ifdef __cplusplus
}
#endif
#endif
