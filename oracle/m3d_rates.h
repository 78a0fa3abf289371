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
  /* 'largest' with preconditioning (the tracked split contributions) */
  int split_offset, use_flags, do_scat, npw;
  const float *Irem, *escape, *chi;             /* (f, cell): the solver's split parts, total chi */
  const float *split_opac1, *split_opac2, *split_chi1, *split_chi2;  /* (f, cell) */
  const int32_t *split_id, *split_chi_id;       /* (f, cell) */
  const float *chi_scat, *dJnu;                 /* (f, cell), do_scat */
  const int32_t *ntrans;                        /* (nline) transition numbers (psi_flags bits) */
  const int32_t *psi_flags_in;                  /* (npw, ncell) words, cells fastest */
  int32_t *psi_flags_out;                       /* (npw, ncell) raised bits */
} orc_bb_args;

void orc_bb_rates(const orc_bb_args *a, float *gamma_ij, float *gamma_ji);

/* add_bound_bound_opacities (SRC/rt_patch_mod.f90:2689-2831) with add_largest
 * (3150-3169) and the direction's wphi share */
typedef struct {
  int nline, nf, ncell, nk;
  int lte, split, split_offset;
  const int32_t *iblue, *ired, *lev_i, *lev_j;
  const double *Bij;
  const float *lower_off, *q;                   /* (nline), (nline, nf) */
  const double *wq;                             /* (nline, nf) */
  const int32_t *wphi_line;                     /* (nline) 1 = sum wphi for the line; NULL = none */
  const double *b, *nstar, *pop_ratio;          /* (nk, ncell) x2, (ncell) */
  const float *inv_df, *vproj, *e_kT, *adamp;   /* (ncell) x3, (nline, ncell) */
  float wmu;
  size_t sf, sc;
  float *xhat, *ehat;                           /* (f, cell) */
  float *split_opac1, *split_opac2, *split_chi1, *split_chi2;
  int32_t *split_id, *split_chi_id;
  float *wphi;                                  /* (nline, ncell), += wmu * sum phi wq */
} orc_bbop_args;

void orc_bb_opacity(const orc_bbop_args *a);

/* bound_free_rates (SRC/rt_patch_mod.f90:3193-3259) */
typedef struct {
  int ncont, nline, nf, ncell, nk, precond;
  double cut_x;
  const int32_t *iblue, *ired, *lev_i, *lev_j, *active;  /* (ncont), 1-based */
  const double *ev;                             /* (nk) */
  const double *wnu;                            /* (ncont, nf) */
  const float *alpha;                           /* (ncont, nf) */
  const float *twohnu3_c2;                      /* (nf) */
  const double *e_kT;                           /* (ncell) */
  const double *nstar, *b;                      /* (nk, ncell) */
  const float *Jnu, *beta, *PsiJ;               /* (f, cell) */
  size_t sf, sc;
} orc_bf_args;

/* gamma arrays (nline + ncont, ncell); the continua at rows nline.. */
void orc_bf_rates(const orc_bf_args *a, float *gamma_ij, float *gamma_ji);

/* work_to_transfer_units (SRC/rt_patch_mod.f90:3571-3651) over n (f, cell) points of
 * any common layout: work1 = xhat -> chi, work2 = ehat -> S.  split: eta from the
 * tracked parts; scat: chi_scat, dJnu; returns the number of floored points. */
long orc_work_to_transfer(size_t n, float *work1, float *work2, const float *beta, const float *so1,
                          const float *so2, const float *chi_scat, const float *dJnu, int split,
                          int scat, int suppress_masers, double maser_chi_floor);

/* lte_background_emission (SRC/rt_patch_mod.f90:2127-2328), elementwise over n points:
 * with scat_ali the continuum-scattering ALI step J = (J_FS - op J_ali_old)/(1 - op),
 * op = clamp(sigma PsiJ, 0, scat_op_max), in FP64 (negative -> J_FS), J_ali = J; with
 * j_to_b J = B first; then J -= B, eta = chi_cont B + wscat sigma J, dJnu = wscat J.
 * Returns the count of negative ALI values; diag[3] = op_raw max, op max, amp max. */
long orc_background_emission(size_t n, float *J, const float *B, const float *chi_cont, const float *sigma,
                             const float *PsiJ, const float *Jali_old, float *Jali, float *eta, float *dJnu,
                             int scat_ali, int j_to_b, float wscat, double scat_op_max, double *diag);

/* convert_to_work_units (:3518-3536): xhat = chi_cont + ehat, then ehat /= beta */
void orc_convert_to_work_units(size_t n, const float *chi_cont, float *xhat, float *ehat, const float *beta);

#ifdef __cplusplus
}
#endif
#endif
