/* TEST INFRASTRUCTURE ONLY: the sampled background-table restatement (m3d_optab.c). */
#ifndef M3D_OPTAB_H
#define M3D_OPTAB_H
typedef struct {
  int nnu, ntemp, nne;
  double mu[2], rot[4], c0[2], dc[2];  /* rot column-major, as the reference's rot(2,2) */
  const double *lam;                   /* (nnu) */
  const float *val, *scat;             /* (nnu, ntemp, nne), nnu fastest */
} orc_optab;
/* node slopes along lam, the first and the second PC axis, each shaped like val */
void orc_optab_slopes(const orc_optab *t, const float *f, float *dl, float *dt, float *de);
void orc_optab_interpolate(const orc_optab *t, const float *dl, const float *dt, const float *de, const float *sl,
                           const float *st, const float *se, int ncell, const double *temp, const double *ne,
                           const float *chi500, int nf, const double *lam_q, float *back, float *scat);
#endif
