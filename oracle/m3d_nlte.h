/* m3d_nlte.h -- TEST INFRASTRUCTURE ONLY: oracle restatement of the per-cell
 * statistical-equilibrium update (SRC/atom_mod.f90, HOP/op_math_mod.f90). */
#ifndef M3D_NLTE_H
#define M3D_NLTE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* the model atom: nk levels, nrad transitions (lines first), npair collisional pairs */
typedef struct {
  int nk, nrad, nline, npair;
  const int32_t *ti, *tj;       /* (nrad) 1-based lower and upper level */
  const int32_t *active;        /* (nrad) */
  const double *g, *ev;         /* (nk) statistical weight, energy [eV] */
  const double *Aji;            /* (nline) */
  const int32_t *pi, *pj;       /* (npair) 1-based, pi < pj */
} orc_atom;

typedef struct {
  double max_update, min_pop_frac, conv_pop_floor, sor_w, cswitch, cut_x, bf_cut_rate, tmax_update;
} orc_se_params;

typedef struct {
  double maxerr;     /* signed largest weighted relative change */
  int epos, elvl;    /* its cell (0-based) and level (0-based) */
  long nclamp;       /* corrections cut by the limiter */
  double *lvlerr;    /* (nk) per-level signed largest change, caller-owned */
} orc_se_result;

void orc_lineq_no_pivot(double *a, double *b, int n);
void orc_normalize_rates(const orc_atom *at, int ncell, const double *nstar, const float *wphi,
                         float *gamma_ij, float *gamma_ji);
void orc_transition_matrix(const orc_atom *at, const orc_se_params *p, int ncell, int c,
                           const double *e_kT, const double *totn, const double *nstar,
                           const double *b, const float *gij, const float *gji, const double *cup,
                           double *w, double *e);
int orc_update_populations(const orc_atom *at, const orc_se_params *p, int ncell, const double *e_kT,
                           const double *temp, const double *totn, const double *nstar, double *b,
                           const float *gij, const float *gji, const double *cup, orc_se_result *res);

#ifdef __cplusplus
}
#endif
#endif
