/* m3d_qlg.h -- TEST INFRASTRUCTURE ONLY: oracle restatement of QLG-bot and
 * add_opacity_equi (HOP/qlg_mod.f90, HOP/linelist_mod.f90 of the reference). */
#ifndef M3D_QLG_H
#define M3D_QLG_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float ih, bL, gG, cZH, cL2; } orc_qlg_coef;     /* qlg_bot_coef_t */
typedef struct { float D, D2, SQ, D2h, D2G, D2L; } orc_qlg_cell; /* qlg_bot_cell_t */

void orc_qlg_init(void);
void orc_qlg_tables(float *tgx, float *th2x);     /* 9020 entries each */
void orc_qlg_w32(double *a, double *L);           /* Weideman N=32 coefficients */
void orc_qlg_ref_f64(double a, double v0, double dv, int n, double *K);
void orc_qlg_bot_sweep(float a, float v0, float dv, int n, float *K);
void orc_qlg_bot_setup(float gL, float D, float SQ, float S, orc_qlg_coef *c);
float orc_qlg_bot_line_eval(const orc_qlg_coef *c, float x);
void orc_qlg_bot_cell(float D, orc_qlg_cell *w);
void orc_qlg_bot_line(const orc_qlg_cell *w, float gL, float S, float ref, float lnq,
                      orc_qlg_coef *c, float *hw);
void orc_qlg_bot_line_sweep(const orc_qlg_cell *w, float gL, float S, float ref, float lnq, int nf,
                            const double *llam, double llc, double pos, double idl, float cap,
                            int *fa, int *fb, float *hw, float *K);
float orc_qlg_bot_profile(float a, float q, float vpr, float idf);

/* one isotope's structure-of-arrays line list (HOP/line_types_mod.f90:22-74) */
typedef struct {
  int nline;
  const double *llam0;
  const float *lower_ev, *gpref, *grad, *gw1, *gw2, *gstk;
  const int32_t *unsold, *lev_i, *lev_j, *ipart;  /* ipart NULL: every line in group 1 */
  int32_t *used;
  float *strength, *reach;
  /* the list's reaches (has_reach: every line carries one, finalize_linelist,
   * HOP/linelist_mod.f90:375-378) and active flags; NULL: none / every line */
  const float *reach_in;
  const int32_t *active;
} orc_linelist;

/* select_lines (HOP/linelist_mod.f90:1110-1142): the active lines of l0..l1
 * (1-based, narrowed in place) whose reach window can touch llam(1:nf), into sel
 * (1-based, nsel of them); hb the half grid steps at the two ends.  vmax is
 * max |vproj| over the call's cells. */
void orc_select_lines(const orc_linelist *ll, int nf, const double *llam, double vmax, int *l0, int *l1,
                      int32_t *sel, int *nsel, double hb[2]);

void orc_add_opacity_equi(const orc_linelist *ll, int nf, const double *llam, int ncell,
                          float *chi, float *eta, const float *cont, const float *logref,
                          const float *vproj, const float *inv_df, const float *vbar2,
                          const float *e_kT, const float *ne, const float *tothi,
                          const float *stark_pe, int np, const float *lnbase, int nlev,
                          const float *dep, int do_nlte, const float *one_m_eps, float *chi_scat,
                          int do_scat, float tol, float line_reject_frac);
/* add_opacity (HOP/linelist_mod.f90:469-719) on an ascending grid of any spacing;
 * arguments as orc_add_opacity_equi */
void orc_add_opacity(const orc_linelist *ll, int nf, const double *llam, int ncell, float *chi, float *eta,
                     const float *cont, const float *logref, const float *vproj, const float *inv_df,
                     const float *vbar2, const float *e_kT, const float *ne, const float *tothi,
                     const float *stark_pe, int np, const float *lnbase, int nlev, const float *dep,
                     int do_nlte, const float *one_m_eps, float *chi_scat, int do_scat, float tol,
                     float line_reject_frac);

#ifdef __cplusplus
}
#endif
#endif
