/*
 * m3d_nlte.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the per-cell
 * statistical-equilibrium update, the checker for the HIP rate-solve kernel.
 * Citations are relative to experiments/OpenMULTI3D/ of
 * Bergemann-Group-development/dispatch.
 *
 *   normalize_profiles        SRC/atom_mod.f90:377-431
 *   setup_transition_matrix   SRC/atom_mod.f90:740-851
 *   line_rate_coefficients    SRC/atom_mod.f90:862-884
 *   lineq_no_pivot            HOP/op_math_mod.f90:420-449
 *   update_populations        SRC/atom_mod.f90:439-732
 *
 * Pinning: lineq_no_pivot against the reference's compiled routine
 * (tests/golden/hop_lineq.npz); the assembly by detailed balance (with J = B_nu
 * the radiative pairs satisfy n*_i R_ij = n*_j R_ji and LTE is a fixed point of the
 * update), by particle conservation and by numpy's solve of the same matrix; the
 * limiter and floor by their closed forms.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "m3d_nlte.h"

/* HOP/op_math_mod.f90:420-449, column-major a(n,n) */
void orc_lineq_no_pivot(double *a, double *b, int n)
{
#define A(i, j) a[(i) + (size_t)n * (j)]
  for (int k = 0; k < n - 1; k++) {
    for (int i = k + 1; i < n; i++) A(i, k) = A(i, k) / A(k, k);
    for (int j = k + 1; j < n; j++)
      if (A(k, j) != 0)
        for (int i = k + 1; i < n; i++) A(i, j) = A(i, j) - A(i, k) * A(k, j);
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) b[i] = b[i] - A(i, j) * b[j];
  for (int i = n - 1; i >= 0; i--) {
    for (int j = i + 1; j < n; j++) b[i] = b[i] - A(i, j) * b[j];
    b[i] = b[i] / A(i, i);
  }
#undef A
}

/* normalize_profiles (after iteration 1): lines times wphi (already inverted),
 * continua gamma_ji times n*_i/n*_j in FP64, stored back in FP32 */
void orc_normalize_rates(const orc_atom *at, int ncell, const double *nstar, const float *wphi,
                         float *gamma_ij, float *gamma_ji)
{
  for (int kr = 0; kr < at->nrad; kr++) {
    if (!at->active[kr]) continue;
    for (int c = 0; c < ncell; c++) {
      const size_t o = (size_t)kr * ncell + c;
      if (kr < at->nline) {
        if (wphi) {
          gamma_ij[o] = gamma_ij[o] * wphi[o];
          gamma_ji[o] = gamma_ji[o] * wphi[o];
        }
      } else {
        const int i = at->ti[kr] - 1, j = at->tj[kr] - 1;
        gamma_ji[o] = (float)((double)gamma_ji[o] *
                              (nstar[(size_t)i * ncell + c] / nstar[(size_t)j * ncell + c]));
      }
    }
  }
}

/* setup_transition_matrix of one cell: w (nk,nk) column-major, e (nk) */
static void transition_matrix(const orc_atom *at, const orc_se_params *p, int ncell, int c,
                              const double *e_kT, const double *totn, const double *nstar,
                              const double *b, const float *gij, const float *gji, const double *cup,
                              double *w, double *e, double *bn)
{
  const int nk = at->nk;
#define W(i, j) w[(i) + (size_t)nk * (j)]
#define NS(i) nstar[(size_t)(i) * ncell + c]
  memset(w, 0, sizeof(double) * nk * nk);
  memset(e, 0, sizeof(double) * nk);
  for (int kr = 0; kr < at->nrad; kr++) {
    if (!at->active[kr]) continue;
    const int i = at->ti[kr] - 1, j = at->tj[kr] - 1;
    double r_up, r_down;
    if ((at->ev[j] - at->ev[i]) * e_kT[c] > p->cut_x) {
      const double cut = kr < at->nline ? at->Aji[kr] : p->bf_cut_rate;
      W(i, j) = W(i, j) + cut;
      W(j, i) = W(j, i) + cut * NS(j) / NS(i);
      continue;
    }
    if (kr < at->nline) {
      /* line_rate_coefficients, SRC/atom_mod.f90:862-884 */
      const double th = gij[(size_t)kr * ncell + c], dev = gji[(size_t)kr * ncell + c];
      const double g_ij = at->g[i] / at->g[j];
      if (NS(i) > 0 && NS(j) > 0) {
        const double beta = NS(j) * g_ij / NS(i);
        r_up = th * (beta / (1.0 - beta)) + dev;
        r_down = g_ij * (th / (1.0 - beta) + dev);
      } else {
        r_up = 0.0;
        r_down = 0.0;
      }
    } else {
      r_up = gij[(size_t)kr * ncell + c];
      r_down = gji[(size_t)kr * ncell + c];
    }
    W(j, i) = W(j, i) + r_up;
    W(i, j) = W(i, j) + r_down;
  }
  if (p->cswitch > 0 && cup) {
    for (int q = 0; q < at->npair; q++) {
      const int i = at->pi[q] - 1, j = at->pj[q] - 1;
      const double cu = cup[(size_t)q * ncell + c];
      if (cu == 0) continue;
      W(j, i) = W(j, i) + cu * p->cswitch;
      W(i, j) = W(i, j) + cu * (NS(i) / NS(j)) * p->cswitch;
    }
  }
  for (int i = 0; i < nk; i++) {
    double s = 0.0;
    for (int k = 0; k < nk; k++) s += W(k, i);
    W(i, i) = -s;
  }
  int pos = 0;
  for (int i = 0; i < nk; i++) {
    bn[i] = b[(size_t)i * ncell + c] * NS(i);
    if (bn[i] > bn[pos]) pos = i;
  }
  e[pos] = totn[c];
  for (int i = 0; i < nk; i++)
    for (int j = 0; j < nk; j++) W(i, j) = W(i, j) * bn[j];
  for (int j = 0; j < nk; j++) W(pos, j) = bn[j];
#undef W
#undef NS
}

/* The assembled matrix and right-hand side of cell c, for tests. */
void orc_transition_matrix(const orc_atom *at, const orc_se_params *p, int ncell, int c,
                           const double *e_kT, const double *totn, const double *nstar,
                           const double *b, const float *gij, const float *gji, const double *cup,
                           double *w, double *e)
{
  double *bn = (double *)malloc(sizeof(double) * at->nk);
  transition_matrix(at, p, ncell, c, e_kT, totn, nstar, b, gij, gji, cup, w, e, bn);
  free(bn);
}

/* update_populations, SRC/atom_mod.f90:439-732, over all cells in order.  b is
 * updated in place; returns 0, or 1 + the first cell without a finite solution. */
int orc_update_populations(const orc_atom *at, const orc_se_params *p, int ncell, const double *e_kT,
                           const double *temp, const double *totn, const double *nstar, double *b,
                           const float *gij, const float *gji, const double *cup, orc_se_result *res)
{
  const int nk = at->nk;
  double *w = (double *)malloc(sizeof(double) * nk * nk);
  double *e = (double *)malloc(sizeof(double) * nk);
  double *bn = (double *)malloc(sizeof(double) * nk);
  double *bold = (double *)malloc(sizeof(double) * nk);
  const double eup = p->max_update > 1.0 ? p->max_update : 1.0, elo = 1.0 / eup;
  int status = 0;
  res->maxerr = 0.0;
  res->epos = -1;
  res->elvl = -1;
  res->nclamp = 0;
  for (int i = 0; i < nk; i++) res->lvlerr[i] = 0.0;
  for (int c = 0; c < ncell; c++) {
    if (!(totn[c] > 0 && temp[c] < p->tmax_update)) continue;
    transition_matrix(at, p, ncell, c, e_kT, totn, nstar, b, gij, gji, cup, w, e, bn);
    orc_lineq_no_pivot(w, e, nk);
    for (int i = 0; i < nk; i++)
      if (!(fabs(e[i]) <= 1.79769313486231570e308)) {
        status = c + 1;
        goto done;
      }
    for (int i = 0; i < nk; i++) {
      if (e[i] < elo || e[i] > eup) res->nclamp++;
      if (e[i] < elo)
        e[i] = elo * exp((e[i] > 0 ? e[i] : 0.0) / elo - 1.0);
      else if (e[i] > eup)
        e[i] = eup * (2.0 - eup / e[i]);
    }
    if (p->sor_w != 1.0)
      for (int i = 0; i < nk; i++) e[i] = pow(e[i], p->sor_w);
    double sum = 0.0;
    for (int i = 0; i < nk; i++) {
      const size_t o = (size_t)i * ncell + c;
      bold[i] = b[o];
      b[o] = b[o] * e[i];
      sum += b[o] * nstar[o];
    }
    const double norm = totn[c] / sum;
    for (int i = 0; i < nk; i++) {
      const size_t o = (size_t)i * ncell + c;
      b[o] = b[o] * norm;
      if (nstar[o] > 0) {
        double fl = p->min_pop_frac * totn[c] / nstar[o];
        if (fl > 1.0) fl = 1.0;
        if (b[o] < fl) b[o] = fl;
      }
    }
    int imax = 0;
    double wmax = -1.0;
    for (int i = 0; i < nk; i++) {
      const size_t o = (size_t)i * ncell + c;
      e[i] = bold[i] > 0 ? b[o] / bold[i] - 1.0 : 0.0;
      if (fabs(e[i]) > fabs(res->lvlerr[i])) res->lvlerr[i] = e[i];
      double ew = fabs(e[i]);
      if (p->conv_pop_floor > 0) {
        double f = b[o] * nstar[o] / (p->conv_pop_floor * totn[c]);
        if (f > 1.0) f = 1.0;
        ew = fabs(e[i]) * f;
      }
      if (ew > wmax) {
        wmax = ew;
        imax = i;
      }
    }
    if (wmax > fabs(res->maxerr)) {
      res->maxerr = copysign(wmax, e[imax]);
      res->epos = c;
      res->elvl = imax;
    }
  }
done:
  free(w);
  free(e);
  free(bn);
  free(bold);
  return status;
}
