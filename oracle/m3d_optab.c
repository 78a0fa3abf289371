/*
 * m3d_optab.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the sampled
 * background-opacity table of HOP (experiments/OpenMULTI3D/HOP/optab_mod.f90 of
 * Bergemann-Group-development/dispatch), the checker for dispatch_amd/csrc/optab.hip.
 *
 *   sample_build_derivatives   optab_mod.f90:127-166: node slopes along wavelength
 *                              (pchip_slopes8 on the FP64 nodes and values) and along
 *                              the two PC axes (node_slopes_dim, :438-468: pchip_mono4
 *                              evaluated at its own nodes, i.e. pchip_slopes4's slopes)
 *   sample_interpolate         optab_mod.f90:181-287: separable tensor-product cubic
 *                              Hermite, wavelength -> PC axis 1 -> PC axis 2
 *   pchip_slopes4 / 8          HOP/interpol_mod.f90:193-283
 *   hermite_coefs              HOP/interpol_mod.f90:171-188
 *   searchsorted8              HOP/op_math_mod.f90:122-199
 *
 * In the reference's precision: pchip_slopes4 forms dx, dy and the quotients
 * dy/dx in FP32 and the harmonic mean in FP64; pchip_slopes8 all in FP64; the
 * per-frequency bracket and the per-cell PCA coordinates in FP64, the Hermite
 * coefficients and the collapse in FP32, each product and sum rounded apart in the
 * reference's order (the Makefile's -ffp-contract=off).
 *
 * Pinning: slopes, totals and scattering parts bit for bit against the reference's
 * compiled optab_mod (tests/golden/hop_optab_sample.npz); node values and linear
 * data reproduced exactly (tests/test_oracle_optab.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "m3d_optab.h"

#define V(a, k, i, j) a[(size_t)(k) + (size_t)t->nnu * ((size_t)(i) + (size_t)t->ntemp * (size_t)(j))]

/* pchip_slopes8 with imin = 1, imax = n: a line of n values at stride s */
static void slopes8(int n, const double *x, const float *f, size_t s, float *out)
{
  double *y = (double *)malloc(sizeof(double) * n), *d = (double *)calloc(n, sizeof(double));
  double *dx = (double *)malloc(sizeof(double) * n), *dy = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) y[i] = (double)f[i * s];
  for (int i = 0; i + 1 < n; i++) {
    dx[i] = x[i + 1] - x[i];
    dy[i] = y[i + 1] - y[i];
  }
  if (n == 2) {
    d[0] = dy[0] / dx[0];
    d[1] = d[0];
  }
  for (int i = 1; i < n - 1; i++) {
    if (dy[i - 1] == 0 || dy[i] == 0) continue;
    const double d1 = dy[i - 1] / dx[i - 1], d2 = dy[i] / dx[i];
    if ((d1 < 0) == (d2 < 0)) {
      const double alpha = (1.0 + dx[i - 1] / (dx[i] + dx[i - 1])) / 3.0;
      d[i] = 1.0 / ((1.0 - alpha) / d1 + alpha / d2);
    }
  }
  if (n >= 2) {
    d[0] = 2.0 * dy[0] / dx[0] - d[1];
    d[n - 1] = 2.0 * dy[n - 2] / dx[n - 2] - d[n - 2];
  }
  for (int i = 0; i < n; i++) out[i * s] = (float)d[i];
  free(y); free(d); free(dx); free(dy);
}

/* pchip_slopes4 with imin = 1, imax = n (node_slopes_dim's values at the nodes) */
static void slopes4(int n, const float *x, const float *y, size_t s, float *out)
{
  float *d = (float *)calloc(n, sizeof(float));
  float *dx = (float *)malloc(sizeof(float) * n), *dy = (float *)malloc(sizeof(float) * n);
  for (int i = 0; i + 1 < n; i++) {
    dx[i] = x[i + 1] - x[i];
    dy[i] = y[(i + 1) * s] - y[i * s];
  }
  if (n == 2) {
    d[0] = dy[0] / dx[0];
    d[1] = d[0];
  }
  for (int i = 1; i < n - 1; i++) {
    if (dy[i - 1] == 0 || dy[i] == 0) continue;
    const double d1 = (double)(dy[i - 1] / dx[i - 1]), d2 = (double)(dy[i] / dx[i]);
    if ((d1 < 0) == (d2 < 0)) {
      const double alpha = (double)((1.0f + dx[i - 1] / (dx[i] + dx[i - 1])) / 3.0f);
      d[i] = (float)(1.0 / ((1.0 - alpha) / d1 + alpha / d2));
    }
  }
  if (n >= 2) {
    d[0] = 2.0f * dy[0] / dx[0] - d[1];
    d[n - 1] = 2.0f * dy[n - 2] / dx[n - 2] - d[n - 2];
  }
  for (int i = 0; i < n; i++) out[i * s] = d[i];
  free(d); free(dx); free(dy);
}

void orc_optab_slopes(const orc_optab *t, const float *f, float *dl, float *dt, float *de)
{
  float *tax = (float *)malloc(sizeof(float) * t->ntemp), *eax = (float *)malloc(sizeof(float) * t->nne);
  for (int i = 0; i < t->ntemp; i++) tax[i] = (float)(t->c0[0] + i * t->dc[0]);
  for (int i = 0; i < t->nne; i++) eax[i] = (float)(t->c0[1] + i * t->dc[1]);
  for (int j = 0; j < t->nne; j++)
    for (int i = 0; i < t->ntemp; i++) slopes8(t->nnu, t->lam, &V(f, 0, i, j), 1, &V(dl, 0, i, j));
  for (int j = 0; j < t->nne; j++)
    for (int k = 0; k < t->nnu; k++) slopes4(t->ntemp, tax, &V(f, k, 0, j), (size_t)t->nnu, &V(dt, k, 0, j));
  for (int i = 0; i < t->ntemp; i++)
    for (int k = 0; k < t->nnu; k++)
      slopes4(t->nne, eax, &V(f, k, i, 0), (size_t)t->nnu * t->ntemp, &V(de, k, i, 0));
  free(tax); free(eax);
}

/* hermite_coefs, HOP/interpol_mod.f90:171-188 */
static void hcoef(float t, float h, float *p, float *c0, float *c1, float *cm0, float *cm1)
{
  const float t2 = t * t, t3 = t2 * t;
  *p = 3.0f * t2 - 2.0f * t3;
  *c0 = 1.0f - *p;
  *c1 = *p;
  *cm0 = h * (t3 - 2.0f * t2 + t);
  *cm1 = h * (t3 - t2);
}

/* searchsorted8 with outbounds_ok: 0 below, n above, else the l of BinarySearch8 */
static int ssorted(const double *a, int n, double x)
{
  if (x < a[0]) return 0;
  if (x > a[n - 1]) return n;
  int l = 1, r = n;
  while (r - l > 1) {
    const int mid = l + (r - l) / 2;
    if (a[mid - 1] >= x) r = mid; else l = mid;
  }
  return l;
}

typedef struct { float pl, wc0, wc1, wcm0, wcm1; int k; } lamc;
typedef struct { float pt, tc0, tc1, tcm0, tcm1, nc0, nc1, ncm0, ncm1; int it, je; } cellc;

static float collapse(const orc_optab *t, const float *val, const float *dl, const float *dt, const float *de,
                      const lamc *L, const cellc *c)
{
  const int k = L->k - 1, k1 = k + 1, it = c->it - 1, je = c->je - 1;
  const float v00 = L->wc0 * V(val, k, it, je) + L->wc1 * V(val, k1, it, je) + L->wcm0 * V(dl, k, it, je) +
                    L->wcm1 * V(dl, k1, it, je);
  const float v10 = L->wc0 * V(val, k, it + 1, je) + L->wc1 * V(val, k1, it + 1, je) + L->wcm0 * V(dl, k, it + 1, je) +
                    L->wcm1 * V(dl, k1, it + 1, je);
  const float v01 = L->wc0 * V(val, k, it, je + 1) + L->wc1 * V(val, k1, it, je + 1) + L->wcm0 * V(dl, k, it, je + 1) +
                    L->wcm1 * V(dl, k1, it, je + 1);
  const float v11 = L->wc0 * V(val, k, it + 1, je + 1) + L->wc1 * V(val, k1, it + 1, je + 1) +
                    L->wcm0 * V(dl, k, it + 1, je + 1) + L->wcm1 * V(dl, k1, it + 1, je + 1);
  const float q = 1.0f - L->pl, p = L->pl;
  const float gt00 = q * V(dt, k, it, je) + p * V(dt, k1, it, je);
  const float gt10 = q * V(dt, k, it + 1, je) + p * V(dt, k1, it + 1, je);
  const float gt01 = q * V(dt, k, it, je + 1) + p * V(dt, k1, it, je + 1);
  const float gt11 = q * V(dt, k, it + 1, je + 1) + p * V(dt, k1, it + 1, je + 1);
  const float ge00 = q * V(de, k, it, je) + p * V(de, k1, it, je);
  const float ge10 = q * V(de, k, it + 1, je) + p * V(de, k1, it + 1, je);
  const float ge01 = q * V(de, k, it, je + 1) + p * V(de, k1, it, je + 1);
  const float ge11 = q * V(de, k, it + 1, je + 1) + p * V(de, k1, it + 1, je + 1);
  const float f0 = c->tc0 * v00 + c->tc1 * v10 + c->tcm0 * gt00 + c->tcm1 * gt10;
  const float f1 = c->tc0 * v01 + c->tc1 * v11 + c->tcm0 * gt01 + c->tcm1 * gt11;
  const float ge0 = (1.0f - c->pt) * ge00 + c->pt * ge10;
  const float ge1 = (1.0f - c->pt) * ge01 + c->pt * ge11;
  return c->nc0 * f0 + c->nc1 * f1 + c->ncm0 * ge0 + c->ncm1 * ge1;
}

void orc_optab_interpolate(const orc_optab *t, const float *dl, const float *dt, const float *de, const float *sl,
                           const float *st, const float *se, int ncell, const double *temp, const double *ne,
                           const float *chi500, int nf, const double *lam_q, float *back, float *scat)
{
  const double rd1 = 1.0 / t->dc[0], rd2 = 1.0 / t->dc[1];
  const float dc1 = (float)t->dc[0], dc2 = (float)t->dc[1];
  lamc *L = (lamc *)malloc(sizeof(lamc) * nf);
  for (int f = 0; f < nf; f++) {
    int k = ssorted(t->lam, t->nnu, lam_q[f]);
    k = k < 1 ? 1 : k;
    k = k > t->nnu - 1 ? t->nnu - 1 : k;
    L[f].k = k;
    const float hl = (float)(t->lam[k] - t->lam[k - 1]);
    const float tl = (float)((lam_q[f] - t->lam[k - 1]) / (t->lam[k] - t->lam[k - 1]));
    hcoef(tl, hl, &L[f].pl, &L[f].wc0, &L[f].wc1, &L[f].wcm0, &L[f].wcm1);
  }
  for (int c = 0; c < ncell; c++) {
    cellc C;
    const double d1 = log(temp[c]) - t->mu[0], d2 = log(ne[c]) - t->mu[1];
    const double p1 = t->rot[0] * d1 + t->rot[1] * d2;   /* rot(1,1) d1 + rot(2,1) d2 */
    const double p2 = t->rot[2] * d1 + t->rot[3] * d2;   /* rot(1,2) d1 + rot(2,2) d2 */
    const double ft = (p1 - t->c0[0]) * rd1 + 1.0;
    int it = (int)floor(ft);
    it = it < 1 ? 1 : (it > t->ntemp - 1 ? t->ntemp - 1 : it);
    hcoef((float)(ft - it), dc1, &C.pt, &C.tc0, &C.tc1, &C.tcm0, &C.tcm1);
    const double fe = (p2 - t->c0[1]) * rd2 + 1.0;
    int je = (int)floor(fe);
    je = je < 1 ? 1 : (je > t->nne - 1 ? t->nne - 1 : je);
    float pe;
    hcoef((float)(fe - je), dc2, &pe, &C.nc0, &C.nc1, &C.ncm0, &C.ncm1);
    C.it = it;
    C.je = je;
    for (int f = 0; f < nf; f++) {
      const float kap = collapse(t, t->val, dl, dt, de, &L[f], &C);
      const float ks = collapse(t, t->scat, sl, st, se, &L[f], &C);
      back[(size_t)c * nf + f] = (kap + ks) * chi500[c];
      scat[(size_t)c * nf + f] = ks * chi500[c];
    }
  }
  free(L);
}
