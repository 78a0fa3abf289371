/*
 * m3d_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference
 * algorithms that the HIP kernels are checked against.  Nothing in the product
 * (dispatch_amd/) links or calls this file; only tests/, smoke() and bench.py's
 * cpu_baseline leg do.
 *
 * FP32 where the reference computes in real(4), FP64 where it uses real(8),
 * operation order as written there.  Built with -ffp-contract=off so that no
 * multiply-add is fused behind the reference's back.  Citations are relative to
 * experiments/OpenMULTI3D/ of Bergemann-Group-development/dispatch.
 *
 * Pinning (what each function is checked against, tests/test_oracle_*.py):
 *   exp_neg, linear_weights, parab_moments: FP64 closed forms to the ulp bounds
 *     the reference states (rt_solvers_mod.f90:772-774, 857-859);
 *   cc_kernel/set_stencil: Keys' piecewise kernel, partition of unity, linear
 *     reproduction;
 *   gauss quadrature: weight sums, unit vectors, the tabulated nodes;
 *   short_char2/long_char2/short_char/long_char_quadratic: constant-S exact
 *     answer, Eddington-Barbier for linear S(tau), horizontally homogeneous 3D
 *     box equal to the 1D column solution, brute-force FP64 integration, and the
 *     FP64 twin.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "m3d_oracle.h"

#define ORC_MAXN 4096

/* ---------------------------------------------------------------- FP32 pieces */

/* HOP/exp_neg.inc:20-37 -- E = exp(-t), u0 = 1 - exp(-t), neither from the other */
void orc_exp_neg(float t, float *E, float *u0)
{
  const float EXP_LOG2E = 1.44269504f, EXP_C1 = 0.693359375f, EXP_C2 = -2.12194440e-4f;
  const float Q0 = +1.000000000e+00f, Q1 = -5.000000000e-01f, Q2 = +1.666666716e-01f,
              Q3 = -4.167019576e-02f, Q4 = +8.353385143e-03f, Q5 = -1.351550105e-03f;
  float tc = t < 87.0f ? t : 87.0f;
  float n = roundf(tc * EXP_LOG2E);          /* Fortran anint: half away from zero */
  float r = (tc - n * EXP_C1) - n * EXP_C2;
  float m = r * (Q0 + r * (Q1 + r * (Q2 + r * (Q3 + r * (Q4 + r * Q5)))));
  int32_t bits = (int32_t)((uint32_t)(127 - (int)n) << 23);
  float s;
  memcpy(&s, &bits, 4);
  *E = s * (1.0f - m);
  *u0 = (1.0f - s) + s * m;
}

/* HOP/exp_neg.inc:43-49 */
void orc_exp_neg0(float t, float *E, float *u0)
{
  orc_exp_neg(t > -87.0f ? t : -87.0f, E, u0);
  if (!(t <= 87.0f)) *E = 0.0f;
}

/* SRC/rt_solvers_mod.f90:803-809 (degree-7 Psi = t*P(t)) */
float orc_lw_series(float t)
{
  const float P0 = +5.000000000e-01f, P1 = -1.666666716e-01f, P2 = +4.166666791e-02f,
              P3 = -8.333330974e-03f, P4 = +1.388740959e-03f, P5 = -1.983609400e-04f,
              P6 = +2.522676368e-05f, P7 = -2.834696488e-06f;
  return t * (P0 + t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * (P5 + t * (P6 + t * P7)))))));
}

/* SRC/rt_solvers_mod.f90:785-797 */
void orc_linear_weights(float t, float *E, float *W, float *Psi, float *Esc)
{
  float u0, ut, ps;
  int thin;
  orc_exp_neg(t, E, &u0);
  ut = u0 / t;
  ps = orc_lw_series(t);
  thin = fabsf(t) <= 1.0f;
  *Psi = thin ? ps : 1.0f - ut;
  *Esc = thin ? 1.0f - ps : ut;
  *W = thin ? u0 - ps : ut - *E;
}

/* SRC/rt_solvers_mod.f90:862-877 */
void orc_parab_moments(float t, float *E, float *U0, float *U1, float *q)
{
  const float R0 = +3.333333433e-01f, R1 = -8.333329856e-02f, R2 = +1.666668802e-02f,
              R3 = -2.777901944e-03f, R4 = +3.968053206e-04f, R5 = -4.948913192e-05f,
              R6 = +5.514686109e-06f, R7 = -5.894584660e-07f, R8 = +5.153443894e-08f;
  float qs;
  orc_exp_neg(t, E, U0);
  *U1 = fabsf(t) <= 1.0f ? t * orc_lw_series(t) : t - *U0;
  qs = t * t * (R0 + t * (R1 + t * (R2 + t * (R3 + t * (R4 + t * (R5 + t * (R6 + t * (R7 + t * R8))))))));
  *q = fabsf(t) <= 2.0f ? qs : t - 2.0f * (*U1 / t);
}

void orc_linear_weights_n(int n, const float *t, float *E, float *W, float *Psi, float *Esc)
{
  for (int i = 0; i < n; i++) orc_linear_weights(t[i], E + i, W + i, Psi + i, Esc + i);
}
void orc_parab_moments_n(int n, const float *t, float *E, float *U0, float *U1, float *q)
{
  for (int i = 0; i < n; i++) orc_parab_moments(t[i], E + i, U0 + i, U1 + i, q + i);
}
void orc_exp_neg_n(int n, const float *t, float *E, float *u0)
{
  for (int i = 0; i < n; i++) orc_exp_neg(t[i], E + i, u0 + i);
}

/* ---------------------------------------------------------------- FP64 weights */

/* Exact linear-step weights in FP64: E = e^-t, u0 = 1 - e^-t, Psi = 1 - u0/t,
 * W = u0 - Psi, Esc = 1 - Psi, with the Taylor series of Psi below |t| = 1e-2. */
static void lw64(double t, double *E, double *W, double *Psi, double *Esc)
{
  double u0 = -expm1(-t);
  *E = exp(-t);
  if (fabs(t) < 1e-2) {
    double ps = t * (0.5 + t * (-1.0 / 6 + t * (1.0 / 24 + t * (-1.0 / 120 + t * (1.0 / 720
                + t * (-1.0 / 5040 + t * (1.0 / 40320)))))));
    *Psi = ps;
    *Esc = 1.0 - ps;
    *W = u0 - ps;
  } else {
    double ut = u0 / t;
    *Psi = 1.0 - ut;
    *Esc = ut;
    *W = ut - *E;
  }
}

/* ---------------------------------------------------------------- geometry */

/* SRC/ray_mod.f90:581-596: Keys cubic convolution, a = -1/2 */
void orc_cc_kernel(double s, double k[4])
{
  double s2 = s * s, s3 = s2 * s;
  k[0] = 0.5 * (-s3 + 2 * s2 - s);
  k[1] = 0.5 * (3 * s3 - 5 * s2 + 2);
  k[2] = 0.5 * (-3 * s3 + 4 * s2 + s);
  k[3] = 0.5 * (s3 - s2);
}

/* SRC/ray_mod.f90:493-508 */
void orc_set_stencil(const double ds[3], const double vec[3], orc_ray_t *ray)
{
  double wt1, wt2, h1, h2;
  ray->dsup = ds[2] / fabs(vec[2]);
  wt1 = -vec[0] / ds[0] * ray->dsup;
  wt2 = -vec[1] / ds[1] * ray->dsup;
  h1 = wt1 > 0 ? wt1 : wt1 + 1;
  h2 = wt2 > 0 ? wt2 : wt2 + 1;
  ray->upo1 = wt1 > 0 ? -1 : -2;
  ray->upo2 = wt2 > 0 ? -1 : -2;
  orc_cc_kernel(h1, ray->xkern);
  orc_cc_kernel(h2, ray->ykern);
}

/* SRC/ray_mod.f90:269-311 (nodes and weights are default-real literals in the
 * reference, hence the float rounding) and 392-467 (nphi_copies). PI is the
 * reference's const_mod value (HOP/const_mod.f90:14). */
int orc_gauss_quadrature(int ntheta, int nphi_in, int jnu_weights, int one_dimensional,
                         double *vec, double *wts)
{
  static const float MU[6][6] = {
    {1.0f},
    {0.78867513f, 0.21132487f},
    {0.88729833f, 0.50000000f, 0.11270167f},
    {0.93056816f, 0.66999052f, 0.33000948f, 0.06943184f},
    {0.95308992f, 0.76923466f, 0.50000000f, 0.23076534f, 0.04691008f},
    {0.96623476f, 0.83060469f, 0.61930959f, 0.38069041f, 0.16939531f, 0.03376524f}};
  static const float WMU[6][6] = {
    {1.0f},
    {0.50000000f, 0.50000000f},
    {0.27777778f, 0.44444444f, 0.27777778f},
    {0.17392742f, 0.32607258f, 0.32607258f, 0.17392742f},
    {0.11846344f, 0.23931434f, 0.28444444f, 0.23931434f, 0.11846344f},
    {0.08566225f, 0.18038079f, 0.23395697f, 0.23395697f, 0.18038079f, 0.08566225f}};
  const double PI = 3.14159265359;
  double mu[6], wmu[6];
  int nrays, includes_mu1 = 0, nmuphi, n1, nphi;
  if (ntheta < 1 || ntheta > 6) return -1;
  for (int i = 0; i < ntheta; i++) {
    mu[i] = (double)MU[ntheta - 1][i];
    wmu[i] = (double)WMU[ntheta - 1][i];
    if (mu[i] == 1.0) includes_mu1 = 1;
  }
  nrays = ntheta;
  nphi = one_dimensional ? 1 : nphi_in;
  nmuphi = includes_mu1 ? (nrays - 1) * nphi + 1 : nrays * nphi;
  if (jnu_weights) nmuphi *= 2;
  if (includes_mu1) {
    wts[0] = wmu[0];
    vec[0] = 0; vec[1] = 0; vec[2] = 1;
    n1 = 2;
  } else {
    n1 = 1;
  }
  for (int nt = n1; nt <= nrays; nt++)
    for (int np = 1; np <= nphi; np++) {
      int i = (nt - n1) * nphi + np;
      if (includes_mu1) i++;
      double xphi = 2.0 * PI / (double)nphi * (np - 1);
      double theta = acos(mu[nt - 1]);
      vec[3 * (i - 1) + 0] = sin(theta) * cos(xphi);
      vec[3 * (i - 1) + 1] = sin(theta) * sin(xphi);
      vec[3 * (i - 1) + 2] = mu[nt - 1];
      wts[i - 1] = wmu[nt - 1] / (double)nphi;
    }
  {
    int half = jnu_weights ? nmuphi / 2 : nmuphi;
    for (int i = 0; i < 3 * half; i++)
      if (fabs(vec[i]) < 1e-3) vec[i] = 0.0;
    if (jnu_weights) {
      for (int i = 0; i < half; i++) {
        wts[half + i] = wts[i];
        for (int c = 0; c < 3; c++) vec[3 * (half + i) + c] = -vec[3 * i + c];
      }
      for (int i = 0; i < nmuphi; i++) wts[i] /= 2;
    } else {
      for (int i = 0; i < nmuphi; i++) wts[i] *= vec[3 * i + 2];
    }
  }
  return nmuphi;
}

/* ---------------------------------------------------------------- solvers */

#define MIN_(a, b) ((a) < (b) ? (a) : (b))
#define MAX_(a, b) ((a) > (b) ? (a) : (b))

#define REAL float
#define FN(n) n##_f32
#define WEIGHTS orc_linear_weights
#define TINY FLT_MIN
#define fabs_ fabsf
#define copysign_ copysignf
#include "sc2_body.inc"
#undef REAL
#undef FN
#undef WEIGHTS
#undef TINY
#undef fabs_
#undef copysign_

#define REAL double
#define FN(n) n##_f64
#define WEIGHTS lw64
#define TINY FLT_MIN
#define fabs_ fabs
#define copysign_ copysign
#include "sc2_body.inc"
#undef REAL
#undef FN
#undef WEIGHTS
#undef TINY
#undef fabs_
#undef copysign_

/* short_char, SRC/rt_solvers_mod.f90:19-189 (no amr): chi, S and Inu by the 4x4
 * cubic kernel, each clamped from below by the inner 2x2 minimum; dt from the
 * mean of chi at the upwind point and the cell. */
void orc_short_char_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                        float *Psi, float *Irem, float *escape)
{
  const int nf = p->nf, nx = p->nx, ny = p->ny, nz = p->nz;
  const size_t sx = (size_t)(p->nf_stride ? p->nf_stride : nf), sy = sx * nx, sz = sy * ny;
  orc_ray_t ray;
  float xkern[4], ykern[4], dsup05;
  int xup_t[4 * ORC_MAXN], yup_t[4 * ORC_MAXN];
  const float *wk[3] = {chi, S, I};
  float *row_red = (float *)malloc(sizeof(float) * (size_t)nf * 3 * nx * ny);
  float *wup = (float *)malloc(sizeof(float) * (size_t)nf * 3);
  float *wmin = (float *)malloc(sizeof(float) * (size_t)nf * 3);
  const int zs = p->dir > 0 ? 1 : nz, ze = p->dir > 0 ? nz : 1;
  const float mdt = -(float)p->maser_dtau_max, imax = (float)p->maser_Inu_max;
#define RR(f, i, x, y) row_red[(f) + (size_t)nf * ((i) + 3 * ((x) + (size_t)nx * (y)))]
#define W_(a, f, x, y, z) a[(f) + sx * (x) + sy * (y) + sz * (size_t)(z)]
  orc_set_stencil(p->ds, p->vec, &ray);
  for (int i = 0; i < 4; i++) { xkern[i] = (float)ray.xkern[i]; ykern[i] = (float)ray.ykern[i]; }
  dsup05 = 0.5f * (float)ray.dsup;
  for (int x = 0; x < nx; x++)
    for (int ii = 0; ii < 4; ii++) {
      int v = (ii + 1) + (x + 1) + ray.upo1 - 2;
      xup_t[4 * x + ii] = ((v % nx) + nx) % nx;
    }
  for (int y = 0; y < ny; y++)
    for (int jj = 0; jj < 4; jj++) {
      int v = (jj + 1) + (y + 1) + ray.upo2 - 2;
      yup_t[4 * y + jj] = ((v % ny) + ny) % ny;
    }
  for (int z = zs; p->dir > 0 ? z <= ze : z >= ze; z += p->dir) {
    const int zup = z - p->dir;
    memset(row_red, 0, sizeof(float) * (size_t)nf * 3 * nx * ny);
    for (int yup = 0; yup < ny; yup++)
      for (int x = 0; x < nx; x++)
        for (int ii = 0; ii < 4; ii++) {
          const int xup = xup_t[4 * x + ii];
          const float kx = xkern[ii];
          for (int i = 0; i < 3; i++)
            for (int f = 0; f < nf; f++)
              RR(f, i, x, yup) = RR(f, i, x, yup) + kx * W_(wk[i], f, xup, yup, zup);
        }
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        for (int i = 0; i < 3 * nf; i++) { wup[i] = 0.0f; wmin[i] = FLT_MAX; }
        for (int jj = 0; jj < 4; jj++) {
          const int yup = yup_t[4 * y + jj];
          const float ky = ykern[jj];
          for (int i = 0; i < 3; i++)
            for (int f = 0; f < nf; f++) wup[f + nf * i] = wup[f + nf * i] + ky * RR(f, i, x, yup);
        }
        for (int jj = 1; jj <= 2; jj++)
          for (int ii = 1; ii <= 2; ii++) {
            const int xup = xup_t[4 * x + ii], yup = yup_t[4 * y + jj];
            for (int i = 0; i < 3; i++)
              for (int f = 0; f < nf; f++)
                wmin[f + nf * i] = MIN_(wmin[f + nf * i], W_(wk[i], f, xup, yup, zup));
          }
        for (int i = 0; i < 3 * nf; i++) wup[i] = MAX_(wup[i], wmin[i]);
        for (int f = 0; f < nf; f++) {
          const float xxup = wup[f], ssup = wup[f + nf], Inuup = wup[f + 2 * nf];
          const float xx = W_(chi, f, x, y, z), ss = W_(S, f, x, y, z);
          float dt = dsup05 * (xxup + xx), E, W, Ps, Es, Inu;
          dt = MAX_(dt, mdt);
          orc_linear_weights(dt, &E, &W, &Ps, &Es);
          Inu = Inuup * E + W * ssup + Ps * ss;
          if (p->split_rates) {
            const int off = dt <= 0 || xx <= 0 || fabsf(Inu) > imax;
            const float Ir = Inuup * E + W * ssup;
            Inu = MAX_(-imax, MIN_(imax, Inu));
            Ps = off ? 0.0f : Ps;
            if (Irem) W_(Irem, f, x, y, z - 1) = off ? Inu : Ir;
            if (escape) W_(escape, f, x, y, z - 1) = off ? 1.0f : Es;
          } else {
            Inu = MAX_(-imax, MIN_(imax, Inu));
          }
          W_(I, f, x, y, z) = Inu;
          if (p->get_rates && Psi) W_(Psi, f, x, y, z - 1) = (dt > 0 && xx > 0) ? Ps / xx : 0.0f;
        }
      }
  }
  free(row_red); free(wup); free(wmin);
#undef RR
#undef W_
}

/* long_char_quadratic + quad_step, SRC/rt_solvers_mod.f90:891-1020 */
static void quad_step(float xxup, float xxc, float xxdn, float ssup, float ssc, float ssdn,
                      float Inuup, int at_target, float rdsup, float dsup1, float dsup2,
                      float mdtau, float imax, int get_rates, float *Inu, float *plam)
{
  const float dtmin = 1e-20f;
  float t1, t2, t3, dtup, dtdn, dsdtau, dt, rdt, ex, U0, U1, q, A, B, C;
  t1 = (xxc - xxup) * rdsup;
  t3 = (xxdn - xxc) * rdsup;
  t2 = 0.5f * (t1 + t3);
  t2 = (copysignf(1.0f, t1) + copysignf(1.0f, t3)) * MIN_(MIN_(fabsf(t1), fabsf(t3)), fabsf(0.5f * t2));
  dtup = dsup1 * (xxup + xxc) + dsup2 * (t1 - t2);
  dtdn = dsup1 * (xxdn + xxc) + dsup2 * (t2 - t3);
  dtup = MAX_(copysignf(MAX_(fabsf(dtup), dtmin), dtup), mdtau);
  dtdn = MAX_(copysignf(MAX_(fabsf(dtdn), dtmin), dtdn), mdtau);
  dt = at_target ? dtup : dtdn;
  rdt = 1.0f / dt;
  orc_parab_moments(dt, &ex, &U0, &U1, &q);
  if (at_target) {
    t1 = (ssc - ssup) * rdt;
    t3 = (ssdn - ssc) / dtdn;
  } else {
    t1 = (ssc - ssup) / dtup;
    t3 = (ssdn - ssc) * rdt;
  }
  t2 = 0.5f * (t1 + t3);
  dsdtau = (copysignf(1.0f, t1) + copysignf(1.0f, t3)) * MIN_(MIN_(fabsf(t1), fabsf(t3)), fabsf(0.5f * t2));
  *plam = 0.0f;
  if (at_target) {
    B = (2.0f * U1 - q) * rdt;
    A = U0 - B;
    C = q - U1;
    *Inu = Inuup * ex + A * ssup + B * ssc + C * dsdtau;
    if (get_rates && dt > 0 && xxc > 0) *plam = B / xxc;
  } else {
    A = q * rdt;
    B = U0 - A;
    C = U1 - q;
    *Inu = Inuup * ex + A * ssdn + B * ssc + C * dsdtau;
    if (get_rates && dt > 0 && xxdn > 0) *plam = A / xxdn;
  }
  *Inu = MAX_(-imax, MIN_(imax, *Inu));
}

void orc_long_char_quadratic_f32(const orc_sc_params *p, const float *chi, const float *S,
                                 float *I, float *Psi)
{
  const int nf = p->nf, nx = p->nx, ny = p->ny, nz = p->nz, dir = p->dir;
  const size_t sx = (size_t)(p->nf_stride ? p->nf_stride : nf), sy = sx * nx, sz = sy * ny;
  const float ds = (float)p->ds[2];
  const float dsup = (float)((double)ds / fabs(p->vec[2]));
  const float rdsup = 1.0f / dsup, dsup1 = 0.5f * dsup, dsup2 = 1.0f / 12.0f * dsup * dsup;
  const float mdt = -(float)p->maser_dtau_max, imax = (float)p->maser_Inu_max;
  const int zs = dir > 0 ? 1 : nz, ze = dir > 0 ? nz : 1;
#define W_(a, f, x, y, z) a[(f) + sx * (x) + sy * (y) + sz * (size_t)(z)]
  for (int z = zs; z != ze; z += dir) {
    const int zup = z - dir, zdn = z + dir;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        for (int f = 0; f < nf; f++) {
          float Inu, pl;
          quad_step(W_(chi, f, x, y, zup), W_(chi, f, x, y, z), W_(chi, f, x, y, zdn),
                    W_(S, f, x, y, zup), W_(S, f, x, y, z), W_(S, f, x, y, zdn),
                    W_(I, f, x, y, zup), 1, rdsup, dsup1, dsup2, mdt, imax, p->get_rates, &Inu, &pl);
          W_(I, f, x, y, z) = Inu;
          if (p->get_rates && Psi) W_(Psi, f, x, y, z - 1) = pl;
        }
  }
  {
    const int z = ze, z0 = z - dir, zup = z0 - dir;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        for (int f = 0; f < nf; f++) {
          float Inu, pl;
          quad_step(W_(chi, f, x, y, zup), W_(chi, f, x, y, z0), W_(chi, f, x, y, z),
                    W_(S, f, x, y, zup), W_(S, f, x, y, z0), W_(S, f, x, y, z),
                    W_(I, f, x, y, z0), 0, rdsup, dsup1, dsup2, mdt, imax, p->get_rates, &Inu, &pl);
          W_(I, f, x, y, z) = Inu;
          if (p->get_rates && Psi) W_(Psi, f, x, y, z - 1) = pl;
        }
  }
#undef W_
}

/* SRC/rt_patch_mod.f90:3399-3431; I and Psi hold planes 1..nz (no ghosts) */
void orc_integrate_intensity(int nf, int nx, int ny, int nz, float wmu, const float *I,
                             const float *Psi, float *J, float *PsiJ)
{
  const float wmu2 = 2 * wmu;
  const size_t n = (size_t)nf * nx * ny * nz;
  for (size_t i = 0; i < n; i++) J[i] = J[i] + wmu2 * I[i];
  if (PsiJ && Psi)
    for (size_t i = 0; i < n; i++) PsiJ[i] = PsiJ[i] + wmu2 * Psi[i];
}

/* CPU baseline driver: the reference's parallelisation of the formal solution
 * (SRC/rt_column_mod.f90:1-15, 252-321; SRC/ray_patch_mod.f90:97-108): the
 * frequency axis split into n_windows independent columns, each swept by one
 * OpenMP thread through every plane, followed by integrate_intensity
 * (SRC/rt_patch_mod.f90:3399-3431) over the window.  Work arrays as for
 * orc_short_char2_f32; J is (nf, nx, ny, nz).  wmu2 is the weight of J += wmu2*I. */
void orc_sweep_windows_f32(const orc_sc_params *p, int nwin, const float *chi, const float *S,
                           float *I, const float *chi500, const float *chi500_up, const float *g500,
                           float wmu2, float *J)
{
  const int nf = p->nf;
  if (nwin < 1) nwin = 1;
  if (nwin > nf) nwin = nf;
#pragma omp parallel for schedule(dynamic, 1)
  for (int w = 0; w < nwin; w++) {
    const int fs = (int)((long)nf * w / nwin), fe = (int)((long)nf * (w + 1) / nwin);
    orc_sc_params q = *p;
    q.nf = fe - fs;
    q.nf_stride = nf;
    q.get_rates = 0;
    q.split_rates = 0;
    orc_short_char2_f32(&q, chi + fs, S + fs, I + fs, chi500, chi500_up, g500, NULL, NULL,
                        NULL, NULL, NULL);
    const size_t sx = (size_t)nf, sy = sx * p->nx, sz = sy * p->ny;
    for (int z = 1; z <= p->nz; z++)
      for (int y = 0; y < p->ny; y++)
        for (int x = 0; x < p->nx; x++) {
          const size_t o = sx * x + sy * y;
          for (int f = fs; f < fe; f++)
            J[o + sz * (z - 1) + f] = J[o + sz * (z - 1) + f] + wmu2 * I[o + sz * z + f];
        }
  }
}
