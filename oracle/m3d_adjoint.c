/*
 * m3d_adjoint.c -- TEST INFRASTRUCTURE ONLY: the reverse-mode derivative of the
 * oracle's FP64 short_char2 (SRC/rt_solvers_mod.f90:224-507 of experiments/
 * OpenMULTI3D/ in Bergemann-Group-development/dispatch; the S step of 409-440),
 * the checker for the HIP adjoint sweep.  The reference has no adjoint; this is
 * the derivative of the forward restatement, one plane at a time in reverse sweep
 * order:
 *
 *   I = max(-Imax, min(Imax, E Iup + W Sup + Psi S)),  Iup = max(cubic(I'), min2x2(I')),
 *   Sup = bilinear(S'),  xup = bilinear(chi' / chi500'),  dt = max(dt500 (xup + chi/chi500)/2, -dtmax)
 *
 * with primes on the upwind plane.  The cotangent of I feeds S (Psi), the upwind
 * intensity (E, through the cubic weights or, where the clamp holds, to the smallest
 * inner value), the upwind source (W, bilinear) and the step (E' Iup + W' Sup + Psi' S)
 * shared by chi and the upwind chi.  Ties of the clamps go where the forward's
 * MAX_/MIN_ send the value: the clamp holds when the cubic value is not above the
 * minimum, and the minimum is the last smallest of (x2,y2), (x3,y2), (x2,y3), (x3,y3).
 *
 * Pinning (tests/test_oracle_adjoint.py): central finite differences of
 * orc_short_char2_f64 for chi, S and the ghost plane's chi, S and I in both sweep
 * directions; the weight derivatives against differences of the weights.
 */
#include <math.h>
#include <stdlib.h>
#include "m3d_adjoint.h"

#define MIN_(a, b) ((a) < (b) ? (a) : (b))
#define MAX_(a, b) ((a) > (b) ? (a) : (b))

void orc_lin_weights_deriv_f64(double t, double *E, double *W, double *Psi, double *dE, double *dW,
                               double *dPsi)
{
  const double e = exp(-t), u0 = -expm1(-t);
  double ps, dps;
  if (fabs(t) < 5e-2) {
    /* Psi = sum_{n>=1} (-1)^(n+1) t^n / (n+1)!,  Psi' = sum_{n>=2} (-1)^n (n-1)/n! t^(n-2) */
    ps = t * (1.0 / 2 + t * (-1.0 / 6 + t * (1.0 / 24 + t * (-1.0 / 120 + t * (1.0 / 720
         + t * (-1.0 / 5040 + t * (1.0 / 40320 + t * (-1.0 / 362880))))))));
    dps = 1.0 / 2 + t * (-1.0 / 3 + t * (1.0 / 8 + t * (-1.0 / 30 + t * (1.0 / 144
          + t * (-1.0 / 840 + t * (1.0 / 5760 + t * (-1.0 / 45360)))))));
  } else {
    ps = 1.0 - u0 / t;
    dps = (u0 - t * e) / (t * t);
  }
  *E = e;
  *Psi = ps;
  *W = u0 - ps;
  *dE = -e;
  *dPsi = dps;
  *dW = e - dps;
}

void orc_short_char2_adjoint_f64(const orc_sc_params *p, const double *chi, const double *S, const double *I,
                                 const double *chi500, const double *chi500_up, const double *g500,
                                 double *Ibar, double *chibar, double *Sbar)
{
  const int nf = p->nf, nx = p->nx, ny = p->ny, nz = p->nz, dir = p->dir;
  const size_t sx = (size_t)nf, sy = sx * nx, sz = sy * ny, nxy = (size_t)nx * ny;
  const double mdt = -p->maser_dtau_max, imax = p->maser_Inu_max;
  orc_ray_t ray;
  double xk[4], yk[4];
  orc_set_stencil(p->ds, p->vec, &ray);
  for (int i = 0; i < 4; i++) { xk[i] = ray.xkern[i]; yk[i] = ray.ykern[i]; }
  const double hx = -xk[0] + xk[2] + 2 * xk[3], hy = -yk[0] + yk[2] + 2 * yk[3];
  const double wx[2] = {1 - hx, hx}, wy[2] = {1 - hy, hy};
  const double dsup = ray.dsup;
  const int zs = dir > 0 ? 1 : nz, ze = dir > 0 ? nz : 1;
  double *r500 = (double *)malloc(sizeof(double) * nxy);
#define W_(a, f, x, y, z) a[(f) + sx * (x) + sy * (y) + sz * (size_t)(z)]
#define C3(a, x, y, z) a[(x) + (size_t)nx * (y) + nxy * (size_t)((z) - 1)]
  for (int z = ze; dir > 0 ? z >= zs : z <= zs; z -= dir) {
    const int zup = z - dir;
    const int ghost = zup < 1 || zup > nz;
    for (size_t i = 0; i < nxy; i++) r500[i] = ghost ? 1.0 / g500[i] : 1.0 / chi500[i + nxy * (size_t)(zup - 1)];
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        int xs[4], ys[4];
        for (int i = 0; i < 4; i++) {
          xs[i] = (((i + x + ray.upo1) % nx) + nx) % nx;
          ys[i] = (((i + y + ray.upo2) % ny) + ny) % ny;
        }
        const double c = C3(chi500, x, y, z), cu = C3(chi500_up, x, y, z), r = cu / c;
        const double dt500 = fabs(r - 1.0) < 1e-3 ? dsup * c * (1.0 + (r - 1.0) * (0.5 - (r - 1.0) / 12.0))
                                                  : dsup * c * (r - 1.0) / log(r);
        const double rc = 1.0 / c;
        for (int f = 0; f < nf; f++) {
          double xxup = 0, ssup = 0, iraw = 0;
          for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
              const int xi = xs[i + 1], yj = ys[j + 1];
              xxup += wx[i] * wy[j] * W_(chi, f, xi, yj, zup) * r500[xi + nx * yj];
              ssup += wx[i] * wy[j] * W_(S, f, xi, yj, zup);
            }
          for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) iraw += yk[j] * xk[i] * W_(I, f, xs[i], ys[j], zup);
          /* the last smallest of the inner four, as MIN_ chains pick it */
          const int ci[4] = {1, 2, 1, 2}, cj[4] = {1, 1, 2, 2};
          double mn = W_(I, f, xs[1], ys[1], zup);
          int arg = 0;
          for (int k = 1; k < 4; k++) {
            const double v = W_(I, f, xs[ci[k]], ys[cj[k]], zup);
            if (!(mn < v)) { mn = v; arg = k; }
          }
          const int clamp = !(iraw > mn);
          const double Iup = clamp ? mn : iraw;
          const double chv = W_(chi, f, x, y, z), ss = W_(S, f, x, y, z);
          const double xx = chv * rc;
          const double dtr = 0.5 * (xxup + xx) * dt500;
          const double dt = MAX_(dtr, mdt);
          double E, W, Ps, dE, dW, dPs;
          orc_lin_weights_deriv_f64(dt, &E, &W, &Ps, &dE, &dW, &dPs);
          const double Inu = Iup * E + W * ssup + Ps * ss;
          const double g = (Inu >= -imax && Inu <= imax) ? W_(Ibar, f, x, y, z) : 0.0;
          if (g == 0.0) continue;
          W_(Sbar, f, x, y, z) += g * Ps;
          const double gup = g * E, gss = g * W;
          const double gdt = dtr > mdt ? g * (Iup * dE + ssup * dW + ss * dPs) : 0.0;
          const double gxx = 0.5 * gdt * dt500;
          W_(chibar, f, x, y, z) += gxx * rc;
          for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
              const int xi = xs[i + 1], yj = ys[j + 1];
              W_(Sbar, f, xi, yj, zup) += gss * wx[i] * wy[j];
              W_(chibar, f, xi, yj, zup) += gxx * wx[i] * wy[j] * r500[xi + nx * yj];
            }
          if (clamp) {
            W_(Ibar, f, xs[ci[arg]], ys[cj[arg]], zup) += gup;
          } else {
            for (int j = 0; j < 4; j++)
              for (int i = 0; i < 4; i++) W_(Ibar, f, xs[i], ys[j], zup) += gup * yk[j] * xk[i];
          }
        }
      }
  }
  free(r500);
#undef W_
#undef C3
}
