/*
 * m3d_hermite.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of short_char_hermite
 * (SRC/rt_solvers_mod.f90:1917-2248 of experiments/OpenMULTI3D/ in
 * Bergemann-Group-development/dispatch), the 3D short-characteristics solver with
 * quadratic (parabolic-moment) weights, the checker for the HIP kernel.
 *
 * Upwind values of chi, S, I and, past the slab's first plane, of chi/Pe and S/B
 * (S (1-beta)/beta) from the 4x4 Keys stencil, each clamped from below by the inner
 * 2x2 minimum.  The step length is the trapezoid plus the Euler-Maclaurin term with
 * the target slope of chi from the product rule chi/Pe * dPe/ds + Pe d(chi/Pe)/ds,
 * MC-limited; the source slope dS/dtau = S (dB/dtau)/B + B d(S/B)/dtau, clamped to
 * [0, 2x] the secant; I = I_up e^-dt + A S_up + B S + C dS/dtau from parab_moments,
 * with A, B, C in FP64 as the reference keeps them.  The slab's first plane takes
 * the linear step.  The AMR boundary stencil is not restated (amr = '').
 *
 * Pinning (tests/test_oracle_hermite.py): exact for a source linear along the ray in
 * an isothermal medium of constant chi/Pe; the first plane equals short_char's
 * step; a constant source stays fixed.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "m3d_hermite.h"

#define MIN_(a, b) ((a) < (b) ? (a) : (b))
#define MAX_(a, b) ((a) > (b) ? (a) : (b))

void orc_short_char_hermite_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                                const float *beta, const float *pe4, const float *temp, const float *dpe_dx,
                                const float *dpe_dy, const float *dpe_dz, const float *dlnt_dx,
                                const float *dlnt_dy, const float *dlnt_dz, const float *hknu, float *Psi,
                                float *Irem, float *escape)
{
  const int nf = p->nf, nx = p->nx, ny = p->ny, nz = p->nz, dir = p->dir;
  const size_t sx = (size_t)nf, sy = sx * nx, sz = sy * ny, nxy = (size_t)nx * ny;
  const float dtmin = 1e-20f;
  const double mdt = -p->maser_dtau_max;
  const float mdt4 = -(float)p->maser_dtau_max, imax = (float)p->maser_Inu_max;
  orc_ray_t ray;
  float xk[4], yk[4];
  int xup_t[4 * 4096], yup_t[4 * 4096];
  orc_set_stencil(p->ds, p->vec, &ray);
  for (int i = 0; i < 4; i++) { xk[i] = (float)ray.xkern[i]; yk[i] = (float)ray.ykern[i]; }
  const float dsup05 = 0.5f * (float)ray.dsup;
  const float rpath = 0.5f / dsup05, dsc2 = dsup05 * dsup05 / 3.0f;
  const double sg = p->vec[2] >= 0 ? 1.0 : -1.0;
  const float v1 = (float)(p->vec[0] * dir * sg), v2 = (float)(p->vec[1] * dir * sg),
              v3 = (float)(p->vec[2] * dir * sg);
  for (int x = 0; x < nx; x++)
    for (int ii = 0; ii < 4; ii++) xup_t[4 * x + ii] = ((ii + x + ray.upo1) % nx + nx) % nx;
  for (int y = 0; y < ny; y++)
    for (int jj = 0; jj < 4; jj++) yup_t[4 * y + jj] = ((jj + y + ray.upo2) % ny + ny) % ny;
  float *plane_r = (float *)malloc(sizeof(float) * nf * nxy * 2);
  float *row_red = (float *)malloc(sizeof(float) * nf * nxy * 5);
  float wup[5], wmin[5];
  const int zs = dir > 0 ? 1 : nz, ze = dir > 0 ? nz : 1;
#define W_(a, f, x, y, z) a[(f) + sx * (x) + sy * (y) + sz * (size_t)(z)]
#define C_(a, x, y, z) a[(x) + (size_t)nx * (y) + nxy * (size_t)((z) - 1)]
#define B_(f, x, y, z) beta[(f) + sx * (x) + sy * (y) + sz * (size_t)((z) - 1)]
#define PR(f, i, x, y) plane_r[(f) + (size_t)nf * ((x) + nx * (size_t)(y) + nxy * (i))]
#define RR(f, i, x, y) row_red[(f) + (size_t)nf * ((i) + 5 * ((x) + nx * (size_t)(y)))]
  for (int z = zs; dir > 0 ? z <= ze : z >= ze; z += dir) {
    const int zup = z - dir;
    const int chain = z != zs;
    if (chain)
      for (int yu = 0; yu < ny; yu++)
        for (int xu = 0; xu < nx; xu++) {
          const float rpe = 1.0f / C_(pe4, xu, yu, zup);
          for (int f = 0; f < nf; f++) {
            const float bet = B_(f, xu, yu, zup);
            PR(f, 0, xu, yu) = W_(chi, f, xu, yu, zup) * rpe;
            PR(f, 1, xu, yu) = W_(S, f, xu, yu, zup) * (1.0f - bet) / bet;
          }
        }
    memset(row_red, 0, sizeof(float) * nf * nxy * 5);
    for (int yu = 0; yu < ny; yu++)
      for (int x = 0; x < nx; x++)
        for (int ii = 0; ii < 4; ii++) {
          const int xu = xup_t[4 * x + ii];
          const float kx = xk[ii];
          for (int f = 0; f < nf; f++) {
            RR(f, 0, x, yu) = RR(f, 0, x, yu) + kx * W_(chi, f, xu, yu, zup);
            RR(f, 1, x, yu) = RR(f, 1, x, yu) + kx * W_(S, f, xu, yu, zup);
            RR(f, 2, x, yu) = RR(f, 2, x, yu) + kx * W_(I, f, xu, yu, zup);
            if (chain) {
              RR(f, 3, x, yu) = RR(f, 3, x, yu) + kx * PR(f, 0, xu, yu);
              RR(f, 4, x, yu) = RR(f, 4, x, yu) + kx * PR(f, 1, xu, yu);
            }
          }
        }
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        float gpe = 0, aT = 0, pec = 1, rpec = 1;
        if (chain) {
          gpe = v1 * C_(dpe_dx, x, y, z) + v2 * C_(dpe_dy, x, y, z) + v3 * C_(dpe_dz, x, y, z);
          aT = (v1 * C_(dlnt_dx, x, y, z) + v2 * C_(dlnt_dy, x, y, z) + v3 * C_(dlnt_dz, x, y, z)) / C_(temp, x, y, z);
          pec = C_(pe4, x, y, z);
          rpec = 1.0f / pec;
        }
        for (int f = 0; f < nf; f++) {
          for (int i = 0; i < 5; i++) { wup[i] = 0.0f; wmin[i] = FLT_MAX; }
          for (int jj = 0; jj < 4; jj++) {
            const int yu = yup_t[4 * y + jj];
            for (int i = 0; i < 5; i++) wup[i] = wup[i] + yk[jj] * RR(f, i, x, yu);
          }
          for (int jj = 1; jj <= 2; jj++)
            for (int ii = 1; ii <= 2; ii++) {
              const int xu = xup_t[4 * x + ii], yu = yup_t[4 * y + jj];
              wmin[0] = MIN_(wmin[0], W_(chi, f, xu, yu, zup));
              wmin[1] = MIN_(wmin[1], W_(S, f, xu, yu, zup));
              wmin[2] = MIN_(wmin[2], W_(I, f, xu, yu, zup));
              if (chain) {
                wmin[3] = MIN_(wmin[3], PR(f, 0, xu, yu));
                wmin[4] = MIN_(wmin[4], PR(f, 1, xu, yu));
              }
            }
          for (int i = 0; i < 5; i++) wup[i] = MAX_(wup[i], wmin[i]);
          const float xxup = wup[0], ssup = wup[1], Inuup = wup[2];
          const float xx = W_(chi, f, x, y, z), ss = W_(S, f, x, y, z);
          float Inu, psi_out = 0.0f, rem = 0.0f, esc = 1.0f;
          if (chain) {
            const float rxpup = wup[3], rhou = wup[4], betc = B_(f, x, y, z);
            const float dchi = xx - xxup, t1 = dchi * rpath, sgn = copysignf(1.0f, t1);
            const float rxpc = xx * rpec;
            float mc = rxpc * gpe + pec * (rxpc - rxpup) * rpath;
            mc = sgn * MAX_(0.0f, MIN_(sgn * mc, 2.0f * fabsf(t1)));
            double dt = (double)(dsup05 * (xxup + xx) + dsc2 * (t1 - mc));
            dt = MAX_(copysign(MAX_(fabs(dt), (double)dtmin), dt), mdt);
            const double rdt = 1.0 / dt;
            const float rdt4 = (float)rdt;
            float ex, U0, U1, q;
            orc_parab_moments((float)dt, &ex, &U0, &U1, &q);
            const float r1mbc = 1.0f / (1.0f - betc), Bc = betc * r1mbc;
            const float gc = hknu[f] * r1mbc * aT / xx;
            const float rhoc = ss / Bc, dss = ss - ssup, t1s = dss * rdt4, sgns = copysignf(1.0f, t1s);
            float dsdtau = ss * gc + Bc * (rhoc - rhou) * rdt4, kf, srem;
            if (sgns * dsdtau <= 0.0f) {
              dsdtau = 0.0f; kf = 0.0f; srem = 0.0f;
            } else if (sgns * dsdtau >= 2.0f * fabsf(t1s)) {
              dsdtau = 2.0f * t1s; kf = 2.0f * rdt4; srem = -2.0f * rdt4 * ssup;
            } else {
              kf = gc + rdt4; srem = -Bc * rhou * rdt4;
            }
            const double B = (2.0 * U1 - q) * rdt, A = U0 - B, Cc = q - U1;
            Inu = Inuup * ex + (float)A * ssup + (float)B * ss + (float)Cc * dsdtau;
            Inu = MAX_(-imax, MIN_(imax, Inu));
            const float Psi_loc = (float)(B + Cc * (double)kf);
            if (dt > 0 && xx > 0 && Psi_loc >= 0.0f && Psi_loc <= 1.0f && fabsf(Inu) < imax) {
              psi_out = Psi_loc / xx;
              rem = Inuup * ex + (float)A * ssup + (float)Cc * srem;
              esc = (float)(1.0 - (B + Cc * (double)kf));
            } else {
              psi_out = 0.0f;
              rem = Inu;
              esc = 1.0f;
            }
          } else {
            float dtl = dsup05 * (xxup + xx), El, Wl, Ps4, Es4;
            dtl = MAX_(dtl, mdt4);
            orc_linear_weights(dtl, &El, &Wl, &Ps4, &Es4);
            Inu = Inuup * El + Wl * ssup + Ps4 * ss;
            const int off = dtl <= 0 || xx <= 0 || fabsf(Inu) > imax;
            const float Ir = Inuup * El + Wl * ssup;
            Inu = MAX_(-imax, MIN_(imax, Inu));
            if (p->split_rates) {
              Ps4 = off ? 0.0f : Ps4;
              rem = off ? Inu : Ir;
              esc = off ? 1.0f : Es4;
            }
            psi_out = (dtl > 0 && xx > 0) ? Ps4 / xx : 0.0f;
          }
          W_(I, f, x, y, z) = Inu;
          const size_t o = (size_t)f + sx * x + sy * y + sz * (size_t)(z - 1);
          if (p->get_rates && Psi) Psi[o] = psi_out;
          if (p->split_rates && (p->get_rates || !chain)) {
            if (Irem) Irem[o] = rem;
            if (escape) escape[o] = esc;
          }
        }
      }
  }
  free(plane_r);
  free(row_red);
#undef W_
#undef C_
#undef B_
#undef PR
#undef RR
}
