/*
 * m3d_b1d.c -- TEST INFRASTRUCTURE ONLY: CPU oracle of the batched 1D LTE
 * synthesis (the checker for dispatch_amd/csrc/batch1d.hip).  Citations are
 * relative to experiments/OpenMULTI3D/ of Bergemann-Group-development/dispatch.
 *
 * Per frequency of one atmosphere, from the bottom point up: chi/chi500 = the
 * continuum plus every line inside its window, s K(a, v) with K the QLG-bot profile
 * (bot_setup and qlg_bot_line_eval, HOP/qlg_mod.f90:524-556, restated in
 * m3d_qlg.c), S = B_nu(T) = 1 / (exp(h c / (k T lam)) - 1), and for each upward mu
 * long_char2's linear step in tau500 (SRC/rt_solvers_mod.f90:527-663) with exact
 * weights in FP64; the bottom point thermalised, I = S.
 *
 * What is in FP32, as the kernel forms it: the window test (|lam_f - lam_l|, both as
 * offsets from lam0, against the half-width) and the profile's table index (bot_setup's
 * ih, the Doppler scale (c/lam_l)(1/v_D) ih, xs = |dl| times it, snapped at 800 per
 * unit): these are decisions, and a decision taken differently is not a rounding
 * error.  Everything else -- the profile's value from the same table entries, the
 * sums, the Planck function, the exponentials of the step -- is FP64.
 *
 * Pinning (tests/test_b1d_cpu.py): a line-free atmosphere whose source function is
 * linear in tau500 gives the Eddington-Barbier intensity a + b mu exactly; an
 * isothermal one gives I = B; the profile equals the oracle's qlg_bot_sweep (itself
 * bit-exact with the compiled reference) and Re w(v + ia) to QLG's 5.7e-3; outside
 * its window a line stays below the tolerance against the continuum.
 */
#include <math.h>
#include <stddef.h>
#include "m3d_b1d.h"
#include "m3d_qlg.h"

#define QLG_NX 9019

static float tgx[QLG_NX + 1], th2x[QLG_NX + 1];
static int have_tab = 0;

/* exact linear-step weights: E = e^-t, Psi = 1 - (1 - E)/t, W = (1 - E) - Psi */
static void weights(double t, double *E, double *W, double *Ps)
{
  const double u0 = -expm1(-t);
  *E = exp(-t);
  if (fabs(t) < 1e-4) {
    *Ps = t * (0.5 - t * (1.0 / 6.0 - t / 24.0));
  } else {
    *Ps = 1.0 - u0 / t;
  }
  *W = u0 - *Ps;
}

void orc_b1d_atmosphere(int nd, int nf, int nmu, int nline, double lam0, double dlam, const float *dtau,
                        const float *hckT, const float *ivth, const float *cont_a, const float *cont_b,
                        const float *line_off, const float *line_cl, const float *line_rec, const float *mu,
                        const float *wflux, double lam_ref, double *I_top, double *flux, double *chi_out)
{
  if (!have_tab) {
    orc_qlg_tables(tgx, th2x);
    have_tab = 1;
  }
  for (int f = 0; f < nf; f++) {
    const double lam = lam0 + (double)f * dlam;
    const float xf = (float)((double)f * dlam);
    const double q3 = pow(lam / lam_ref, 3.0);
    double I[8], xu = 0.0, Su = 0.0;
    for (int d = nd - 1; d >= 0; d--) {
      double x = (double)cont_a[d] + (double)cont_b[d] * q3;
      for (int l = 0; l < nline; l++) {
        const float *r = line_rec + 4 * ((size_t)d * nline + l);
        const float dl = fabsf(xf - line_off[l]);
        if (!(dl <= r[2])) continue;
        orc_qlg_coef c;
        orc_qlg_bot_setup(r[0], 1.0f, 0.8325546112f, 1.7724538509f, &c);
        const float xsc = (line_cl[l] * ivth[d]) * c.ih;
        const float xs = dl * xsc;
        const float sn = xs * 800.0f;
        const int ig = (int)roundf(sn < (float)QLG_NX ? sn : (float)QLG_NX);
        const double xd = (double)xs, L = 1.0 / (1.0 + xd * xd);
        const double K = (double)tgx[ig] * c.gG + L * ((double)c.bL + L * ((double)c.cL2 + (double)c.cZH * th2x[ig]));
        x += (double)r[1] * K;
      }
      const double S = 1.0 / expm1((double)hckT[d] / lam);
      if (chi_out) chi_out[(size_t)d * nf + f] = x;
      if (d == nd - 1) {
        for (int k = 0; k < nmu; k++) I[k] = S;
      } else {
        for (int k = 0; k < nmu; k++) {
          const double t = 0.5 * (xu + x) * (double)dtau[d] / (double)mu[k];
          double E, W, Ps;
          weights(t, &E, &W, &Ps);
          I[k] = I[k] * E + W * Su + Ps * S;
        }
      }
      xu = x;
      Su = S;
    }
    double F = 0.0;
    for (int k = 0; k < nmu; k++) {
      I_top[(size_t)k * nf + f] = I[k];
      F += (double)wflux[k] * I[k];
    }
    flux[f] = F;
  }
}
