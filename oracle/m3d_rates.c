/*
 * m3d_rates.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the line-rate
 * integrals of bound_bound_rates (SRC/rt_patch_mod.f90:2852-3142 of
 * experiments/OpenMULTI3D/ in Bergemann-Group-development/dispatch) for one
 * direction, the checker for the HIP rates kernel.
 *
 * Two formulations: 'original' (sum_A, sum_I, sum_P with the preconditioning
 * correction, split into th/dev, SRC/rt_patch_mod.f90:3112-3125) and 'largest'
 * without preconditioning (sum_T = sum phi w Bnu/n_ref, sum_U = sum phi w (I - Bnu),
 * :3100-3107).  Plain FP32 sums as the reference, or Kahan-compensated ones.
 * Pinning: detailed balance (for I = Bnu and no preconditioning, dev = 0 and the
 * thermal integral is A times the profile-weighted Bnu/n_ref), the FP64 sum.
 */
#include <math.h>
#include <stdlib.h>
#include "m3d_qlg.h"
#include "m3d_rates.h"

static void kadd(float *s, float *c, float v, int kahan)
{
  if (!kahan) {
    *s = *s + v;
    return;
  }
  const float y = v - *c;
  const float t = *s + y;
  *c = (t - *s) - y;
  *s = t;
}

void orc_bb_rates(const orc_bb_args *a, float *gamma_ij, float *gamma_ji)
{
  const int nc = a->ncell;
  orc_qlg_init();
  for (int l = 0; l < a->nline; l++) {
    const int fs = a->iblue[l] - 1, fe = a->ired[l] - 1;
    const int li = a->lev_i[l] - 1, lj = a->lev_j[l] - 1;
    const float lower_off = a->lower_off[l];
    for (int c = 0; c < nc; c++) {
      const double nsi = a->nstar[(size_t)li * nc + c], nsj = a->nstar[(size_t)lj * nc + c];
      float ni = (float)(a->b[(size_t)li * nc + c] * nsi);
      float nj = (float)(a->b[(size_t)lj * nc + c] * nsj);
      const float idf = a->inv_df[c], vpr = a->vproj[c], ekt = a->e_kT[c];
      const float adamp = a->adamp[(size_t)l * nc + c];
      float inv_nref = 0.0f;
      if (nsi > 0 && nsj > 0) {
        const double beta_ref = nsj * a->g[li] / (nsi * a->g[lj]);
        double v = (1.0 - beta_ref) / beta_ref;
        inv_nref = (float)(v < 1e35 ? v : 1e35);
      }
      /* the real(4) exp of the reference, taken correctly rounded */
      const float eoff = lower_off != 0 ? (float)exp((double)(-lower_off * ekt)) : 1.0f;
      if (lower_off != 0) {
        ni = ni * eoff;
        nj = nj * eoff;
      }
      float eta_coef = 0, g_ratio = 0;
      if (a->precond) {
        float split_bj = (float)a->b[(size_t)lj * nc + c];
        if (lower_off != 0) split_bj = split_bj * eoff;
        float split_Rij = (float)(nsi * a->pop_ratio[c]);
        split_Rij = (float)((double)split_Rij * a->Bij[l]);
        eta_coef = split_Rij * split_bj;
        g_ratio = 0.0f;
        if (nj > 0.0f) g_ratio = (float)((double)(ni / nj) * (a->g[lj] / a->g[li]));
      }
      float sA = 0, cA = 0, sI = 0, cI = 0, sP = 0, cP = 0, sw = 0, cw = 0;
      for (int f = fs; f <= fe; f++) {
        const size_t o = (size_t)f * a->sf + (size_t)c * a->sc;
        const float wmunu = (float)((double)a->wmu * a->wnu[(size_t)l * a->nf + f] * a->Aij[l]);
        const float phi = orc_qlg_bot_profile(adamp, a->q[(size_t)l * a->nf + f], vpr, idf);
        const float vvv = phi * wmunu;
        if (a->own_norm) kadd(&sw, &cw, phi * (float)a->wq[(size_t)l * a->nf + f], a->kahan);
        if (a->largest) {
          kadd(&sA, &cA, vvv * (a->Bnu[o] * inv_nref), a->kahan);      /* sum_T */
          kadd(&sI, &cI, vvv * (a->Inu[o] - a->Bnu[o]), a->kahan);     /* sum_U */
        } else {
          kadd(&sA, &cA, vvv * (a->Bnu[o] * inv_nref), a->kahan);
          kadd(&sI, &cI, vvv * a->Inu[o], a->kahan);
          kadd(&sP, &cP, vvv * phi * a->Psi[o] * a->beta[o], a->kahan);
        }
      }
      float g_th, g_dev;
      if (a->largest) {
        g_th = sA;
        g_dev = sI;
      } else {
        g_th = sA;
        g_dev = sI;
        if (a->precond) {
          g_th = g_th - sP * eta_coef * (g_ratio - 1.0f);
          g_dev = g_dev - sP * eta_coef;
        }
        if (inv_nref > 0.0f) g_dev = g_dev - g_th / inv_nref;
      }
      if (a->own_norm && sw > 0.0f) {
        g_th = g_th / (sw * a->wtot);
        g_dev = g_dev / (sw * a->wtot);
      }
      gamma_ij[(size_t)l * nc + c] = gamma_ij[(size_t)l * nc + c] + g_th;
      gamma_ji[(size_t)l * nc + c] = gamma_ji[(size_t)l * nc + c] + g_dev;
    }
  }
}
