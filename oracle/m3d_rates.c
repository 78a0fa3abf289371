/*
 * m3d_rates.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the ALI line and
 * continuum terms of experiments/OpenMULTI3D/ in
 * Bergemann-Group-development/dispatch, the checker for the HIP kernels:
 *
 *  orc_bb_rates    bound_bound_rates for one direction (SRC/rt_patch_mod.f90:2852-3142):
 *                  'original' (sum_A, sum_I, sum_P with the preconditioning correction,
 *                  :3112-3125), 'largest' without preconditioning (sum_T, sum_U,
 *                  :3100-3107) and 'largest' with it (the split integrands of
 *                  Rybicki & Hummer 1992 from Irem, escape and the tracked
 *                  contributions, :3062-3098); plain FP32 sums as the reference, or
 *                  Kahan-compensated ones.
 *  orc_bb_opacity  add_bound_bound_opacities (:2689-2831) with add_largest (:3150-3169)
 *                  and the direction's share of wphi.
 *  orc_bf_rates    bound_free_rates (:3193-3259).
 *  orc_work_to_transfer, orc_convert_to_work_units: the (xhat, ehat) <-> (chi, S)
 *                  conversions (:3518-3536, 3571-3651).
 *
 * Pinning (tests/test_oracle_rates.py): closed form of the normalised thermal
 * integral, detailed balance, the preconditioning invariant of the net rate, the
 * normalisation scalings; the split tracking against a from-scratch recount of
 * every (f, cell) sum and the 'largest' split path reducing to the unsplit one
 * where Psi = 0; bf rates by their LTE detailed balance.
 */
#include <float.h>
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

/* other_contribution, SRC/rt_patch_mod.f90:3174-3187 */
static float other_contribution(float value, int index, float largest, float rest, int largest_index)
{
  return index == largest_index ? rest : largest + (rest - value);
}

/* add_largest, SRC/rt_patch_mod.f90:3150-3169 */
static void add_largest(float value, int index, float *largest, float *rest, int32_t *largest_index)
{
  if (fabsf(value) > fabsf(*largest)) {
    *rest = *rest + *largest;
    *largest = value;
    *largest_index = index;
  } else {
    *rest = *rest + value;
  }
}

void orc_bb_rates(const orc_bb_args *a, float *gamma_ij, float *gamma_ji)
{
  const int nc = a->ncell;
  const int split_pre = a->largest && a->precond;
  orc_qlg_init();
  float *vph = (float *)malloc(sizeof(float) * (a->nf > 0 ? a->nf : 1));
  for (int l = 0; l < a->nline; l++) {
    const int fs = a->iblue[l] - 1, fe = a->ired[l] - 1;
    const int li = a->lev_i[l] - 1, lj = a->lev_j[l] - 1;
    const float lower_off = a->lower_off[l];
    const int index = a->split_offset + l + 1;
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
      float eta_coef = 0, g_ratio = 0, abs_coef = 0;
      if (a->precond) {
        float split_bj = (float)a->b[(size_t)lj * nc + c];
        if (lower_off != 0) split_bj = split_bj * eoff;
        float split_Rij = (float)(nsi * a->pop_ratio[c]);
        split_Rij = (float)((double)split_Rij * a->Bij[l]);
        eta_coef = split_Rij * split_bj;
        float split_bi = (float)a->b[(size_t)li * nc + c];
        if (lower_off != 0) split_bi = split_bi * eoff;
        abs_coef = split_Rij * split_bi;
        g_ratio = 0.0f;
        if (nj > 0.0f) g_ratio = (float)((double)(ni / nj) * (a->g[lj] / a->g[li]));
      }
      float sA = 0, cA = 0, sI = 0, cI = 0, sP = 0, cP = 0, sw = 0, cw = 0;
      for (int f = fs; f <= fe; f++)
        vph[f] = orc_qlg_bot_profile(adamp, a->q[(size_t)l * a->nf + f], vpr, idf);
      if (split_pre) {
        int in_force = 0, raise = 0, pw = 0, pb = 0;
        if (a->use_flags) {
          pw = (a->ntrans[l] - 1) / 32;
          pb = (a->ntrans[l] - 1) % 32;
          in_force = (a->psi_flags_in[(size_t)pw * nc + c] >> pb) & 1;
        }
        for (int f = fs; f <= fe; f++) {
          const size_t o = (size_t)f * a->sf + (size_t)c * a->sc;
          float psi = a->Psi[o];
          const float eta_u = vph[f] * eta_coef;
          const float eta_t = eta_u * a->beta[o];
          const float chi_t = vph[f] * abs_coef - eta_t;
          const float eta_other = a->beta[o] * other_contribution(eta_u, index, a->split_opac1[o],
                                                                  a->split_opac2[o], a->split_id[o]);
          const float chi_other = other_contribution(vph[f] * abs_coef, index, a->split_chi1[o],
                                                     a->split_chi2[o], a->split_chi_id[o]) - eta_other;
          float other = eta_other;
          if (a->do_scat) other = other + a->chi_scat[o] * a->dJnu[o];
          if (chi_t < 0.0f || in_force) psi = 0.0f;
          raise = raise || psi == 0.0f;
          const float up = psi == 0.0f ? a->Inu[o] : a->Irem[o] + psi * other;
          const float spont = psi == 0.0f ? 1.0f : (chi_other + a->escape[o] * chi_t) / a->chi[o];
          const float wmunu = (float)((double)a->wmu * a->wnu[(size_t)l * a->nf + f] * a->Aij[l]);
          const float vvv = vph[f] * wmunu;
          const float nb = a->Bnu[o];
          kadd(&sA, &cA, vvv * spont * (nb * inv_nref), a->kahan);   /* sum_T */
          kadd(&sI, &cI, vvv * (up - nb * spont), a->kahan);         /* sum_U */
        }
        if (a->use_flags && raise && !in_force) a->psi_flags_out[(size_t)pw * nc + c] |= (int32_t)(1u << pb);
        if (a->own_norm)
          for (int f = fs; f <= fe; f++) kadd(&sw, &cw, vph[f] * (float)a->wq[(size_t)l * a->nf + f], a->kahan);
      } else {
        for (int f = fs; f <= fe; f++) {
          const size_t o = (size_t)f * a->sf + (size_t)c * a->sc;
          const float wmunu = (float)((double)a->wmu * a->wnu[(size_t)l * a->nf + f] * a->Aij[l]);
          const float phi = vph[f];
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
      }
      float g_th = sA, g_dev = sI;
      if (!a->largest) {
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
  free(vph);
}

void orc_bb_opacity(const orc_bbop_args *a)
{
  const int nc = a->ncell;
  orc_qlg_init();
  float *vph = (float *)malloc(sizeof(float) * (a->nf > 0 ? a->nf : 1));
  for (int l = 0; l < a->nline; l++) {
    const int fs = a->iblue[l] - 1, fe = a->ired[l] - 1;
    const int li = a->lev_i[l] - 1, lj = a->lev_j[l] - 1;
    const float lower_off = a->lower_off[l];
    const int index = a->split_offset + l + 1;
    const int sum_wphi = a->wphi_line && a->wphi_line[l];
    for (int c = 0; c < nc; c++) {
      const float idf = a->inv_df[c], ekt = a->e_kT[c];
      const float ni = (float)(a->nstar[(size_t)li * nc + c] * a->pop_ratio[c]);
      float bi = 1.0f, bj = 1.0f;
      if (!a->lte) {
        bi = (float)a->b[(size_t)li * nc + c];
        bj = (float)a->b[(size_t)lj * nc + c];
      }
      if (lower_off != 0) {
        const float eoff = (float)exp((double)(-lower_off * ekt));
        bi = bi * eoff;
        bj = bj * eoff;
      }
      for (int f = fs; f <= fe; f++)
        vph[f] = orc_qlg_bot_profile(a->adamp[(size_t)l * nc + c], a->q[(size_t)l * a->nf + f], a->vproj[c], idf);
      const float Rij = (float)((double)ni * a->Bij[l]);
      const float eta_fac = Rij * bj, chi_fac = Rij * bi;
      for (int f = fs; f <= fe; f++) {
        const size_t o = (size_t)f * a->sf + (size_t)c * a->sc;
        a->xhat[o] = a->xhat[o] + vph[f] * chi_fac;
        a->ehat[o] = a->ehat[o] + vph[f] * eta_fac;
      }
      if (a->split)
        for (int f = fs; f <= fe; f++) {
          const size_t o = (size_t)f * a->sf + (size_t)c * a->sc;
          add_largest(vph[f] * eta_fac, index, &a->split_opac1[o], &a->split_opac2[o], &a->split_id[o]);
          add_largest(vph[f] * chi_fac, index, &a->split_chi1[o], &a->split_chi2[o], &a->split_chi_id[o]);
        }
      if (sum_wphi) {
        double s = 0.0;  /* sum(phi * line%wq): real(4) times real(8), summed in real(8) */
        for (int f = fs; f <= fe; f++) s += (double)vph[f] * a->wq[(size_t)l * a->nf + f];
        a->wphi[(size_t)l * nc + c] = a->wphi[(size_t)l * nc + c] + a->wmu * (float)s;
      }
    }
  }
  free(vph);
}

void orc_bf_rates(const orc_bf_args *a, float *gamma_ij, float *gamma_ji)
{
  const int nc = a->ncell;
  float *wa = (float *)malloc(sizeof(float) * (a->nf > 0 ? a->nf : 1));
  for (int j = 0; j < a->ncont; j++) {
    if (a->active && !a->active[j]) continue;
    const int row = a->nline + j;
    const int fs = a->iblue[j] - 1, fe = a->ired[j] - 1;
    const int li = a->lev_i[j] - 1, lj = a->lev_j[j] - 1;
    for (int f = fs; f <= fe; f++) {
      const float wmunu = (float)(0.5 * a->wnu[(size_t)j * a->nf + f]);
      wa[f] = a->twohnu3_c2[f] * a->alpha[(size_t)j * a->nf + f] * wmunu;
    }
    const double dev = a->ev[lj] - a->ev[li];
    for (int c = 0; c < nc; c++) {
      if (dev * a->e_kT[c] > a->cut_x) continue;  /* radiative_cut, SRC/atom_mod.f90:1157-1162 */
      float sum_abs = 0, sum_ems = 0, sum_pre = 0;
      for (int f = fs; f <= fe; f++) {
        const size_t o = (size_t)f * a->sf + (size_t)c * a->sc;
        const float vij = wa[f];
        const float vji = vij * a->beta[o];
        const float phi = a->PsiJ[o] * a->alpha[(size_t)j * a->nf + f];
        sum_abs = sum_abs + vij * a->Jnu[o];
        sum_ems = sum_ems + vji * (a->Jnu[o] + 1.0f);
        sum_pre = sum_pre + vji * phi;
      }
      float g_ij = sum_abs, g_ji = sum_ems;
      if (a->precond) {
        const double nsi = a->nstar[(size_t)li * nc + c];
        g_ij = (float)((double)g_ij - (double)sum_pre * nsi * a->b[(size_t)lj * nc + c]);
        g_ji = (float)((double)g_ji - (double)sum_pre * nsi * a->b[(size_t)li * nc + c]);
      }
      gamma_ij[(size_t)row * nc + c] = gamma_ij[(size_t)row * nc + c] + g_ij;
      gamma_ji[(size_t)row * nc + c] = gamma_ji[(size_t)row * nc + c] + g_ji;
    }
  }
  free(wa);
}

long orc_work_to_transfer(size_t n, float *work1, float *work2, const float *beta, const float *so1,
                          const float *so2, const float *chi_scat, const float *dJnu, int split,
                          int scat, int suppress_masers, double maser_chi_floor)
{
  long nfl = 0;
  const float fl = (float)maser_chi_floor;
  for (size_t i = 0; i < n; i++) {
    float eta, chi;
    if (split) {
      const float eta_rest = beta[i] * so2[i];
      eta = beta[i] * so1[i] + eta_rest;
      chi = work1[i] - eta;
      if (scat) eta = beta[i] * so1[i] + (eta_rest + chi_scat[i] * dJnu[i]);
    } else {
      eta = work2[i] * beta[i];
      chi = work1[i] - eta;
      if (scat) eta = eta + chi_scat[i] * dJnu[i];
    }
    const float xhat = work1[i];
    float floor = fl * fabsf(xhat);
    if (floor < FLT_MIN) floor = FLT_MIN;       /* max(., tiny(1.0)) */
    if (suppress_masers) {
      if (chi < floor) chi = floor;
    } else if (fabsf(chi) < floor) {
      chi = copysignf(floor, chi);
      nfl++;
    }
    work1[i] = chi;
    work2[i] = eta / chi;
  }
  return nfl;
}

void orc_convert_to_work_units(size_t n, const float *chi_cont, float *xhat, float *ehat, const float *beta)
{
  for (size_t i = 0; i < n; i++) {
    xhat[i] = chi_cont[i] + ehat[i];
    ehat[i] = ehat[i] / beta[i];
  }
}

long orc_background_emission(size_t n, float *J, const float *B, const float *chi_cont, const float *sigma,
                             const float *PsiJ, const float *Jali_old, float *Jali, float *eta, float *dJnu,
                             int scat_ali, int j_to_b, float wscat, double scat_op_max, double *diag)
{
  long nneg = 0;
  double op_raw = 0, op_max = 0, amp_max = 0;
  if (scat_ali)
    for (size_t i = 0; i < n; i++) {
      double op = (double)sigma[i] * (double)PsiJ[i];
      if (op > op_raw) op_raw = op;
      op = op < scat_op_max ? op : scat_op_max;
      op = op > 0.0 ? op : 0.0;
      if (op > op_max) op_max = op;
      const double Jfs = J[i];
      double Jn = (Jfs - op * (double)Jali_old[i]) / (1.0 - op);
      if (Jn < 0.0) {
        nneg++;
        Jn = Jfs;
      }
      if (Jfs > 0.0 && Jn / Jfs > amp_max) amp_max = Jn / Jfs;
      J[i] = (float)Jn;
      if (Jali) Jali[i] = J[i];
    }
  for (size_t i = 0; i < n; i++) {
    float j = J[i];
    if (j_to_b) j = B[i];
    j = j - B[i];
    J[i] = j;
    eta[i] = chi_cont[i] * B[i] + wscat * sigma[i] * j;
    if (dJnu) dJnu[i] = wscat * j;
  }
  if (diag) {
    diag[0] = op_raw;
    diag[1] = op_max;
    diag[2] = amp_max;
  }
  return nneg;
}
