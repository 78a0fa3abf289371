// Line-rate integrals of one direction, bound_bound_rates
// (SRC/rt_patch_mod.f90:2852-3142): per line and cell, the frequency sums of the
// QLG-bot profile times the intensity, Bnu/n_ref and the Lambda* correction, split
// into the thermal (th) and deviation (dev) integrals that line_rate_coefficients
// turns into rates.  Formulations: 'original' (with the preconditioning
// correction) and 'largest' without preconditioning.
//
// Mapping.  One lane per cell; the line loop and each line's frequency window are
// uniform across the wavefront, so the per-frequency line data (abscissa q and
// quadrature weights) are scalar loads and the per-(f, cell) fields are one
// contiguous run per wavefront in the native layout.  QLG-bot's tables sit in LDS.
// The sums are FP32 with Kahan compensation.
#include "m3drt_internal.hpp"
#include <cmath>
#include <algorithm>

#pragma clang fp contract(off)

namespace m3drt {

namespace {

constexpr int QLG_NX = 9019;

struct RatesArgs {
  int nline, nf, ncell, nk, precond, own_norm, largest;
  const int32_t *iblue, *ired, *lev_i, *lev_j;
  const double *Aij, *Bij, *wnu, *wq, *g, *b, *nstar, *pop_ratio;
  const float *lower_off, *q, *inv_df, *vproj, *e_kT, *adamp;
  const float *Inu, *Psi, *Bnu, *beta;
  int64_t sf, sc;
  float wmu, wtot;
  float *gij, *gji;
  const float *tab;
};

struct Kahan {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

__global__ void __launch_bounds__(256) bb_rates_kernel(const RatesArgs a) {
  extern __shared__ float tab[];
  for (int i = threadIdx.x; i < 2 * (QLG_NX + 1); i += blockDim.x) tab[i] = a.tab[i];
  __syncthreads();
  const float *TGX = tab, *TH2X = tab + QLG_NX + 1;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  const int nc = a.ncell;
  const float idf = a.inv_df[c], vpr = a.vproj[c], ekt = a.e_kT[c];
  for (int l = 0; l < a.nline; ++l) {
    // a window or a level outside the arrays skips the line (documented in the header)
    const int fs = max(a.iblue[l] - 1, 0), fe = min(a.ired[l] - 1, a.nf - 1);
    const int li = a.lev_i[l] - 1, lj = a.lev_j[l] - 1;
    if (li < 0 || li >= a.nk || lj < 0 || lj >= a.nk) continue;
    const float lower_off = a.lower_off[l];
    const double nsi = a.nstar[(int64_t)li * nc + c], nsj = a.nstar[(int64_t)lj * nc + c];
    float ni = (float)(a.b[(int64_t)li * nc + c] * nsi);
    float nj = (float)(a.b[(int64_t)lj * nc + c] * nsj);
    float inv_nref = 0.0f;
    if (nsi > 0 && nsj > 0) {
      const double beta_ref = nsj * a.g[li] / (nsi * a.g[lj]);
      inv_nref = (float)fmin((1.0 - beta_ref) / beta_ref, 1e35);
    }
    float eta_coef = 0.0f, g_ratio = 0.0f;
    const float eoff = lower_off != 0 ? (float)exp((double)(-lower_off * ekt)) : 1.0f;
    if (lower_off != 0) {
      ni = ni * eoff;
      nj = nj * eoff;
    }
    if (a.precond) {
      float split_bj = (float)a.b[(int64_t)lj * nc + c];
      if (lower_off != 0) split_bj = split_bj * eoff;
      float split_Rij = (float)(nsi * a.pop_ratio[c]);
      split_Rij = (float)((double)split_Rij * a.Bij[l]);
      eta_coef = split_Rij * split_bj;
      if (nj > 0.0f) g_ratio = (float)((double)(ni / nj) * (a.g[lj] / a.g[li]));
    }
    // QLG-bot setup for damping a, on Doppler units (qlg_bot_sweep's, HOP/qlg_mod.f90:230-240)
    const float ad = a.adamp[(int64_t)l * nc + c];
    const float d = 1.0f + ad * 0.60602209212341318f;
    const float ih = d / (ad * d + 0.8325546112f);
    const float r = ad * ih, qq = r * (1.0f - r), scl = 1.7724538509f * ih, qs = qq * scl;
    const float x = 2.0f * r - 1.0f;
    const float bL = (0.3183098862f * r) * scl;
    const float gG = (0.4697186f * (1.0f - r)) * scl + qs * (-0.064924393191013452f + x * (-0.11019326278794868f));
    const float cZH = qs * (0.38494195930078284f + x * (-0.10567135279610407f));
    const float cL2 = qs * (0.04054415746208978f + x * (0.16241029253712097f));
    Kahan sA, sI, sP, sw;
    const int64_t lo = (int64_t)l * a.nf;
    for (int f = fs; f <= fe; ++f) {
      const float qf = a.q[lo + f];
      // qlg_bot_profile (HOP/qlg_mod.f90:664-671): profile per unit frequency
      const float v = fabsf(qf - vpr) * idf;
      const float xs = v * ih, u = xs * xs, L = 1.0f / (1.0f + u);
      const int ig = (int)roundf(fminf(xs * 800.0f, (float)QLG_NX));
      float p = cL2 + cZH * TH2X[ig];
      p = bL + L * p;
      const float K = TGX[ig] * gG + L * p;
      const float phi = K * 0.5641895835f * idf * (qf + 1.0f);
      const float wmunu = (float)((double)a.wmu * a.wnu[lo + f] * a.Aij[l]);
      const float vvv = phi * wmunu;
      const int64_t o = (int64_t)f * a.sf + (int64_t)c * a.sc;
      const float nb = a.Bnu[o], Iv = a.Inu[o];
      if (a.own_norm) sw.add(phi * (float)a.wq[lo + f]);
      sA.add(vvv * (nb * inv_nref));
      if (a.largest) {
        sI.add(vvv * (Iv - nb));
      } else {
        sI.add(vvv * Iv);
        sP.add(vvv * phi * a.Psi[o] * a.beta[o]);
      }
    }
    float g_th = sA.s, g_dev = sI.s;
    if (!a.largest) {
      if (a.precond) {
        g_th = g_th - sP.s * eta_coef * (g_ratio - 1.0f);
        g_dev = g_dev - sP.s * eta_coef;
      }
      if (inv_nref > 0.0f) g_dev = g_dev - g_th / inv_nref;
    }
    if (a.own_norm && sw.s > 0.0f) {
      g_th = g_th / (sw.s * a.wtot);
      g_dev = g_dev / (sw.s * a.wtot);
    }
    const int64_t go = (int64_t)l * nc + c;
    a.gij[go] = a.gij[go] + g_th;
    a.gji[go] = a.gji[go] + g_dev;
  }
}

}  // namespace

int bb_rates(m3drt_ctx *ctx, const m3drt_bb_rates_args *r) {
  int rc = qlg_tables(ctx);
  if (rc) return rc;
  RatesArgs a;
  a.nline = r->nline; a.nf = r->nf; a.ncell = r->ncell; a.nk = r->nk;
  a.precond = (r->flags & M3DRT_RATES_PRECOND) ? 1 : 0;
  a.own_norm = (r->flags & M3DRT_RATES_OWN_NORM) ? 1 : 0;
  a.largest = (r->flags & M3DRT_RATES_LARGEST) ? 1 : 0;
  a.iblue = r->iblue; a.ired = r->ired; a.lev_i = r->lev_i; a.lev_j = r->lev_j;
  a.Aij = r->Aij; a.Bij = r->Bij; a.wnu = r->wnu; a.wq = r->wq; a.g = r->g; a.b = r->b;
  a.nstar = r->nstar; a.pop_ratio = r->pop_ratio; a.lower_off = r->lower_off; a.q = r->q;
  a.inv_df = r->inv_df; a.vproj = r->vproj; a.e_kT = r->e_kT; a.adamp = r->adamp;
  a.Inu = r->Inu; a.Psi = r->Psi; a.Bnu = r->Bnu; a.beta = r->beta;
  a.sf = r->sf; a.sc = r->sc; a.wmu = r->wmu; a.wtot = r->wtot;
  a.gij = r->gamma_ij; a.gji = r->gamma_ji;
  a.tab = ctx->qlg_tab;
  const size_t lds = 2 * (QLG_NX + 1) * sizeof(float);
  M3DRT_HIP(hipFuncSetAttribute((const void *)bb_rates_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(bb_rates_kernel, dim3((r->ncell + 255) / 256), dim3(256), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
