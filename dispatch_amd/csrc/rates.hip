// The ALI line and continuum terms of one direction / one window, for gfx950:
//
//  bb_opacity_kernel  add_bound_bound_opacities (SRC/rt_patch_mod.f90:2689-2831): the
//                     model atom's lines into xhat/ehat, add_largest tracking
//                     (3150-3169), the direction's wphi share;
//  bb_rates_kernel    bound_bound_rates (2852-3142): 'original' with the
//                     preconditioning correction, 'largest' without it, and
//                     'largest' with it (Rybicki & Hummer integrands from Irem,
//                     escape and other_contribution, 3174-3187; psi_flags);
//  bf_rates_kernel    bound_free_rates (3193-3259).
//
// Mapping.  One lane per cell; the line (continuum) loop and each window's
// frequency loop are uniform across the wavefront, so the per-frequency line data
// are scalar loads and the per-(f, cell) fields are one contiguous run per
// wavefront in the native layout.  A cell's lines are visited in list order, which
// is what add_largest's order dependence needs; no two lanes touch one point.
// QLG-bot's tables sit in LDS.  Line-rate sums are FP32 with Kahan compensation.
#include "m3drt_internal.hpp"
#include <cfloat>
#include <cmath>
#include <algorithm>

#pragma clang fp contract(off)

namespace m3drt {

namespace {

constexpr int QLG_NX = 9019;

struct Kahan {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// QLG-bot on Doppler units for damping a (qlg_bot_sweep's setup, HOP/qlg_mod.f90:230-240)
// and the model-atom profile of qlg_bot_profile (HOP/qlg_mod.f90:682-689)
struct Qlg {
  float ih, bL, gG, cZH, cL2;
  __device__ __forceinline__ explicit Qlg(float ad) {
    const float d = 1.0f + ad * 0.60602209212341318f;
    ih = d / (ad * d + 0.8325546112f);
    const float r = ad * ih, qq = r * (1.0f - r), scl = 1.7724538509f * ih, qs = qq * scl;
    const float x = 2.0f * r - 1.0f;
    bL = (0.3183098862f * r) * scl;
    gG = (0.4697186f * (1.0f - r)) * scl + qs * (-0.064924393191013452f + x * (-0.11019326278794868f));
    cZH = qs * (0.38494195930078284f + x * (-0.10567135279610407f));
    cL2 = qs * (0.04054415746208978f + x * (0.16241029253712097f));
  }
  __device__ __forceinline__ float profile(const float *TGX, const float *TH2X, float qf, float vpr,
                                           float idf) const {
    const float v = fabsf(qf - vpr) * idf;
    const float xs = v * ih, u = xs * xs, L = 1.0f / (1.0f + u);
    const int ig = (int)roundf(fminf(xs * 800.0f, (float)QLG_NX));
    float p = cL2 + cZH * TH2X[ig];
    p = bL + L * p;
    const float K = TGX[ig] * gG + L * p;
    return K * 0.5641895835f * idf * (qf + 1.0f);
  }
};

__device__ __forceinline__ void load_tables(float *tab, const float *src) {
  for (int i = threadIdx.x; i < 2 * (QLG_NX + 1); i += blockDim.x) tab[i] = src[i];
  __syncthreads();
}

// other_contribution, SRC/rt_patch_mod.f90:3174-3187
__device__ __forceinline__ float other_contribution(float value, int index, float largest, float rest,
                                                    int largest_index) {
  return index == largest_index ? rest : largest + (rest - value);
}

// ---------------------------------------------------------------- line rates

struct RatesArgs {
  int nline, nf, ncell, nk, precond, own_norm, largest, scat, use_flags, split_offset;
  const int32_t *iblue, *ired, *lev_i, *lev_j;
  const double *Aij, *Bij, *wnu, *wq, *g, *b, *nstar, *pop_ratio;
  const float *lower_off, *q, *inv_df, *vproj, *e_kT, *adamp;
  const float *Inu, *Psi, *Bnu, *beta;
  const float *Irem, *esc, *chi, *so1, *so2, *sc1, *sc2, *chi_scat, *dJnu;
  const int32_t *sid, *scid, *ntrans, *flags_in;
  int32_t *flags_out;
  int64_t sf, sc;
  float wmu, wtot;
  float *gij, *gji;
  const float *tab;
  int lchunk, nchunk, ncp;  // a lane's run of lines, the runs per cell, cells per run padded to 256
};

// Lane (c, k): cell c, lines [k * lchunk, (k + 1) * lchunk), consecutive lanes on
// consecutive cells; each run's cells padded to whole workgroups, so k is the
// workgroup's (the line parameters stay scalar loads).  Each line's sums are its own
// (one lane per (line, cell)), so the runs change no result.
__global__ void __launch_bounds__(256) bb_rates_kernel(const RatesArgs a) {
  extern __shared__ float tab[];
  load_tables(tab, a.tab);
  const float *TGX = tab, *TH2X = tab + QLG_NX + 1;
  const int base = blockIdx.x * blockDim.x;  // the lane count fits int32 (checked at launch)
  const int k = base / a.ncp, c = base - k * a.ncp + threadIdx.x;
  if (k >= a.nchunk || c >= a.ncell) return;
  const int nc = a.ncell;
  const bool split_pre = a.largest && a.precond;
  const float idf = a.inv_df[c], vpr = a.vproj[c], ekt = a.e_kT[c];
  const int l_hi = min(a.nline, (k + 1) * a.lchunk);
  for (int l = k * a.lchunk; l < l_hi; ++l) {
    // a window or a level outside the arrays skips the line (documented in the header)
    const int fs = max(a.iblue[l] - 1, 0), fe = min(a.ired[l] - 1, a.nf - 1);
    const int li = a.lev_i[l] - 1, lj = a.lev_j[l] - 1;
    if (li < 0 || li >= a.nk || lj < 0 || lj >= a.nk) continue;
    const float lower_off = a.lower_off[l];
    const int index = a.split_offset + l + 1;
    const double nsi = a.nstar[(int64_t)li * nc + c], nsj = a.nstar[(int64_t)lj * nc + c];
    float ni = (float)(a.b[(int64_t)li * nc + c] * nsi);
    float nj = (float)(a.b[(int64_t)lj * nc + c] * nsj);
    float inv_nref = 0.0f;
    if (nsi > 0 && nsj > 0) {
      const double beta_ref = nsj * a.g[li] / (nsi * a.g[lj]);
      inv_nref = (float)fmin((1.0 - beta_ref) / beta_ref, 1e35);
    }
    float eta_coef = 0.0f, g_ratio = 0.0f, abs_coef = 0.0f;
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
      float split_bi = (float)a.b[(int64_t)li * nc + c];
      if (lower_off != 0) split_bi = split_bi * eoff;
      abs_coef = split_Rij * split_bi;
      if (nj > 0.0f) g_ratio = (float)((double)(ni / nj) * (a.g[lj] / a.g[li]));
    }
    const Qlg Q(a.adamp[(int64_t)l * nc + c]);
    Kahan sA, sI, sP, sw;
    const int64_t lo = (int64_t)l * a.nf;
    if (split_pre) {
      bool in_force = false, raise = false;
      int pw = 0, pb = 0;
      if (a.use_flags) {
        pw = (a.ntrans[l] - 1) / 32;
        pb = (a.ntrans[l] - 1) % 32;
        in_force = (a.flags_in[(int64_t)pw * nc + c] >> pb) & 1;
      }
      for (int f = fs; f <= fe; ++f) {
        // every read of the point first, unconditionally, so that they are in flight
        // together (reads under the psi == 0 choice would wait for Psi's); the
        // frequency's own values through the scalar cache
        const float qf = dev::ro(a.q)[lo + f];
        const double wnu = dev::ro(a.wnu)[lo + f];
        const double wqf = a.own_norm ? dev::ro(a.wq)[lo + f] : 0.0;
        const int64_t o = (int64_t)f * a.sf + (int64_t)c * a.sc;
        float psi = a.Psi[o];
        const float bt = a.beta[o], Iv = a.Inu[o], Ir = a.Irem[o], es = a.esc[o], chv = a.chi[o], nb = a.Bnu[o];
        const float s1 = a.so1[o], s2 = a.so2[o], c1 = a.sc1[o], c2 = a.sc2[o];
        const int32_t si = a.sid[o], ci = a.scid[o];
        const float sct = a.scat ? a.chi_scat[o] * a.dJnu[o] : 0.0f;
        const float vph = Q.profile(TGX, TH2X, qf, vpr, idf);
        const float eta_u = vph * eta_coef;
        const float eta_t = eta_u * bt;
        const float chi_t = vph * abs_coef - eta_t;
        const float eta_other = bt * other_contribution(eta_u, index, s1, s2, si);
        const float chi_other = other_contribution(vph * abs_coef, index, c1, c2, ci) - eta_other;
        float other = eta_other;
        if (a.scat) other = other + sct;
        if (chi_t < 0.0f || in_force) psi = 0.0f;
        raise = raise || psi == 0.0f;
        const bool zero = psi == 0.0f;
        const float up = zero ? Iv : Ir + psi * other;
        const float spont = zero ? 1.0f : (chi_other + es * chi_t) / chv;
        const float wmunu = (float)((double)a.wmu * wnu * a.Aij[l]);
        const float vvv = vph * wmunu;
        sA.add(vvv * spont * (nb * inv_nref));  // sum_T
        sI.add(vvv * (up - nb * spont));        // sum_U
        if (a.own_norm) sw.add(vph * (float)wqf);
      }
      if (a.use_flags && raise && !in_force) atomicOr(&a.flags_out[(int64_t)pw * nc + c], (int32_t)(1u << pb));
    } else {
      for (int f = fs; f <= fe; ++f) {
        const float qf = dev::ro(a.q)[lo + f];
        const float phi = Q.profile(TGX, TH2X, qf, vpr, idf);
        const float wmunu = (float)((double)a.wmu * dev::ro(a.wnu)[lo + f] * a.Aij[l]);
        const float vvv = phi * wmunu;
        const int64_t o = (int64_t)f * a.sf + (int64_t)c * a.sc;
        const float nb = a.Bnu[o], Iv = a.Inu[o];
        if (a.own_norm) sw.add(phi * (float)dev::ro(a.wq)[lo + f]);
        sA.add(vvv * (nb * inv_nref));
        if (a.largest) {
          sI.add(vvv * (Iv - nb));
        } else {
          sI.add(vvv * Iv);
          sP.add(vvv * phi * a.Psi[o] * a.beta[o]);
        }
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

// ---------------------------------------------------------------- line opacity

struct OpacArgs {
  int nline, nf, ncell, nk, lte, split, split_offset;
  const int32_t *iblue, *ired, *lev_i, *lev_j, *wphi_line;
  const double *Bij, *wq, *b, *nstar, *pop_ratio;
  const float *lower_off, *q, *inv_df, *vproj, *e_kT, *adamp;
  float *xhat, *ehat, *so1, *so2, *sc1, *sc2, *wphi;
  int32_t *sid, *scid;
  int64_t sf, sc;
  float wmu;
  const float *tab;
};

// add_largest, SRC/rt_patch_mod.f90:3150-3169
__device__ __forceinline__ void add_largest(float value, int index, float *largest, float *rest,
                                            int32_t *largest_index) {
  const float lg = *largest;
  if (fabsf(value) > fabsf(lg)) {
    *rest = *rest + lg;
    *largest = value;
    *largest_index = index;
  } else {
    *rest = *rest + value;
  }
}

// One line's frequencies of one cell.  The arrays are distinct (restrict): the
// compiler may then issue the reads of the next frequencies before this one's
// read-modify-writes, which otherwise order every iteration's loads after the
// previous iteration's stores.
template <bool SPLIT>
__device__ __forceinline__ double opacity_line(const Qlg &Q, const float *TGX, const float *TH2X,
                                               const float *__restrict__ q, const double *__restrict__ wq,
                                               bool sum_wphi, float vpr, float idf, float chi_fac, float eta_fac,
                                               int index, int fs, int fe, int64_t sf, int64_t co,
                                               float *__restrict__ xhat, float *__restrict__ ehat,
                                               float *__restrict__ so1, float *__restrict__ so2,
                                               int32_t *__restrict__ sid, float *__restrict__ sc1,
                                               float *__restrict__ sc2, int32_t *__restrict__ scid) {
  // the accumulations rounded product and sum apart, as the reference rounds them
#pragma clang fp contract(off)
  double ws = 0.0;
#pragma unroll 4
  for (int f = fs; f <= fe; ++f) {
    const float vph = Q.profile(TGX, TH2X, dev::ro(q)[f], vpr, idf);
    const int64_t o = (int64_t)f * sf + co;
    xhat[o] = xhat[o] + vph * chi_fac;
    ehat[o] = ehat[o] + vph * eta_fac;
    if (SPLIT) {
      add_largest(vph * eta_fac, index, &so1[o], &so2[o], &sid[o]);
      add_largest(vph * chi_fac, index, &sc1[o], &sc2[o], &scid[o]);
    }
    if (sum_wphi) ws += (double)vph * dev::ro(wq)[f];  // sum(phi * line%wq) in real(8)
  }
  return ws;
}

__global__ void __launch_bounds__(256) bb_opacity_kernel(const OpacArgs a) {
  extern __shared__ float tab[];
  load_tables(tab, a.tab);
  const float *TGX = tab, *TH2X = tab + QLG_NX + 1;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  const int nc = a.ncell;
  const float idf = a.inv_df[c], vpr = a.vproj[c], ekt = a.e_kT[c];
  for (int l = 0; l < a.nline; ++l) {
    const int fs = max(a.iblue[l] - 1, 0), fe = min(a.ired[l] - 1, a.nf - 1);
    const int li = a.lev_i[l] - 1, lj = a.lev_j[l] - 1;
    if (li < 0 || li >= a.nk || lj < 0 || lj >= a.nk) continue;
    const float lower_off = a.lower_off[l];
    const int index = a.split_offset + l + 1;
    const float ni = (float)(a.nstar[(int64_t)li * nc + c] * a.pop_ratio[c]);
    float bi = 1.0f, bj = 1.0f;
    if (!a.lte) {
      bi = (float)a.b[(int64_t)li * nc + c];
      bj = (float)a.b[(int64_t)lj * nc + c];
    }
    if (lower_off != 0) {
      const float eoff = (float)exp((double)(-lower_off * ekt));
      bi = bi * eoff;
      bj = bj * eoff;
    }
    const float Rij = (float)((double)ni * a.Bij[l]);
    const float eta_fac = Rij * bj, chi_fac = Rij * bi;
    const Qlg Q(a.adamp[(int64_t)l * nc + c]);
    const int64_t lo = (int64_t)l * a.nf;
    const bool sum_wphi = a.wphi_line && a.wphi_line[l];
    const int64_t co = (int64_t)c * a.sc;
    const double ws =
        a.split ? opacity_line<true>(Q, TGX, TH2X, a.q + lo, a.wq + lo, sum_wphi, vpr, idf, chi_fac, eta_fac, index,
                                     fs, fe, a.sf, co, a.xhat, a.ehat, a.so1, a.so2, a.sid, a.sc1, a.sc2, a.scid)
                : opacity_line<false>(Q, TGX, TH2X, a.q + lo, a.wq + lo, sum_wphi, vpr, idf, chi_fac, eta_fac,
                                      index, fs, fe, a.sf, co, a.xhat, a.ehat, nullptr, nullptr, nullptr, nullptr,
                                      nullptr, nullptr);
    if (sum_wphi) {
      const int64_t go = (int64_t)l * nc + c;
      a.wphi[go] = a.wphi[go] + a.wmu * (float)ws;
    }
  }
}

// ---------------------------------------------------------------- one window, fused

struct WinArgs {
  OpacArgs o;
  const float *cont, *Bnu, *beta;
  float floor_frac;
  int suppress;
  unsigned long long *n_floored;
  int fchunk, nchunk, ncp;  // a lane's run of frequencies, the runs per cell, cells per run padded to 256
};

// convert_to_work_units, the split tracking seeded with the continuum as contribution
// 0, add_bound_bound_opacities with add_largest, and work_to_transfer_units, for one
// cell's points of a window, frequency by frequency: each point's working values stay
// in registers from the continuum to (chi, S), and are written once.  The same
// operations in the same order as m3drt_convert_to_work_units, m3drt_bb_opacity
// (lines in index order at each point) and m3drt_work_to_transfer.  A line's per-cell
// factors are formed when the frequency loop enters its window (lines cover
// contiguous frequency ranges, so a lane re-forms them only where windows overlap).
// The frequency loop of one cell; the arrays are distinct (restrict), so the reads of
// the next points need not wait for this point's writes.
template <bool SPLIT>
__device__ __forceinline__ unsigned long long window_cell(
    const OpacArgs &a, const float *TGX, const float *TH2X, int c, int f_lo, int f_hi, float floor_frac, int suppress,
    const float *__restrict__ cont, const float *__restrict__ Bnu, const float *__restrict__ beta,
    float *__restrict__ xhat, float *__restrict__ ehat, float *__restrict__ so1, float *__restrict__ so2,
    int32_t *__restrict__ sidp, float *__restrict__ sc1, float *__restrict__ sc2, int32_t *__restrict__ scidp) {
  // every product and sum rounded on its own, as the separate kernels and the
  // reference round them
#pragma clang fp contract(off)
  const int nc = a.ncell;
  const float idf = a.inv_df[c], vpr = a.vproj[c], ekt = a.e_kT[c];
  unsigned long long floored = 0;
  int cur = -1;
  float chi_fac = 0.0f, eta_fac = 0.0f;
  Qlg Q(0.0f);
  // The lines covering the current frequency as a mask (windows of up to 64 lines),
  // re-formed only where some line's window starts or ends: a frequency loop that
  // otherwise scans every line of the window at every frequency.
  const bool masked = a.nline <= 64;
  uint64_t cover = 0;
  int f_next = 0;
  // one line's contribution at frequency f, lines in index order
  auto add_line = [&](int l, int f, float &x, float &e, float &s1, float &s2, int32_t &sid, float &c1, float &c2,
                      int32_t &cid) {
    const int li = dev::ro(a.lev_i)[l] - 1, lj = dev::ro(a.lev_j)[l] - 1;
    if (l != cur) {
      const float lower_off = dev::ro(a.lower_off)[l];
      const float ni = (float)(a.nstar[(int64_t)li * nc + c] * a.pop_ratio[c]);
      float bi = 1.0f, bj = 1.0f;
      if (!a.lte) {
        bi = (float)a.b[(int64_t)li * nc + c];
        bj = (float)a.b[(int64_t)lj * nc + c];
      }
      if (lower_off != 0) {
        const float eoff = (float)exp((double)(-lower_off * ekt));
        bi = bi * eoff;
        bj = bj * eoff;
      }
      const float Rij = (float)((double)ni * dev::ro(a.Bij)[l]);
      eta_fac = Rij * bj;
      chi_fac = Rij * bi;
      Q = Qlg(a.adamp[(int64_t)l * nc + c]);
      cur = l;
    }
    const float vph = Q.profile(TGX, TH2X, dev::ro(a.q)[(int64_t)l * a.nf + f], vpr, idf);
    const int index = a.split_offset + l + 1;
    x = x + vph * chi_fac;
    e = e + vph * eta_fac;
    if (SPLIT) {
      add_largest(vph * eta_fac, index, &s1, &s2, &sid);
      add_largest(vph * chi_fac, index, &c1, &c2, &cid);
    }
  };
  auto live = [&](int l, int f) {
    const int fs = max(dev::ro(a.iblue)[l] - 1, 0), fe = min(dev::ro(a.ired)[l] - 1, a.nf - 1);
    const int li = dev::ro(a.lev_i)[l] - 1, lj = dev::ro(a.lev_j)[l] - 1;
    return f >= fs && f <= fe && li >= 0 && li < a.nk && lj >= 0 && lj < a.nk;
  };
  // the next point's reads are issued before this point's writes (a deeper ring of
  // four points measured slower: 14.4 against 12.4 ms on the 432 000-cell window, its
  // registers costing a wave per SIMD)
  int64_t o = (int64_t)c * a.sc + (int64_t)f_lo * a.sf;
  float cc = cont[o], bn = Bnu[o], bt = beta[o];
  for (int f = f_lo; f < f_hi; ++f) {
    const int64_t on = o + a.sf;
    float ccn = 0.0f, bnn = 0.0f, btn = 0.0f;
    if (f + 1 < f_hi) {
      ccn = cont[on];
      bnn = Bnu[on];
      btn = beta[on];
    }
    if (masked && f >= f_next) {
      cover = 0;
      f_next = a.nf;
      for (int l = 0; l < a.nline; ++l) {
        const int fs = max(dev::ro(a.iblue)[l] - 1, 0), fe = min(dev::ro(a.ired)[l] - 1, a.nf - 1);
        if (fs > f) {
          f_next = min(f_next, fs);
        } else if (fe >= f) {
          f_next = min(f_next, fe + 1);
          if (live(l, f)) cover |= 1ull << l;
        }
      }
    }
    // the continuum (eta = chi B) in working units (convert_to_work_units)
    const float e0 = cc * bn;
    float x = cc + e0;
    float e = e0 / bt;
    float s1 = e, s2 = 0.0f, c1 = x, c2 = 0.0f;
    int32_t sid = 0, cid = 0;
    if (masked) {
      for (uint64_t m = cover; m; m &= m - 1) add_line((int)__builtin_ctzll(m), f, x, e, s1, s2, sid, c1, c2, cid);
    } else {
      for (int l = 0; l < a.nline; ++l)
        if (live(l, f)) add_line(l, f, x, e, s1, s2, sid, c1, c2, cid);
    }
    // work_to_transfer_units
    float eta, chi;
    if (SPLIT) {
      const float eta_rest = bt * s2;
      eta = bt * s1 + eta_rest;
    } else {
      eta = e * bt;
    }
    chi = x - eta;
    const float floor = fmaxf(floor_frac * fabsf(x), FLT_MIN);
    if (suppress) {
      chi = fmaxf(chi, floor);
    } else if (fabsf(chi) < floor) {
      chi = copysignf(floor, chi);
      ++floored;
    }
    xhat[o] = chi;
    ehat[o] = eta / chi;
    if (SPLIT) {
      so1[o] = s1;
      so2[o] = s2;
      sidp[o] = sid;
      sc1[o] = c1;
      sc2[o] = c2;
      scidp[o] = cid;
    }
    o = on;
    cc = ccn;
    bn = bnn;
    bt = btn;
  }
  return floored;
}

// convert_to_work_units, the split tracking seeded with the continuum as contribution
// 0, add_bound_bound_opacities with add_largest, and work_to_transfer_units, for one
// cell's points of a window, frequency by frequency: each point's working values stay
// in registers from the continuum to (chi, S), and are written once.  The same
// operations in the same order as m3drt_convert_to_work_units, m3drt_bb_opacity
// (lines in index order at each point) and m3drt_work_to_transfer.  A line's per-cell
// factors are formed when the frequency loop enters its window (lines cover
// contiguous frequency ranges, so a lane re-forms them only where windows overlap).
template <bool SPLIT>
__global__ void __launch_bounds__(256) window_opacity_kernel(const WinArgs w) {
  extern __shared__ float tab[];
  const OpacArgs &a = w.o;
  load_tables(tab, a.tab);
  const float *TGX = tab, *TH2X = tab + QLG_NX + 1;
  // lane (c, k): cell c, frequencies [k * fchunk, (k + 1) * fchunk); consecutive lanes
  // take consecutive cells of one run, each run's cells padded to whole workgroups so
  // that k is the workgroup's
  const int base = blockIdx.x * blockDim.x;  // the lane count fits int32 (checked at launch)
  const int k = base / w.ncp, c = base - k * w.ncp + threadIdx.x;
  if (k >= w.nchunk || c >= a.ncell) return;
  const int f_lo = k * w.fchunk, f_hi = min(a.nf, f_lo + w.fchunk);
  if (f_lo >= f_hi) return;
  const unsigned long long floored =
      window_cell<SPLIT>(a, TGX, TH2X, c, f_lo, f_hi, w.floor_frac, w.suppress, w.cont, w.Bnu, w.beta, a.xhat,
                         a.ehat, a.so1, a.so2, a.sid, a.sc1, a.sc2, a.scid);
  if (w.n_floored && floored) atomicAdd(w.n_floored, floored);
}

// ---------------------------------------------------------------- bound-free rates

struct BfArgs {
  int ncont, nline, nf, ncell, nk, precond;
  double cut_x;
  const int32_t *iblue, *ired, *lev_i, *lev_j, *active;
  const double *ev, *wnu, *e_kT, *nstar, *b;
  const float *alpha, *twohnu3_c2, *Jnu, *beta, *PsiJ;
  int64_t sf, sc;
  float *gij, *gji;
};

__global__ void __launch_bounds__(256) bf_rates_kernel(const BfArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  const int nc = a.ncell;
  const double ekt = a.e_kT[c];
  for (int j = 0; j < a.ncont; ++j) {
    if (a.active && !a.active[j]) continue;
    const int li = a.lev_i[j] - 1, lj = a.lev_j[j] - 1;
    if (li < 0 || li >= a.nk || lj < 0 || lj >= a.nk) continue;
    const double dev = a.ev[lj] - a.ev[li];
    if (dev * ekt > a.cut_x) continue;  // radiative_cut, SRC/atom_mod.f90:1157-1162
    const int fs = max(a.iblue[j] - 1, 0), fe = min(a.ired[j] - 1, a.nf - 1);
    const int64_t lo = (int64_t)j * a.nf;
    float sum_abs = 0.0f, sum_ems = 0.0f, sum_pre = 0.0f;
    for (int f = fs; f <= fe; ++f) {
      const float al = a.alpha[lo + f];
      const float wmunu = (float)(0.5 * a.wnu[lo + f]);
      const float vij = a.twohnu3_c2[f] * al * wmunu;
      const int64_t o = (int64_t)f * a.sf + (int64_t)c * a.sc;
      const float J = a.Jnu[o];
      const float vji = vij * a.beta[o];
      const float phi = a.PsiJ[o] * al;
      sum_abs = sum_abs + vij * J;
      sum_ems = sum_ems + vji * (J + 1.0f);
      sum_pre = sum_pre + vji * phi;
    }
    float g_ij = sum_abs, g_ji = sum_ems;
    if (a.precond) {
      const double nsi = a.nstar[(int64_t)li * nc + c];
      g_ij = (float)((double)g_ij - (double)sum_pre * nsi * a.b[(int64_t)lj * nc + c]);
      g_ji = (float)((double)g_ji - (double)sum_pre * nsi * a.b[(int64_t)li * nc + c]);
    }
    const int64_t go = (int64_t)(a.nline + j) * nc + c;
    a.gij[go] = a.gij[go] + g_ij;
    a.gji[go] = a.gji[go] + g_ji;
  }
}

template <typename K, typename A>
int launch_cells(m3drt_ctx *ctx, K kern, const A &a, int ncell, size_t lds) {
  if (lds) M3DRT_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3((ncell + 255) / 256), dim3(256), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
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
  a.scat = (r->flags & M3DRT_RATES_SCAT) ? 1 : 0;
  a.use_flags = (r->flags & M3DRT_RATES_FLAGS) ? 1 : 0;
  a.split_offset = r->split_offset;
  a.iblue = r->iblue; a.ired = r->ired; a.lev_i = r->lev_i; a.lev_j = r->lev_j;
  a.Aij = r->Aij; a.Bij = r->Bij; a.wnu = r->wnu; a.wq = r->wq; a.g = r->g; a.b = r->b;
  a.nstar = r->nstar; a.pop_ratio = r->pop_ratio; a.lower_off = r->lower_off; a.q = r->q;
  a.inv_df = r->inv_df; a.vproj = r->vproj; a.e_kT = r->e_kT; a.adamp = r->adamp;
  a.Inu = r->Inu; a.Psi = r->Psi; a.Bnu = r->Bnu; a.beta = r->beta;
  a.Irem = r->Irem; a.esc = r->escape; a.chi = r->chi;
  a.so1 = r->split_opac1; a.so2 = r->split_opac2; a.sc1 = r->split_chi1; a.sc2 = r->split_chi2;
  a.sid = r->split_id; a.scid = r->split_chi_id; a.chi_scat = r->chi_scat; a.dJnu = r->dJnu;
  a.ntrans = r->ntrans; a.flags_in = r->psi_flags_in; a.flags_out = r->psi_flags_out;
  a.sf = r->sf; a.sc = r->sc; a.wmu = r->wmu; a.wtot = r->wtot;
  a.gij = r->gamma_ij; a.gji = r->gamma_ji;
  a.tab = ctx->qlg_tab;
  // the lines in runs over enough lanes (about 2M) that the chip holds the grid
  // several times over (one lane per cell leaves, on the 432 000-cell window, 1688
  // workgroups against 1280 resident at this kernel's register use)
  const int64_t want = 1 << 21;
  int nchunk = (int)std::max<int64_t>(1, (want + r->ncell - 1) / std::max(1, r->ncell));
  nchunk = std::max(1, std::min(nchunk, r->nline));
  a.lchunk = (r->nline + nchunk - 1) / nchunk;
  a.nchunk = std::max(1, (r->nline + a.lchunk - 1) / std::max(1, a.lchunk));
  a.ncp = (r->ncell + 255) / 256 * 256;
  const int64_t lanes = (int64_t)a.ncp * a.nchunk;
  if (lanes > (int64_t)INT32_MAX) return set_error(M3DRT_ELIMIT, "m3drt_bb_rates: too many points");
  return launch_cells(ctx, bb_rates_kernel, a, (int)lanes, 2 * (QLG_NX + 1) * sizeof(float));
}

int bb_opacity(m3drt_ctx *ctx, const m3drt_bb_opacity_args *r) {
  int rc = qlg_tables(ctx);
  if (rc) return rc;
  OpacArgs a;
  a.nline = r->nline; a.nf = r->nf; a.ncell = r->ncell; a.nk = r->nk;
  a.lte = (r->flags & M3DRT_BBOP_LTE) ? 1 : 0;
  a.split = (r->flags & M3DRT_BBOP_SPLIT) ? 1 : 0;
  a.split_offset = r->split_offset;
  a.iblue = r->iblue; a.ired = r->ired; a.lev_i = r->lev_i; a.lev_j = r->lev_j; a.wphi_line = r->wphi_line;
  a.Bij = r->Bij; a.wq = r->wq; a.b = r->b; a.nstar = r->nstar; a.pop_ratio = r->pop_ratio;
  a.lower_off = r->lower_off; a.q = r->q; a.inv_df = r->inv_df; a.vproj = r->vproj; a.e_kT = r->e_kT;
  a.adamp = r->adamp;
  a.xhat = r->xhat; a.ehat = r->ehat;
  a.so1 = r->split_opac1; a.so2 = r->split_opac2; a.sc1 = r->split_chi1; a.sc2 = r->split_chi2;
  a.sid = r->split_id; a.scid = r->split_chi_id; a.wphi = r->wphi;
  a.sf = r->sf; a.sc = r->sc; a.wmu = r->wmu;
  a.tab = ctx->qlg_tab;
  return launch_cells(ctx, bb_opacity_kernel, a, r->ncell, 2 * (QLG_NX + 1) * sizeof(float));
}

int bf_rates(m3drt_ctx *ctx, const m3drt_bf_rates_args *r) {
  BfArgs a;
  a.ncont = r->ncont; a.nline = r->nline; a.nf = r->nf; a.ncell = r->ncell; a.nk = r->nk;
  a.precond = r->precond; a.cut_x = r->cut_x;
  a.iblue = r->iblue; a.ired = r->ired; a.lev_i = r->lev_i; a.lev_j = r->lev_j; a.active = r->active;
  a.ev = r->ev; a.wnu = r->wnu; a.e_kT = r->e_kT; a.nstar = r->nstar; a.b = r->b;
  a.alpha = r->alpha; a.twohnu3_c2 = r->twohnu3_c2; a.Jnu = r->Jnu; a.beta = r->beta; a.PsiJ = r->PsiJ;
  a.sf = r->sf; a.sc = r->sc;
  a.gij = r->gamma_ij; a.gji = r->gamma_ji;
  return launch_cells(ctx, bf_rates_kernel, a, r->ncell, 0);
}

int window_opacity(m3drt_ctx *ctx, const m3drt_bb_opacity_args *r, const m3drt_window_args *wa) {
  int rc = qlg_tables(ctx);
  if (rc) return rc;
  WinArgs w;
  OpacArgs &a = w.o;
  a.nline = r->nline; a.nf = r->nf; a.ncell = r->ncell; a.nk = r->nk;
  a.lte = (r->flags & M3DRT_BBOP_LTE) ? 1 : 0;
  a.split = (r->flags & M3DRT_BBOP_SPLIT) ? 1 : 0;
  a.split_offset = r->split_offset;
  a.iblue = r->iblue; a.ired = r->ired; a.lev_i = r->lev_i; a.lev_j = r->lev_j; a.wphi_line = nullptr;
  a.Bij = r->Bij; a.wq = nullptr; a.b = r->b; a.nstar = r->nstar; a.pop_ratio = r->pop_ratio;
  a.lower_off = r->lower_off; a.q = r->q; a.inv_df = r->inv_df; a.vproj = r->vproj; a.e_kT = r->e_kT;
  a.adamp = r->adamp;
  a.xhat = r->xhat; a.ehat = r->ehat;
  a.so1 = r->split_opac1; a.so2 = r->split_opac2; a.sc1 = r->split_chi1; a.sc2 = r->split_chi2;
  a.sid = r->split_id; a.scid = r->split_chi_id; a.wphi = nullptr;
  a.sf = r->sf; a.sc = r->sc; a.wmu = r->wmu;
  a.tab = ctx->qlg_tab;
  w.cont = wa->chi_cont; w.Bnu = wa->Bnu; w.beta = wa->beta;
  w.floor_frac = (float)wa->maser_chi_floor;
  w.suppress = (wa->flags & M3DRT_UNITS_SUPPRESS_MASERS) ? 1 : 0;
  w.n_floored = reinterpret_cast<unsigned long long *>(wa->n_floored);
  const size_t lds = 2 * (QLG_NX + 1) * sizeof(float);
  // Each cell's frequencies in runs of at least 64 over enough lanes (about 2M) that
  // the chip holds the grid several times over: one lane walking all of a cell's
  // frequencies leaves, on the 432 000-cell window, 1688 workgroups against 1536
  // resident at this kernel's register use, and the last 152 run a second full-length
  // round alone.  The points are independent, so the runs change no result.
  const int64_t want = 1 << 21;
  int nchunk = (int)std::max<int64_t>(1, (want + r->ncell - 1) / std::max(1, r->ncell));
  nchunk = std::max(1, std::min(nchunk, r->nf / 64));
  w.fchunk = (r->nf + nchunk - 1) / nchunk;
  w.nchunk = (r->nf + w.fchunk - 1) / w.fchunk;
  w.ncp = (r->ncell + 255) / 256 * 256;
  const int64_t lanes = (int64_t)w.ncp * w.nchunk;
  if (lanes > (int64_t)INT32_MAX) return set_error(M3DRT_ELIMIT, "m3drt_window_opacity: too many points");
  if (a.split) return launch_cells(ctx, window_opacity_kernel<true>, w, (int)lanes, lds);
  return launch_cells(ctx, window_opacity_kernel<false>, w, (int)lanes, lds);
}

}  // namespace m3drt
