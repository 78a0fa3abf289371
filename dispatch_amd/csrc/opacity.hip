// Line opacity: the summation of add_opacity_equi (HOP/linelist_mod.f90:729-960,
// an equidistant ln(lambda) grid) and of add_opacity (:469-719, a grid of any
// spacing) with the QLG-bot Voigt kernel (HOP/qlg_mod.f90:223-256, 524-653), for
// gfx950.  The two share the line loop and its tolerance logic; they differ in how
// a line centre finds its node (index arithmetic, or a forward cursor taking the
// nearer node), the abscissa of the profile (grid steps, or ln(lambda) itself) and
// the window (nearest node +- nint(half-width), or bisection).
//
// Mapping.  Each cell walks its line list in order, because whether a line is
// kept depends on the opacity the earlier lines already deposited at its centre
// (the running tolerance floor, HOP/linelist_mod.f90:853-870).  So one lane owns
// one cell and the line loop is uniform across a wavefront: lanes are consecutive
// cells, and with cells at unit stride (the native layout) every deposit of a
// wavefront at one frequency is one contiguous run.  QLG-bot's two 9020-entry
// tables (72 KB) are staged in LDS once per workgroup.  Up to four lines that pass
// are deposited together, eight frequencies at a time (one read-modify-write of
// the spectrum per frequency for the chunk), with every decision still taken on the
// opacity the sequential order would have at that point.
//
// Arithmetic.  FP32 as the reference, each operation separately rounded (no fused
// multiply-add) so that the window edges and the snapped table index, integer
// decisions taken from floating-point values, come out as the reference's; exp,
// log and pow are taken in FP64 and rounded once.  The reciprocal 1/(1+u) of the
// per-point evaluation uses the hardware reciprocal (1 ulp).
#include "m3drt_internal.hpp"
#include <cmath>
#include <vector>
#include <algorithm>

#pragma clang fp contract(off)

namespace m3drt {

namespace {

constexpr int QLG_NX = 9019;

struct OpacArgs {
  int nline;
  const double *llam0;
  const float *lower_ev, *gpref, *grad, *gw1, *gw2, *gstk;
  const int32_t *unsold, *lev_i, *lev_j, *ipart;
  int32_t *used;
  float *strength, *reach;
  const float *reach_in;  // the list's reaches (has_reach), or NULL
  const int32_t *sel;     // the selection: sel[0] lines, sel[1..] their 0-based indices; NULL: all
  int nf, ncell, npart, nlev, nlte, scat;
  const double *llam;
  const float *vproj, *idf, *vb2, *ekt, *ne, *tothi, *spe, *lnbase, *dep;
  float *chi, *eta, *csc;
  const float *cont, *logref, *ome;
  int64_t sf, sc;
  float tol, rej;
  int lref_from_cont;  // logref = lt + ln cont at the nodes used, -FLT_MAX where cont <= 0
  float lt;            // ln(tol / voigt_peak)
  const float *tgx, *th2x;
};

struct Coef {
  float ih, bL, gG, cZH, cL2;
};

// a line that passed, waiting to be deposited: its profile constants, its window
// fa..fb (1-based), its abscissa (the node-fa offset x0 in grid steps on an
// equidistant grid, the shifted centre llc in ln(lambda) on any other), its
// departure factors and whether it also feeds chi_scat (plain: it does not; always
// plain without the scattering split)
struct Pend {
  Coef c;
  double llc;
  float x0, bi, bj;
  int fa, fb;
  bool plain;
};

// bot_setup (HOP/qlg_mod.f90:524-540): QLG-bot's per-line constants, the paper's fit
__device__ __forceinline__ Coef bot_setup(float gL, float D, float SQ, float S) {
  Coef c;
  const float e = D + gL * 0.60602209212341318f;
  c.ih = e / (gL * e + SQ);
  const float r = gL * c.ih;
  const float q = r * (1.0f - r);
  const float sc = S * c.ih;
  const float qs = q * sc;
  const float x = 2.0f * r - 1.0f;
  c.bL = (0.3183098862f * r) * sc;
  c.gG = (0.4697186f * (1.0f - r)) * sc + qs * (-0.064924393191013452f + x * (-0.11019326278794868f));
  c.cZH = qs * (0.38494195930078284f + x * (-0.10567135279610407f));
  c.cL2 = qs * (0.04054415746208978f + x * (0.16241029253712097f));
  return c;
}

__device__ __forceinline__ int nint(float v) { return (int)roundf(v); }
__device__ __forceinline__ int nint(double v) { return (int)round(v); }

// atomic max of non-negative floats through their ordered bit patterns
__device__ __forceinline__ void atomic_max_pos(float *p, float v) {
  atomicMax(reinterpret_cast<int *>(p), __float_as_int(v));
}

// bisect_nearest (HOP/linelist_mod.f90:1075-1095), 1-based: the first node of
// llam(ws:we) at or above x, one down when the node below is strictly nearer
__device__ __forceinline__ int bisect_nearest(const double *llam, int ws, int we, double x) {
  int lo = ws, hi = we;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (llam[mid - 1] < x) lo = mid + 1;
    else hi = mid;
  }
  if (lo > ws && fabs(llam[lo - 2] - x) < fabs(llam[lo - 1] - x)) lo = lo - 1;
  return lo;
}

// select_lines (HOP/linelist_mod.f90:1110-1142) in one workgroup: max |vproj| over
// the call's cells and the list's largest reach (block maxima), the slice of line
// centres within blue - reach_max .. red + reach_max (lower_bound8 / upper_bound8,
// :1807-1844, on thread 0), then the active lines of the slice whose own reach
// window touches blue..red, compacted in list order.  sel[0] = their count.
__global__ void __launch_bounds__(1024) select_lines_kernel(const double *__restrict__ llam0,
                                                            const float *__restrict__ reach_in,
                                                            const int32_t *__restrict__ active, int nline,
                                                            const double *__restrict__ llam, int nf,
                                                            const float *__restrict__ vproj, int ncell,
                                                            int32_t *__restrict__ sel) {
  __shared__ float red_v[1024], red_r[1024];
  __shared__ int scan[1024];
  __shared__ int l0s, l1s, base;
  __shared__ double blue_s, red_s;
  const int tid = threadIdx.x, nt = blockDim.x;
  float v = 0.0f, r = -3.40282347e38f;
  if (reach_in) {
    for (int c = tid; c < ncell; c += nt) v = fmaxf(v, fabsf(vproj[c]));
    for (int l = tid; l < nline; l += nt) r = fmaxf(r, reach_in[l]);
  }
  red_v[tid] = v;
  red_r[tid] = r;
  __syncthreads();
  for (int h = nt / 2; h > 0; h /= 2) {
    if (tid < h) {
      red_v[tid] = fmaxf(red_v[tid], red_v[tid + h]);
      red_r[tid] = fmaxf(red_r[tid], red_r[tid + h]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    int l0 = 1, l1 = nline;
    double blue = 0.0, redw = 0.0;
    if (reach_in) {
      double hb0 = 0.0, hb1 = 0.0;
      if (nf > 1) {
        hb0 = 0.5 * (llam[1] - llam[0]);
        hb1 = 0.5 * (llam[nf - 1] - llam[nf - 2]);
      }
      const double vmax = (double)red_v[0];
      blue = llam[0] - hb0 - vmax;
      redw = llam[nf - 1] + hb1 + vmax;
      const double rmax = (double)red_r[0];
      // lower_bound8: first centre >= blue - rmax; upper_bound8: last <= red + rmax
      int lo = 1, hi = nline + 1;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (llam0[mid - 1] >= blue - rmax) hi = mid;
        else lo = mid + 1;
      }
      l0 = max(l0, lo);
      lo = 1;
      hi = nline + 1;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (llam0[mid - 1] > redw + rmax) hi = mid;
        else lo = mid + 1;
      }
      l1 = min(l1, lo - 1);
    }
    l0s = l0;
    l1s = l1;
    blue_s = blue;
    red_s = redw;
    base = 0;
  }
  __syncthreads();
  const int l0 = l0s, l1 = l1s;
  for (int start = l0; start <= l1; start += nt) {
    const int l = start + tid;  // 1-based
    int keep = 0;
    if (l <= l1) {
      keep = !active || active[l - 1] != 0;
      if (keep && reach_in) {
        const double rr = (double)reach_in[l - 1];
        keep = !(llam0[l - 1] + rr < blue_s || llam0[l - 1] - rr > red_s);
      }
    }
    scan[tid] = keep;
    __syncthreads();
    for (int h = 1; h < nt; h *= 2) {  // inclusive scan
      const int x = tid >= h ? scan[tid - h] : 0;
      __syncthreads();
      scan[tid] += x;
      __syncthreads();
    }
    if (keep) sel[1 + base + scan[tid] - 1] = l - 1;
    __syncthreads();
    if (tid == nt - 1) base += scan[nt - 1];
    __syncthreads();
  }
  if (tid == 0) sel[0] = base;
}

// qlg_bot_line_eval (HOP/qlg_mod.f90:542-556) at abscissa x
__device__ __forceinline__ float bot_eval(const Coef &c, float x, const float *TGX, const float *TH2X) {
  const float xs = fabsf(x) * c.ih;
  const float u = xs * xs;
  const float L = __builtin_amdgcn_rcpf(1.0f + u);
  const int ig = nint(fminf(xs * 800.0f, (float)QLG_NX));
  const float G = TGX[ig], W = TH2X[ig];
  float p = c.cL2 + c.cZH * W;
  p = c.bL + L * p;
  return G * c.gG + L * p;
}

// DIRECT: chi and eta are the working opacity itself, for a list of at most K lines
// alone in its composition (one deposit per cell, after every read of cont): each
// point's line sum is formed from zero in line order and added to the working value
// at the deposit, as the increments of the composition would be added.  SCAT: the
// call splits off the scattering part (a line then feeds chi_scat unless its
// departure factors are given); without it every line is plain, known at compile
// time, and the scattering path leaves the per-point code.  The direct form never
// takes the split.
template <bool GENERAL, bool DIRECT, bool SCAT>
__global__ void __launch_bounds__(256) opacity_kernel(const OpacArgs a) {
  static_assert(!(DIRECT && SCAT), "the direct form takes no scattering split");
  extern __shared__ float tab[];  // TGX then TH2X
  for (int i = threadIdx.x; i <= QLG_NX; i += blockDim.x) {
    tab[i] = a.tgx[i];
    tab[QLG_NX + 1 + i] = a.th2x[i];
  }
  __syncthreads();
  const float *TGX = tab, *TH2X = tab + QLG_NX + 1;
  const int cidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (cidx >= a.ncell) return;

  const float sqrt_pi = 1.7724538509f, ex_cutoff = -85.0f, voigt_peak = 0.5641895835f;
  const int nf = a.nf;
  const double llam1 = a.llam[0];
  const double dllam = (a.llam[nf - 1] - a.llam[0]) / (double)(nf - 1);
  const double inv_dllam = 1.0 / dllam;
  const float sdl = (float)inv_dllam;

  float *chi = a.chi + cidx * a.sc, *eta = a.eta + cidx * a.sc;
  float *csc_c = a.csc ? a.csc + cidx * a.sc : nullptr;
  const float *ome_c = a.ome ? a.ome + cidx * a.sc : nullptr;
  const float *cont = a.cont + cidx * a.sc, *logref = a.logref + cidx * a.sc;
  const int64_t sf = a.sf;

  // the cell's Doppler width (in grid steps, or in ln(lambda) on a general grid) and
  // its cell-invariant terms (qlg_bot_cell)
  const float idf = a.idf[cidx];
  const float D = GENERAL ? 1.0f / idf : sdl / idf, D2 = D * D, SQ = 0.8325546112f * D2;
  const float D2h = 0.5f * D2, D2G = 1.5f * D2, D2L = 5.0f * D2;
  const float vb2 = a.vb2[cidx];
  const float unsold_c = (float)pow((double)vb2, (double)0.3f) * a.tothi[cidx];
  const float lvb2 = (float)log((double)vb2);
  const float ekt = a.ekt[cidx], ne_c = a.ne[cidx], tothi_c = a.tothi[cidx];
  const float starkpe = a.spe[cidx], vproj_c = a.vproj[cidx];

  // Lines that pass are kept pending, up to K, and deposited together: one batched
  // read-modify-write of the spectrum per frequency for the chunk instead of one per
  // line, the sums at each frequency still in line order.  A new line's running floor
  // needs chi at its centre node as the sequential order has it there: the global
  // value (every earlier chunk deposited) plus the pending lines' deposits at that
  // node, added in line order with the same FP32 operations.
  constexpr int K = 4;
  Pend pend[K];
  int np = 0, fc = 1;
  const double *llam = a.llam;
  auto val = [&](const Pend &q, int f) -> float {
    const float x = GENERAL ? (float)(llam[f - 1] - q.llc) : q.x0 + (float)(f - q.fa);
    return bot_eval(q.c, x, TGX, TH2X);
  };
  auto flush = [&]() {
    if (np == 0) return;
    int Fa = nf + 1, Fb = 0;
    bool anyscat = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k < np) {
        Fa = min(Fa, pend[k].fa);
        Fb = max(Fb, pend[k].fb);
        anyscat = anyscat || !(!SCAT || pend[k].plain);
      }
    }
    constexpr int U = 8;
    for (int f = Fa; f <= Fb; f += U) {
      bool any = false;
#pragma unroll
      for (int k = 0; k < K; ++k) any = any || (k < np && pend[k].fa <= f + U - 1 && pend[k].fb >= f);
      if (!any) continue;
      float cv[U], ev[U], sv[U], ov[U], cb[DIRECT ? U : 1], eb[DIRECT ? U : 1];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        if (f + j <= Fb) {
          const int64_t o = (int64_t)(f + j - 1) * sf;
          if (DIRECT) {
            cb[j] = chi[o];
            eb[j] = eta[o];
            cv[j] = 0.0f;
            ev[j] = 0.0f;
          } else {
            cv[j] = chi[o];
            ev[j] = eta[o];
          }
          if (anyscat) {
            sv[j] = csc_c[o];
            ov[j] = ome_c[o];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (k < np) {
#pragma unroll
          for (int j = 0; j < U; ++j) {
            const int g = f + j;
            if (g >= pend[k].fa && g <= pend[k].fb) {
              const float lte = val(pend[k], g);
              if (!SCAT || pend[k].plain) {
                cv[j] = cv[j] + lte * pend[k].bi;
                ev[j] = ev[j] + lte * pend[k].bj;
              } else {
                cv[j] = cv[j] + lte;
                ev[j] = ev[j] + lte;
                sv[j] = sv[j] + ov[j] * lte;
              }
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < U; ++j) {
        if (f + j <= Fb) {
          const int64_t o = (int64_t)(f + j - 1) * sf;
          if (DIRECT) {
            chi[o] = cb[j] + cv[j];
            eta[o] = eb[j] + ev[j];
          } else {
            chi[o] = cv[j];
            eta[o] = ev[j];
          }
          if (anyscat) csc_c[o] = sv[j];
        }
      }
    }
    np = 0;
  };
  auto queue = [&](const Pend &q) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (k == np) pend[k] = q;
    if (++np == K) flush();
  };

  auto lref_at = [&](int f) -> float {
    if (!a.lref_from_cont) return logref[(f - 1) * sf];
    const float x = cont[(f - 1) * sf];
    return x > 0.0f ? a.lt + (float)log((double)x) : -3.40282347e38f;
  };
  const int nsel = a.sel ? a.sel[0] : a.nline;
  for (int is = 0; is < nsel; ++is) {
    const int l = a.sel ? a.sel[1 + is] : is;
    const double llc = a.llam0[l] - (double)vproj_c;
    const double pos = (llc - llam1) * inv_dllam;
    int f0;
    if (GENERAL) {
      // the cursor only advances: the lines ascend and vproj is the cell's
      // (HOP/linelist_mod.f90:578-590)
      while (fc < nf && !(a.llam[fc - 1] >= llc)) ++fc;
      f0 = fc > 1 ? (fabs(a.llam[fc - 1] - llc) < fabs(a.llam[fc - 2] - llc) ? fc : fc - 1) : 1;
    } else {
      f0 = min(nf, max(1, nint(pos) + 1));
    }
    const float ex = a.lower_ev[l] * ekt;
    const int ip = a.ipart ? a.ipart[l] : 1;
    const float arg = a.lnbase[(int64_t)(ip - 1) * a.ncell + cidx] + a.gpref[l] - ex;
    if (arg < ex_cutoff) continue;
    // a line centred beyond the grid, in a list with reaches: the reference is the
    // continuum at the nearer grid end, without the running floor (:594-602, 844-851)
    const bool outside = a.reach_in && (llc < a.llam[0] || llc > a.llam[nf - 1]);
    float Rij, ref, lnq;
    if (outside) {
      const int fe = llc < a.llam[0] ? 1 : nf;
      const float lref_o = lref_at(fe);
      if (arg <= lref_o) continue;
      Rij = (float)exp((double)arg);
      ref = a.tol * cont[(fe - 1) * sf];
      lnq = arg - lref_o;
    } else {
      // the tolerance reference: tol * continuum, floored by tol^2 * the opacity
      // already at the centre node (:853-870; the reference caches it per node until
      // a line deposits, which is the same value)
      float chi_f0 = DIRECT ? 0.0f : chi[(f0 - 1) * sf];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (k < np && pend[k].fa <= f0 && f0 <= pend[k].fb) {
          const float v = val(pend[k], f0);
          chi_f0 = (!SCAT || pend[k].plain) ? chi_f0 + v * pend[k].bi : chi_f0 + v;
        }
      }
      float refl_c = a.tol * cont[(f0 - 1) * sf], lref_c;
      const float flo = fmaxf(a.tol, a.rej) * a.tol * chi_f0;
      if (flo > refl_c) {
        refl_c = flo;
        lref_c = (float)log((double)(flo * sqrt_pi));
      } else {
        lref_c = lref_at(f0);
      }
      if (arg <= lref_c) continue;
      Rij = (float)exp((double)arg);
      ref = refl_c;
      if (voigt_peak * Rij <= ref) continue;
      lnq = arg - lref_c;
    }

    const bool nlte = a.nlte && a.lev_i[l] > 0 && a.lev_j[l] > 0;
    float bi = 1.0f, bj = 1.0f;
    if (nlte) {
      bi = a.dep[(int64_t)(a.lev_i[l] - 1) * a.ncell + cidx];
      bj = a.dep[(int64_t)(a.lev_j[l] - 1) * a.ncell + cidx];
    }
    // damping, physical then in grid steps (HOP/linelist_mod.f90:895-906)
    float gam = a.grad[l];
    if (a.unsold[l])
      gam = gam + a.gw1[l] * unsold_c;
    else
      gam = gam + a.gw1[l] * (float)exp((double)(a.gw2[l] * lvb2)) * tothi_c;
    if (a.gstk[l] >= 0.0f)
      gam = gam + a.gstk[l] * ne_c;
    else
      gam = gam - a.gstk[l] * starkpe;
    if (!GENERAL) gam = gam * sdl;
    if (!outside && a.strength) atomic_max_pos(&a.strength[l], voigt_peak * Rij / cont[(f0 - 1) * sf]);
    Pend q;
    q.bi = bi;
    q.bj = bj;
    q.plain = !(a.scat && !nlte);
    q.llc = llc;
    q.x0 = 0.0f;
    if (GENERAL) {
      // qlg_bot_line on the ln(lambda) abscissa; the window by bisection unless the
      // line covers the grid (HOP/linelist_mod.f90:660-705); a line beyond the grid
      // keeps the part of its window (capped by its reach) that reaches the grid
      q.c = bot_setup(gam, D, SQ, Rij * D);
      const float AL = 0.3183098862f * gam * (Rij * D) / ref;
      const float hw = sqrtf(fmaxf(fmaxf(D2 * lnq + fminf(AL, D2G),
                                         AL - fmaxf(gam * gam - D2h, 0.0f) + fminf(AL, D2L)), 0.0f));
      q.fa = 1;
      q.fb = nf;
      if (outside) {
        if (a.reach) atomic_max_pos(&a.reach[l], hw);
        const double dl = fmin((double)hw, (double)a.reach_in[l]);
        double hb0 = 0.0, hb1 = 0.0;  // select_lines' half steps at the grid ends
        if (nf > 1) {
          hb0 = 0.5 * (a.llam[1] - a.llam[0]);
          hb1 = 0.5 * (a.llam[nf - 1] - a.llam[nf - 2]);
        }
        if (llc + dl < a.llam[0] - hb0 || llc - dl > a.llam[nf - 1] + hb1) continue;
        if (a.used) a.used[l] = 1;
        q.fa = bisect_nearest(a.llam, 1, nf, llc - dl);
        q.fb = bisect_nearest(a.llam, 1, nf, llc + dl);
      } else {
        if (a.used) a.used[l] = 1;
        if (a.reach) atomic_max_pos(&a.reach[l], hw);
        const double dl = (double)hw;
        if (!(dl >= fmax(a.llam[nf - 1] - llc, llc - a.llam[0]))) {
          q.fa = bisect_nearest(a.llam, 1, f0, llc - dl);
          q.fb = bisect_nearest(a.llam, f0, nf, llc + dl);
        }
      }
      if (q.fa > q.fb) continue;
      queue(q);
      continue;
    }

    // qlg_bot_line_sweep (HOP/qlg_mod.f90:630-653): setup, window, sweep
    const float S = Rij * D;
    const int ic = nint(pos) + 1;
    const int i1 = min(nf, max(1, ic));
    const float x1 = (float)((a.llam[i1 - 1] - llc) * inv_dllam);
    q.c = bot_setup(gam, D, SQ, S);
    const float AL = 0.3183098862f * gam * S / ref;
    const float hw = sqrtf(fmaxf(fmaxf(D2 * lnq + fminf(AL, D2G),
                                       AL - fmaxf(gam * gam - D2h, 0.0f) + fminf(AL, D2L)), 0.0f));
    if (a.reach) atomic_max_pos(&a.reach[l], (float)((double)hw * dllam));
    // the window of a line beyond the grid is capped by its reach (:914-921)
    const float cap = outside ? (float)((double)a.reach_in[l] * inv_dllam) : 3.40282347e38f;
    const int noff = nint(fminf(fminf(hw, cap), (float)(nf + abs(ic))));
    q.fa = max(1, ic - noff);
    q.fb = min(nf, ic + noff);
    if (q.fa > q.fb) continue;
    if (a.used) a.used[l] = 1;
    q.x0 = x1 - (float)(i1 - q.fa);
    queue(q);
  }
  flush();
}

// ------------------------------------------------------------------ tables
// QLG-bot's tables (HOP/qlg_mod.f90:364-375, 382-395): 2^-u and Z_0.2(u)/L^2 on
// u_i = (i/800)^2, built in FP64 and rounded once.  Z_0.2 needs an FP64 Voigt
// function: Weideman's (1994) N = 32 rational approximation, its coefficients
// from the paper's DFT recipe, with Humlicek's region-I asymptote beyond
// |v| + a = 15.

struct W32 {
  double a[32];
  double L;
};

static W32 weideman32() {
  W32 w;
  const int N = 32, M = 2 * N, M2 = 2 * M;
  const double PI = 3.14159265358979323846;
  w.L = std::sqrt(N / std::sqrt(2.0));
  std::vector<double> g(M2, 0.0), f(M2);
  for (int k = -M + 1; k <= M - 1; k++) {
    const double t = w.L * std::tan(k * PI / M / 2);
    g[k + M] = std::exp(-t * t) * (w.L * w.L + t * t);
  }
  for (int i = 0; i < M2; i++) f[i] = g[(i + M) % M2];
  for (int j = 1; j <= N; j++) {
    double re = 0.0;
    for (int i = 0; i < M2; i++) re += f[i] * std::cos(2 * PI * (double)i * j / M2);
    w.a[N - j] = re / M2;
  }
  return w;
}

static double voigt_f64(const W32 &w, double a, double v) {
  const double ISQPI = 0.564189583547756287;
  const double av = std::fabs(v);
  if (av + a >= 15.0) {
    const double tr = a * a - av * av + 0.5, ti = -2.0 * a * av;
    return ISQPI * ((a * tr - av * ti) / (tr * tr + ti * ti));
  }
  const double dr = w.L + a, nr0 = w.L - a, di = -v;
  const double q = 1.0 / (dr * dr + di * di);
  const double zr = (nr0 * dr + v * di) * q, zi = (v * dr - nr0 * di) * q;
  double pr = w.a[0], pim = 0.0;
  for (int j = 1; j < 32; j++) {
    const double t = pr * zr - pim * zi + w.a[j];
    pim = pr * zi + pim * zr;
    pr = t;
  }
  const double ir = dr * q, ii = -di * q;
  return 2.0 * (pr * (ir * ir - ii * ii) - pim * (2.0 * ir * ii)) + ISQPI * ir;
}

}  // namespace

int qlg_tables(m3drt_ctx *ctx) {
  if (ctx->qlg_tab) return M3DRT_OK;
  const W32 w = weideman32();
  std::vector<float> tab(2 * (QLG_NX + 1));
  // anchor of the Z_0.2 slice: damping a* with a*/Gamma(a*) = 0.2, its half-width,
  // peak erfcx(a*) and Lorentz share, as the fitted kernel defines them
  const double A2 = 0.18741813853432507, GAM = 0.93709069267162530, KP = 0.81924853602069690;
  const double ETA = 0.14697983683153040, QRS = 0.16, LN2 = 0.69314718055994531;
  for (int i = 0; i <= QLG_NX; i++) {
    const double x = (double)i / 800.0, u = x * x;
    tab[i] = (float)std::pow(2.0, -u);
    const double L = 1.0 / (1.0 + u), G = std::exp(-LN2 * u);
    const double K = voigt_f64(w, A2, GAM * std::sqrt(u));
    const double P = KP * (ETA * L + (1.0 - ETA) * G);
    tab[QLG_NX + 1 + i] = (float)((K - P) / (QRS * KP) * ((1.0 + u) * (1.0 + u)));
  }
  void *d = nullptr;
  M3DRT_HIP(hipMalloc(&d, tab.size() * sizeof(float)));
  M3DRT_HIP(hipMemcpy(d, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
  ctx->qlg_tab = static_cast<float *>(d);
  return M3DRT_OK;
}

int line_opacity(m3drt_ctx *ctx, const m3drt_linelist *ll, const m3drt_opacity_args *o, bool general,
                 bool lref_from_cont, bool direct) {
  if (direct && (ll->nline > 4 || (o->flags & M3DRT_OPAC_SCAT)))
    return set_error(M3DRT_EINVAL, "line_opacity: the direct deposit takes at most 4 lines and no scattering split");
  int rc = qlg_tables(ctx);
  if (rc) return rc;
  OpacArgs a;
  a.nline = ll->nline;
  a.llam0 = ll->llam0;
  a.lower_ev = ll->lower_ev; a.gpref = ll->gpref; a.grad = ll->grad;
  a.gw1 = ll->gw1; a.gw2 = ll->gw2; a.gstk = ll->gstk;
  a.unsold = ll->unsold; a.lev_i = ll->lev_i; a.lev_j = ll->lev_j; a.ipart = ll->ipart;
  a.used = ll->used; a.strength = ll->strength; a.reach = ll->reach;
  a.reach_in = ll->reach_in;
  a.sel = nullptr;
  if (ll->reach_in || ll->active) {
    hipLaunchKernelGGL(select_lines_kernel, dim3(1), dim3(1024), 0, ctx->stream, ll->llam0, ll->reach_in,
                       ll->active, ll->nline, o->llam, o->nf, o->vproj, o->ncell, ll->sel_work);
    M3DRT_HIP(hipGetLastError());
    a.sel = ll->sel_work;
  }
  a.nf = o->nf; a.ncell = o->ncell; a.npart = o->npart; a.nlev = o->nlev;
  a.nlte = (o->flags & M3DRT_OPAC_NLTE) ? 1 : 0;
  a.scat = (o->flags & M3DRT_OPAC_SCAT) ? 1 : 0;
  a.llam = o->llam;
  a.vproj = o->vproj; a.idf = o->inv_df; a.vb2 = o->vbar2; a.ekt = o->e_kT; a.ne = o->ne;
  a.tothi = o->tothi; a.spe = o->stark_pe; a.lnbase = o->lnbase; a.dep = o->dep;
  a.chi = o->chi; a.eta = o->eta; a.csc = o->chi_scat;
  a.cont = o->cont; a.logref = o->logref; a.ome = o->one_m_eps;
  a.sf = o->sf; a.sc = o->sc;
  a.tol = o->tolerance; a.rej = o->line_reject_frac;
  a.lref_from_cont = lref_from_cont ? 1 : 0;
  a.lt = (float)log((double)(o->tolerance / 0.5641895835f));
  a.tgx = ctx->qlg_tab;
  a.th2x = ctx->qlg_tab + QLG_NX + 1;
  const int threads = 256;
  const int blocks = (o->ncell + threads - 1) / threads;
  const size_t lds = 2 * (QLG_NX + 1) * sizeof(float);
  auto kern = direct ? (general ? opacity_kernel<true, true, false> : opacity_kernel<false, true, false>)
              : a.scat ? (general ? opacity_kernel<true, false, true> : opacity_kernel<false, false, true>)
                       : (general ? opacity_kernel<true, false, false> : opacity_kernel<false, false, false>);
  M3DRT_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
