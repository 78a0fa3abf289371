/*
 * m3d_qlg.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the QLG-bot Voigt
 * kernel and of the equidistant-grid line-opacity summation (add_opacity_equi),
 * the checker for the HIP opacity kernel.  Citations are relative to
 * experiments/OpenMULTI3D/ of Bergemann-Group-development/dispatch.
 *
 * Pinning: qlg_ref_f64 against scipy's Faddeeva function and the reference's own
 * compiled qlg_ref_f64 (tests/golden); QLG-bot against Re w(v + ia) to the 5.7e-3
 * the reference states (HOP/qlg_mod.f90:15-18) and against the compiled
 * qlg_bot_sweep / qlg_bot_line_sweep; add_opacity_equi against the compiled
 * reference on seeded line lists and cells.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "m3d_qlg.h"

/* Scalar constants of QLG (HOP/qlg_mod.f90:86-91, 114-117): default-real literals */
#define LN2f 0.6931471806f
#define RPIf 1.7724538509f
#define OVERPIf 0.3183098862f
#define GNORMf 0.4697186f
#define SQL2f 0.8325546112f
#define ISQPIf 0.5641895835f
#define XSCLf 800.0f
#define QLG_NX 9019

static double w32_a[32];
static double w32_L;
static float TGX[QLG_NX + 1], TH2X[QLG_NX + 1];
static int qlg_ready = 0;

/* Weideman (1994), "Computation of the complex error function", SIAM J. Numer.
 * Anal. 31, 1497: N = 32 rational approximation
 *   w(z) ~ pi^-1/2/(L - iz) + 2/(L - iz)^2 sum_n a_n Z^n,  Z = (L + iz)/(L - iz),
 * L = (N/sqrt 2)^1/2, with the a_n the discrete Fourier coefficients of
 * f(t) = exp(-t^2)(L^2 + t^2) at t = L tan(theta/2), computed here by a direct
 * DFT in FP64 (the paper's recipe), highest power first as in Horner's rule. */
static void weideman32(void)
{
  const int N = 32, M = 2 * N, M2 = 2 * M;
  double f[128];
  const double PI = 3.14159265358979323846;
  w32_L = sqrt(N / sqrt(2.0));
  /* f = [0; f(k)], k = -M+1..M-1, then fftshift (rotate left by M) */
  double g[128];
  g[0] = 0.0;
  for (int k = -M + 1; k <= M - 1; k++) {
    const double th = k * PI / M, t = w32_L * tan(th / 2);
    g[k + M] = exp(-t * t) * (w32_L * w32_L + t * t);
  }
  for (int i = 0; i < M2; i++) f[i] = g[(i + M) % M2];
  /* a(j) = Re DFT(f)(j) / M2, j = 1..N; flipped so that a[0] is the highest power */
  for (int j = 1; j <= N; j++) {
    double re = 0.0;
    for (int i = 0; i < M2; i++) re += f[i] * cos(2 * PI * (double)i * j / M2);
    w32_a[N - j] = re / M2;
  }
}

void orc_qlg_ref_f64(double a, double v0, double dv, int n, double *K)
{
  const double ISQPI8 = 0.564189583547756287;
  const double dr = w32_L + a, nr0 = w32_L - a;
  if (!qlg_ready) orc_qlg_init();
  for (int f = 0; f < n; f++) {
    const double v = v0 + (double)f * dv, av = fabs(v);
    if (av + a >= 15.0) {
      /* Humlicek (1982) region I: w ~ i z / (sqrt(pi) (z^2 - 1/2)) */
      const double tr = a * a - av * av + 0.5, ti = -2.0 * a * av;
      const double q = 1.0 / (tr * tr + ti * ti);
      K[f] = ISQPI8 * ((a * tr + (-av) * ti) * q);
    } else {
      const double di = -v, q = 1.0 / (dr * dr + di * di);
      const double zr = (nr0 * dr + v * di) * q, zi = (v * dr - nr0 * di) * q;
      double pr = w32_a[0], pim = 0.0;
      for (int j = 1; j < 32; j++) {
        const double t = pr * zr - pim * zi + w32_a[j];
        pim = pr * zi + pim * zr;
        pr = t;
      }
      const double ir = dr * q, ii = -di * q;
      const double sr = ir * ir - ii * ii, si = 2.0 * ir * ii;
      K[f] = 2.0 * (pr * sr - pim * si) + ISQPI8 * ir;
    }
  }
}

/* Z_0.2(u)/L^2 of QLG-bot's second table (HOP/qlg_mod.f90:382-395): the residual of
 * the two analytic limits at the anchored damping ratio r* = 0.2 */
static double qlg_wr02(double u)
{
  const double A2 = 0.18741813853432507, GAM = 0.93709069267162530, KP = 0.81924853602069690;
  const double ETA = 0.14697983683153040, QRS = 0.16, DLN2 = 0.69314718055994531;
  const double v = GAM * sqrt(u), L = 1.0 / (1.0 + u), G = exp(-DLN2 * u);
  double Kv;
  orc_qlg_ref_f64(A2, v, 0.0, 1, &Kv);
  const double P = KP * (ETA * L + (1.0 - ETA) * G);
  return (Kv - P) / (QRS * KP) * ((1.0 + u) * (1.0 + u));
}

/* HOP/qlg_mod.f90:364-375: tables on u_i = (i/800)^2, built in FP64, rounded once */
void orc_qlg_init(void)
{
  if (qlg_ready) return;
  weideman32();
  qlg_ready = 1;
  for (int i = 0; i <= QLG_NX; i++) {
    const double x = (double)i / 800.0;
    TGX[i] = (float)pow(2.0, -x * x);
    TH2X[i] = (float)qlg_wr02(x * x);
  }
}

void orc_qlg_tables(float *tgx, float *th2x)
{
  orc_qlg_init();
  memcpy(tgx, TGX, sizeof(TGX));
  memcpy(th2x, TH2X, sizeof(TH2X));
}

void orc_qlg_w32(double *a, double *L)
{
  orc_qlg_init();
  memcpy(a, w32_a, sizeof(w32_a));
  *L = w32_L;
}

/* bot_setup, HOP/qlg_mod.f90:524-540 (the paper's QLG-bot constants) */
void orc_qlg_bot_setup(float gL, float D, float SQ, float S, orc_qlg_coef *c)
{
  const float e = D + gL * 0.60602209212341318f;
  float r, q, sc, qs, x;
  c->ih = e / (gL * e + SQ);
  r = gL * c->ih;
  q = r * (1.0f - r);
  sc = S * c->ih;
  qs = q * sc;
  x = 2.0f * r - 1.0f;
  c->bL = (OVERPIf * r) * sc;
  c->gG = (GNORMf * (1.0f - r)) * sc + qs * (-0.064924393191013452f + x * (-0.11019326278794868f));
  c->cZH = qs * (0.38494195930078284f + x * (-0.10567135279610407f));
  c->cL2 = qs * (0.04054415746208978f + x * (0.16241029253712097f));
}

/* qlg_bot_line_eval, HOP/qlg_mod.f90:542-556: one snapped index into both tables */
float orc_qlg_bot_line_eval(const orc_qlg_coef *c, float x)
{
  const float xs = fabsf(x) * c->ih, u = xs * xs, L = 1.0f / (1.0f + u);
  const float s = xs * XSCLf;
  const int ig = (int)roundf(s < (float)QLG_NX ? s : (float)QLG_NX);
  const float G = TGX[ig], W = TH2X[ig];
  float p = c->cL2 + c->cZH * W;
  p = c->bL + L * p;
  return G * c->gG + L * p;
}

/* qlg_bot_sweep, HOP/qlg_mod.f90:223-256 */
void orc_qlg_bot_sweep(float a, float v0, float dv, int n, float *K)
{
  orc_qlg_coef c;
  orc_qlg_init();
  orc_qlg_bot_setup(a, 1.0f, SQL2f, RPIf, &c);
  for (int f = 0; f < n; f++) K[f] = orc_qlg_bot_line_eval(&c, v0 + (float)f * dv);
}

/* qlg_bot_cell, HOP/qlg_mod.f90:561-571 */
void orc_qlg_bot_cell(float D, orc_qlg_cell *w)
{
  w->D = D;
  w->D2 = D * D;
  w->SQ = SQL2f * w->D2;
  w->D2h = 0.5f * w->D2;
  w->D2G = 1.5f * w->D2;
  w->D2L = 5.0f * w->D2;
}

static float fmax3(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }

/* qlg_bot_line, HOP/qlg_mod.f90:600-611: setup and the tolerance half-width */
void orc_qlg_bot_line(const orc_qlg_cell *w, float gL, float S, float ref, float lnq,
                      orc_qlg_coef *c, float *hw)
{
  float AL;
  orc_qlg_bot_setup(gL, w->D, w->SQ, S, c);
  AL = OVERPIf * gL * S / ref;
  *hw = sqrtf(fmax3(w->D2 * lnq + fminf(AL, w->D2G),
                    AL - fmaxf(gL * gL - w->D2h, 0.0f) + fminf(AL, w->D2L), 0.0f));
}

static int nint_f(float v) { return (int)roundf(v); }
static int nint_d(double v) { return (int)round(v); }

/* qlg_bot_line_sweep, HOP/qlg_mod.f90:630-653; indices 1-based as in the reference */
void orc_qlg_bot_line_sweep(const orc_qlg_cell *w, float gL, float S, float ref, float lnq, int nf,
                            const double *llam, double llc, double pos, double idl, float cap,
                            int *fa, int *fb, float *hw, float *K)
{
  orc_qlg_coef c;
  const int ic = nint_d(pos) + 1;
  const int i1 = ic < 1 ? 1 : (ic > nf ? nf : ic);
  const float x1 = (float)((llam[i1 - 1] - llc) * idl);
  float m;
  int noff;
  orc_qlg_bot_line(w, gL, S, ref, lnq, &c, hw);
  m = fminf(fminf(*hw, cap), (float)(nf + abs(ic)));
  noff = nint_f(m);
  *fa = ic - noff > 1 ? ic - noff : 1;
  *fb = ic + noff < nf ? ic + noff : nf;
  if (*fa > *fb) return;
  {
    const float x0 = x1 - (float)(i1 - *fa);
    for (int f = 1; f <= *fb - *fa + 1; f++) K[f - 1] = orc_qlg_bot_line_eval(&c, x0 + (float)(f - 1));
  }
}

/* qlg_bot_profile, HOP/qlg_mod.f90:664-671: the NLTE-rate convention */
float orc_qlg_bot_profile(float a, float q, float vpr, float idf)
{
  float K;
  const float v = fabsf(q - vpr) * idf;
  orc_qlg_bot_sweep(a, v, 0.0f, 1, &K);
  return K * ISQPIf * idf * (q + 1.0f);
}

/* lower_bound8 / upper_bound8, HOP/linelist_mod.f90:1807-1844: the first index
 * (1-based) with a(i) >= x (n+1 if none), the last with a(i) <= x (0 if none). */
static int lower_bound8(const double *a, int n, double x)
{
  int lo = 1, hi = n + 1;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (a[mid - 1] >= x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

static int upper_bound8(const double *a, int n, double x)
{
  int lo = 1, hi = n + 1;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (a[mid - 1] > x) hi = mid;
    else lo = mid + 1;
  }
  return lo - 1;
}

void orc_select_lines(const orc_linelist *ll, int nf, const double *llam, double vmax, int *l0, int *l1,
                      int32_t *sel, int *nsel, double hb[2])
{
  double blue = 0, red = 0;
  hb[0] = hb[1] = 0;
  if (nf > 1) {
    hb[0] = 0.5 * (llam[1] - llam[0]);
    hb[1] = 0.5 * (llam[nf - 1] - llam[nf - 2]);
  }
  if (ll->reach_in) {
    float rmax = ll->reach_in[0];
    for (int l = 1; l < ll->nline; l++) rmax = ll->reach_in[l] > rmax ? ll->reach_in[l] : rmax;
    blue = llam[0] - hb[0] - vmax;
    red = llam[nf - 1] + hb[1] + vmax;
    const int a = lower_bound8(ll->llam0, ll->nline, blue - (double)rmax);
    const int b = upper_bound8(ll->llam0, ll->nline, red + (double)rmax);
    *l0 = *l0 > a ? *l0 : a;
    *l1 = *l1 < b ? *l1 : b;
  }
  *nsel = 0;
  for (int l = *l0; l <= *l1; l++) {
    if (ll->active && !ll->active[l - 1]) continue;
    if (ll->reach_in) {
      const double r = (double)ll->reach_in[l - 1];
      if (ll->llam0[l - 1] + r < blue || ll->llam0[l - 1] - r > red) continue;
    }
    sel[(*nsel)++] = l;
  }
}

/* max |vproj| over the call's cells, as real(4), widened (select_lines' vmax) */
static double vproj_max(const float *vproj, int ncell)
{
  float v = 0.0f;
  for (int c = 0; c < ncell; c++) v = fabsf(vproj[c]) > v ? fabsf(vproj[c]) : v;
  return (double)v;
}

/*
 * add_opacity_equi, HOP/linelist_mod.f90:729-960, for one isotope's line list over
 * ncell cells (the cell loop flattened, Fortran order) without chunking.
 * chi, eta, cont, logref, one_m_eps, chi_scat are (nf, ncell); lnbase (ncell, np);
 * dep (ncell, nlev).  The lines of select_lines are visited in the list's order; with
 * reaches, a line centred beyond the grid takes its reference at the nearer grid end
 * and its window is capped by its reach (:844-858, 914-921).
 */
void orc_add_opacity_equi(const orc_linelist *ll, int nf, const double *llam, int ncell,
                          float *chi, float *eta, const float *cont, const float *logref,
                          const float *vproj, const float *inv_df, const float *vbar2,
                          const float *e_kT, const float *ne, const float *tothi,
                          const float *stark_pe, int np, const float *lnbase, int nlev,
                          const float *dep, int do_nlte, const float *one_m_eps, float *chi_scat,
                          int do_scat, float tol, float line_reject_frac)
{
  const float sqrt_pi = 1.7724538509f, ex_cutoff = -85.0f, voigt_peak = 0.5641895835f;
  const int nl = ll->nline;
  double llam1, dllam, inv_dllam;
  float sdl;
  float *Kbuf;
  if (nl == 0 || nf < 2) return;
  orc_qlg_init();
  Kbuf = (float *)malloc(sizeof(float) * nf);
  int32_t *sel = (int32_t *)malloc(sizeof(int32_t) * (nl > 0 ? nl : 1));
  int l0 = 1, l1 = nl, nsel;
  double hb[2];
  orc_select_lines(ll, nf, llam, vproj_max(vproj, ncell), &l0, &l1, sel, &nsel, hb);
  llam1 = llam[0];
  dllam = (llam[nf - 1] - llam[0]) / (double)(nf - 1);
  inv_dllam = 1.0 / dllam;
  sdl = (float)inv_dllam;
  for (int cidx = 0; cidx < ncell; cidx++) {
    float *chic = chi + (size_t)nf * cidx, *etac = eta + (size_t)nf * cidx;
    const float *contc = cont + (size_t)nf * cidx, *lrc = logref + (size_t)nf * cidx;
    int f0_prev = -1;
    float refl_c = 0, lref_c = 0;
    const float idf = inv_df[cidx];
    orc_qlg_cell qw;
    orc_qlg_bot_cell(sdl / idf, &qw);
    const float unsold_c = powf(vbar2[cidx], 0.3f) * tothi[cidx];
    const float lvb2 = logf(vbar2[cidx]);
    const float ekt = e_kT[cidx], ne_c = ne[cidx], tothi_c = tothi[cidx];
    const float starkpe = stark_pe[cidx], vproj_c = vproj[cidx];
    for (int is = 0; is < nsel; is++) {
      const int l = sel[is] - 1;
      const double llc = ll->llam0[l] - (double)vproj_c;
      const double pos = (llc - llam1) * inv_dllam;
      int f0 = nint_d(pos) + 1;
      f0 = f0 < 1 ? 1 : (f0 > nf ? nf : f0);
      const float ex = ll->lower_ev[l] * ekt;
      const int ip = ll->ipart ? ll->ipart[l] : 1;
      const float arg = lnbase[cidx + (size_t)ncell * (ip - 1)] + ll->gpref[l] - ex;
      float Rij, ref, a, hw, cap, lnq, bi, bj, lref_o = 0;
      int fa, fb, nlte;
      if (arg < ex_cutoff) continue;
      const int outside = ll->reach_in && (llc < llam[0] || llc > llam[nf - 1]);
      if (outside) {
        const int fe = llc < llam[0] ? 1 : nf;
        lref_o = lrc[fe - 1];
        if (arg <= lref_o) continue;
        Rij = expf(arg);
        ref = tol * contc[fe - 1];
      } else {
        if (f0 != f0_prev) {
          f0_prev = f0;
          refl_c = tol * contc[f0 - 1];
          const float flo = fmaxf(tol, line_reject_frac) * tol * chic[f0 - 1];
          if (flo > refl_c) {
            refl_c = flo;
            lref_c = logf(flo * sqrt_pi);
          } else {
            lref_c = lrc[f0 - 1];
          }
        }
        if (arg <= lref_c) continue;
        Rij = expf(arg);
        ref = refl_c;
        if (voigt_peak * Rij <= ref) continue;
      }
      nlte = do_nlte && ll->lev_i[l] > 0 && ll->lev_j[l] > 0;
      if (nlte) {
        bi = dep[cidx + (size_t)ncell * (ll->lev_i[l] - 1)];
        bj = dep[cidx + (size_t)ncell * (ll->lev_j[l] - 1)];
      } else {
        bi = 1.0f;
        bj = 1.0f;
      }
      a = ll->grad[l];
      if (ll->unsold[l])
        a = a + ll->gw1[l] * unsold_c;
      else
        a = a + ll->gw1[l] * expf(ll->gw2[l] * lvb2) * tothi_c;
      if (ll->gstk[l] >= 0.0f)
        a = a + ll->gstk[l] * ne_c;
      else
        a = a - ll->gstk[l] * starkpe;
      a = a * sdl;
      if (outside) {
        cap = (float)((double)ll->reach_in[l] * inv_dllam);
        lnq = arg - lref_o;
      } else {
        const float st = voigt_peak * Rij / contc[f0 - 1];
        if (st > ll->strength[l]) ll->strength[l] = st;
        cap = 3.40282347e38f;
        lnq = arg - lref_c;
      }
      orc_qlg_bot_line_sweep(&qw, a, Rij * qw.D, ref, lnq, nf, llam, llc, pos, inv_dllam, cap,
                             &fa, &fb, &hw, Kbuf);
      {
        const float rr = (float)(hw * dllam);
        if (rr > ll->reach[l]) ll->reach[l] = rr;
      }
      if (fa > fb) continue;
      ll->used[l] = 1;
      if (do_scat && !nlte) {
        for (int f = fa; f <= fb; f++) {
          const float lte = Kbuf[f - fa];
          chic[f - 1] = chic[f - 1] + lte;
          etac[f - 1] = etac[f - 1] + lte;
          chi_scat[(size_t)nf * cidx + f - 1] = chi_scat[(size_t)nf * cidx + f - 1]
                                                + one_m_eps[(size_t)nf * cidx + f - 1] * lte;
        }
      } else {
        for (int f = fa; f <= fb; f++) {
          const float lte = Kbuf[f - fa];
          chic[f - 1] = chic[f - 1] + lte * bi;
          etac[f - 1] = etac[f - 1] + lte * bj;
        }
      }
      f0_prev = -1;
    }
  }
  free(Kbuf);
  free(sel);
}

/* bisect_nearest, HOP/linelist_mod.f90:1075-1095: the first index of llam(ws:we)
 * (1-based) at or above x, moved one down when the node below is strictly nearer. */
static int bisect_nearest(const double *llam, int ws, int we, double x)
{
  int lo = ws, hi = we;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (llam[mid - 1] < x) lo = mid + 1;
    else hi = mid;
  }
  if (lo > ws && fabs(llam[lo - 2] - x) < fabs(llam[lo - 1] - x)) lo = lo - 1;
  return lo;
}

/*
 * add_opacity, HOP/linelist_mod.f90:469-719, on an ascending ln(lambda) grid of any
 * spacing, for one isotope's line list over ncell cells (no chunking, reaches or
 * per-line windows: every line is selected, select_lines :1110-1142 without
 * reaches).  Per cell a cursor marches forward to each (ascending) line centre and
 * takes the nearer node (:578-590); the profile is set up on the ln(lambda)
 * abscissa itself (qlg_bot_cell(1/idf), the physical damping, qlg_bot_line) and its
 * window found by bisection unless it covers the whole grid (:679-689).  Arrays as
 * orc_add_opacity_equi.
 */
void orc_add_opacity(const orc_linelist *ll, int nf, const double *llam, int ncell, float *chi, float *eta,
                     const float *cont, const float *logref, const float *vproj, const float *inv_df,
                     const float *vbar2, const float *e_kT, const float *ne, const float *tothi,
                     const float *stark_pe, int np, const float *lnbase, int nlev, const float *dep,
                     int do_nlte, const float *one_m_eps, float *chi_scat, int do_scat, float tol,
                     float line_reject_frac)
{
  const float sqrt_pi = 1.7724538509f, ex_cutoff = -85.0f, voigt_peak = 0.5641895835f;
  const int nl = ll->nline;
  (void)nlev;
  (void)np;
  if (nl == 0 || nf < 1) return;
  orc_qlg_init();
  int32_t *sel = (int32_t *)malloc(sizeof(int32_t) * nl);
  int l0 = 1, l1 = nl, nsel;
  double hb[2];
  orc_select_lines(ll, nf, llam, vproj_max(vproj, ncell), &l0, &l1, sel, &nsel, hb);
  for (int cidx = 0; cidx < ncell; cidx++) {
    float *chic = chi + (size_t)nf * cidx, *etac = eta + (size_t)nf * cidx;
    const float *contc = cont + (size_t)nf * cidx, *lrc = logref + (size_t)nf * cidx;
    int f0_prev = -1, fc = 1;
    const int ws = 1, we = nf;
    float refl_c = 0, lref_c = 0;
    const float idf = inv_df[cidx];
    orc_qlg_cell qw;
    orc_qlg_bot_cell(1.0f / idf, &qw);
    const float unsold_c = powf(vbar2[cidx], 0.3f) * tothi[cidx];
    const float lvb2 = logf(vbar2[cidx]);
    const float ekt = e_kT[cidx], ne_c = ne[cidx], tothi_c = tothi[cidx];
    const float starkpe = stark_pe[cidx], vproj_c = vproj[cidx];
    for (int is = 0; is < nsel; is++) {
      const int l = sel[is] - 1;
      const double llc = ll->llam0[l] - (double)vproj_c;
      while (fc < nf && !(llam[fc - 1] >= llc)) fc++;
      int f0;
      if (fc > 1) f0 = fabs(llam[fc - 1] - llc) < fabs(llam[fc - 2] - llc) ? fc : fc - 1;
      else f0 = 1;
      const float ex = ll->lower_ev[l] * ekt;
      const int ip = ll->ipart ? ll->ipart[l] : 1;
      const float arg = lnbase[cidx + (size_t)ncell * (ip - 1)] + ll->gpref[l] - ex;
      float Rij, ref, a, hw, bi, bj, lref_o = 0;
      int fa, fb, nlte;
      if (arg < ex_cutoff) continue;
      const int outside = ll->reach_in && (llc < llam[0] || llc > llam[nf - 1]);
      if (outside) {
        const int fe = llc < llam[0] ? 1 : nf;
        lref_o = lrc[fe - 1];
        if (arg <= lref_o) continue;
        Rij = expf(arg);
        ref = tol * contc[fe - 1];
      } else {
        if (f0 != f0_prev) {
          f0_prev = f0;
          refl_c = tol * contc[f0 - 1];
          const float flo = fmaxf(tol, line_reject_frac) * tol * chic[f0 - 1];
          if (flo > refl_c) {
            refl_c = flo;
            lref_c = logf(flo * sqrt_pi);
          } else {
            lref_c = lrc[f0 - 1];
          }
        }
        if (arg <= lref_c) continue;
        Rij = expf(arg);
        ref = refl_c;
        if (voigt_peak * Rij <= ref) continue;
      }
      nlte = do_nlte && ll->lev_i[l] > 0 && ll->lev_j[l] > 0;
      if (nlte) {
        bi = dep[cidx + (size_t)ncell * (ll->lev_i[l] - 1)];
        bj = dep[cidx + (size_t)ncell * (ll->lev_j[l] - 1)];
      } else {
        bi = 1.0f;
        bj = 1.0f;
      }
      a = ll->grad[l];
      if (ll->unsold[l])
        a = a + ll->gw1[l] * unsold_c;
      else
        a = a + ll->gw1[l] * expf(ll->gw2[l] * lvb2) * tothi_c;
      if (ll->gstk[l] >= 0.0f)
        a = a + ll->gstk[l] * ne_c;
      else
        a = a - ll->gstk[l] * starkpe;
      orc_qlg_coef qc;
      if (outside) {
        orc_qlg_bot_line(&qw, a, Rij * qw.D, ref, arg - lref_o, &qc, &hw);
        double dl = (double)hw;
        if (hw > ll->reach[l]) ll->reach[l] = hw;
        dl = fmin(dl, (double)ll->reach_in[l]);
        if (llc + dl < llam[0] - hb[0] || llc - dl > llam[nf - 1] + hb[1]) continue;
        ll->used[l] = 1;
        fa = bisect_nearest(llam, ws, we, llc - dl);
        fb = bisect_nearest(llam, ws, we, llc + dl);
      } else {
        ll->used[l] = 1;
        {
          const float st = voigt_peak * Rij / contc[f0 - 1];
          if (st > ll->strength[l]) ll->strength[l] = st;
        }
        orc_qlg_bot_line(&qw, a, Rij * qw.D, ref, arg - lref_c, &qc, &hw);
        if (hw > ll->reach[l]) ll->reach[l] = hw;
        const double dl = (double)hw;
        const double span = fmax(llam[we - 1] - llc, llc - llam[ws - 1]);
        if (dl >= span) {
          fa = ws;
          fb = we;
        } else {
          fa = bisect_nearest(llam, ws, f0, llc - dl);
          fb = bisect_nearest(llam, f0, we, llc + dl);
        }
      }
      for (int f = fa; f <= fb; f++) {
        const float lte = orc_qlg_bot_line_eval(&qc, (float)(llam[f - 1] - llc));
        if (do_scat && !nlte) {
          chic[f - 1] = chic[f - 1] + lte;
          etac[f - 1] = etac[f - 1] + lte;
          chi_scat[(size_t)nf * cidx + f - 1] = chi_scat[(size_t)nf * cidx + f - 1]
                                                + one_m_eps[(size_t)nf * cidx + f - 1] * lte;
        } else {
          chic[f - 1] = chic[f - 1] + lte * bi;
          etac[f - 1] = etac[f - 1] + lte * bj;
        }
      }
      f0_prev = -1;
    }
  }
  free(sel);
}
