/*
 * m3d_lch.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of long_char_hermite
 * (SRC/rt_solvers_mod.f90:1052-1504 of experiments/OpenMULTI3D/ in
 * Bergemann-Group-development/dispatch), the cubic-Bezier output-pass solver along
 * columns, the checker for the HIP kernel.  FP32 throughout, every operation
 * separately rounded (the Makefile builds with -ffp-contract=off).
 *
 *   bezier3_weights  :1301-1323 (series P2, P3 :1327-1352, moments above t = 3)
 *   herm_dtau        :1365-1394 (product-rule chi slopes, MC envelope, Euler-Maclaurin)
 *   herm_source      :1403-1459 (rho = S/B parabola, Bezier control points, clamp)
 *   herm_first       :1465-1520 (the zeroed top ghost: secant slope at the upwind end)
 *   herm_plane       :1171-1250 and the plane order of :1100-1160
 *   ghost_thermo     :1257-1266
 * and long_char_linear (:666-757), the trapezoid step with linear_weights along each
 * column, with the split parts of the 'largest' rate formulation, and
 * long_char_hermite2 (:1536-1895): long_char_hermite's product-rule optical depth
 * (slopes clamped to 3x the secant) with long_char_quadratic's MC-limited source
 * slope and parab_moments weights (A, B, C in FP64).
 *
 * Pinning (tests/test_oracle_lch.py): the four weights against numerical
 * quadrature of the Bernstein integrals over -3 <= t <= 87, and their sum against
 * 1 - e^-t; a source linear in tau is reproduced exactly (I = S - dS/dtau on every
 * plane) in all three boundary modes; with S = 0 and chi = (chi/Pe) Pe, chi/Pe
 * quadratic and Pe linear along the column, I = I_ghost exp(-tau) with tau the exact
 * integral (the product-rule slopes and the Euler-Maclaurin step are exact for the
 * cubic chi).  The Lambda diagonal is the reference's heuristic (a bounded
 * surrogate of dI/dS, not the derivative): it is pinned only by 0 <= chi Psi <= 1
 * and by its isothermal, constant-chi value from the quadrature weights.
 * long_char_linear is pinned by test_oracle_lch.py too: along a vertical ray it
 * equals short_char (whose 4x4 stencil then reads the column itself), a source linear
 * in tau with the ghost on the particular solution is exact, and Irem + Psi S = I.
 * long_char_hermite2 likewise: a source linear in tau is exact on every plane in the
 * three boundary kinds (the parabolic step reproduces it), and a cubic chi with
 * S = 0 attenuates by the exact optical depth.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "m3d_lch.h"

#define MIN_(a, b) ((a) < (b) ? (a) : (b))
#define MAX_(a, b) ((a) > (b) ? (a) : (b))

static float lw_select(float a, float b, int first) { return first ? a : b; }

/* P2 = w2/t and P3 = w3/t on |t| <= 3: the reference's minimax coefficients
 * (SRC/rt_solvers_mod.f90:1330-1333, 1344-1347), lowest order first */
static float b3_p2(float t)
{
  static const float c[11] = {+2.500000000e-01f, -1.000000238e-01f, +2.499995567e-02f, -4.761847202e-03f,
                              +7.440865738e-04f, -9.924791084e-05f, +1.156294547e-05f, -1.190877015e-06f,
                              +1.137804801e-07f, -1.100420821e-08f, +7.658396139e-10f};
  float r = c[10];
  for (int k = 9; k >= 0; k--) r = c[k] + t * r;
  return r;
}

static float b3_p3(float t)
{
  static const float c[10] = {+2.500000000e-01f, -4.999998957e-02f, +8.333289996e-03f, -1.190488110e-03f,
                              +1.488493726e-04f, -1.653099389e-05f, +1.640762775e-06f, -1.503805294e-07f,
                              +1.414234507e-08f, -1.025371787e-09f};
  float r = c[9];
  for (int k = 8; k >= 0; k--) r = c[k] + t * r;
  return r;
}

void orc_bezier3_weights(float t, float *E, float *w0, float *w1, float *w2, float *w3)
{
  const float tser = 3.0f;
  float e, M0;
  orc_exp_neg(t, &e, &M0);
  const float tb = MAX_(t, tser), tb2 = tb * tb, tb3 = tb2 * tb, r3 = 1.0f / tb3;
  const float M1 = M0 - tb * e;
  const float M2 = 2.0f * M1 - tb2 * e;
  const float M3 = 3.0f * M2 - tb3 * e;
  const float b0 = M3 * r3;
  const float b1 = 3.0f * (M2 * tb - M3) * r3;
  const float b2 = 3.0f * (M1 * tb2 - 2.0f * M2 * tb + M3) * r3;
  const int ser = t <= tser;
  const float tE = t * e;
  *E = e;
  *w0 = lw_select(tE * b3_p3(-t), b0, ser);
  *w1 = lw_select(tE * b3_p2(-t), b1, ser);
  *w2 = lw_select(t * b3_p2(t), b2, ser);
  *w3 = lw_select(t * b3_p3(t), M0 - b0 - b1 - b2, ser);
}

void orc_bezier3_weights_n(int n, const float *t, float *E, float *w0, float *w1, float *w2, float *w3)
{
  for (int i = 0; i < n; i++) orc_bezier3_weights(t[i], E + i, w0 + i, w1 + i, w2 + i, w3 + i);
}

typedef struct {
  float rdsup, dsup1, dsup2, mdtau, imax, third;
  int get_rates;
} lch_c;

/* herm_dtau: the product-rule chi slopes at both ends, clamped into [0, 2|secant|]
 * with the secant's sign, then the Euler-Maclaurin trapezoid */
static float herm_dtau(const lch_c *k, float xxup, float xxc, float xx3, float peu, float rpeu, float gpu,
                       float pec, float rpec, float gpc, float rpe3, int center_target)
{
  const float dtmin = 1e-20f;
  const float hh = 0.5f * k->rdsup;
  const float t1 = (xxc - xxup) * k->rdsup;
  const float sgn = copysignf(1.0f, t1);
  const float rxpu = xxup * rpeu, rxpc = xxc * rpec, rxp3 = xx3 * rpe3;
  float drxpu, drxpc;
  if (center_target) {
    drxpc = (rxp3 - rxpu) * hh;
    drxpu = (4.0f * rxpc - 3.0f * rxpu - rxp3) * hh;
  } else {
    drxpu = (rxpc - rxp3) * hh;
    drxpc = (rxp3 - 4.0f * rxpu + 3.0f * rxpc) * hh;
  }
  float mup = rxpu * gpu + peu * drxpu;
  float mc = rxpc * gpc + pec * drxpc;
  mup = sgn * MAX_(0.0f, MIN_(sgn * mup, 2.0f * fabsf(t1)));
  mc = sgn * MAX_(0.0f, MIN_(sgn * mc, 2.0f * fabsf(t1)));
  float dt = k->dsup1 * (xxup + xxc) + k->dsup2 * (mup - mc);
  dt = MAX_(copysignf(MAX_(fabsf(dt), dtmin), dt), k->mdtau);
  return dt;
}

/* herm_source: I and the Lambda diagonal over [up, c] from the Bezier weights */
static void herm_source(const lch_c *k, float xxup, float xxc, float ssup, float ssc, float ss3, float betau,
                        float betac, float beta3, float hk, float Inuup, float aTu, float aTc, int center_target,
                        float dt, float ex, float w0, float w1, float w2, float w3, float *Inu, float *plam)
{
  const float hh = 0.5f * k->rdsup;
  const float r1mbu = 1.0f / (1.0f - betau), r1mbc = 1.0f / (1.0f - betac), r1mb3 = 1.0f / (1.0f - beta3);
  const float Bu = betau * r1mbu, Bc = betac * r1mbc, B3 = beta3 * r1mb3;
  const float rhou = ssup / Bu, rhoc = ssc / Bc, rho3 = ss3 / B3;
  float drhu, drhc;
  if (center_target) {
    drhc = (rho3 - rhou) * hh;
    drhu = (4.0f * rhoc - 3.0f * rhou - rho3) * hh;
  } else {
    drhu = (rhoc - rho3) * hh;
    drhc = (rho3 - 4.0f * rhou + 3.0f * rhoc) * hh;
  }
  const float rxxu = 1.0f / xxup, rxxc = 1.0f / xxc;
  const float gu = hk * r1mbu * aTu * rxxu;
  const float gc = hk * r1mbc * aTc * rxxc;
  const float dt3 = dt * k->third;
  float P1 = ssup + dt3 * (ssup * gu + Bu * drhu * rxxu);
  float P2 = ssc - dt3 * (ssc * gc + Bc * drhc * rxxc);
  const float smin = MIN_(ssup, ssc), smax = MAX_(ssup, ssc);
  P1 = MIN_(MAX_(P1, smin), smax);
  P2 = MIN_(MAX_(P2, smin), smax);
  float I = Inuup * ex + w0 * ssup + w1 * P1 + w2 * P2 + w3 * ssc;
  *Inu = MAX_(-k->imax, MIN_(k->imax, I));
  const float d1 = MIN_(MAX_(Bu * k->third / Bc, 0.0f), 1.0f);
  const float d2 = MIN_(MAX_(2.0f * k->third - dt3 * gc, 0.0f), 1.0f);
  *plam = lw_select((w1 * d1 + w2 * d2 + w3) * rxxc, 0.0f, k->get_rates && dt > 0 && xxc > 0);
}

/* herm_first: the interval [ghost, zs] below a zeroed top ghost */
static void herm_first(const lch_c *k, float xxup, float xxc, float xxdn, float ssup, float ssc, float ssdn,
                       float betac, float betadn, float hk, float Inuup, float pec, float rpec, float pedn,
                       float rpedn, float gpc, float aTc, float *Inu, float *plam)
{
  const float dtmin = 1e-20f;
  (void)pedn;
  const float t1 = (xxc - xxup) * k->rdsup;
  const float sgn = copysignf(1.0f, t1);
  const float rxpc = xxc * rpec;
  float mc = rxpc * gpc + pec * (xxdn * rpedn - rxpc) * k->rdsup;
  mc = sgn * MAX_(0.0f, MIN_(sgn * mc, 2.0f * fabsf(t1)));
  float dt = k->dsup1 * (xxup + xxc) + k->dsup2 * (t1 - mc);
  dt = MAX_(copysignf(MAX_(fabsf(dt), dtmin), dt), k->mdtau);
  float ex, w0, w1, w2, w3;
  orc_bezier3_weights(dt, &ex, &w0, &w1, &w2, &w3);
  const float r1mbc = 1.0f / (1.0f - betac), r1mbdn = 1.0f / (1.0f - betadn);
  const float Bc = betac * r1mbc, Bdn = betadn * r1mbdn;
  const float Gc = hk * r1mbc * aTc / xxc;
  const float rhoc = ssc / Bc, rhodn = ssdn / Bdn;
  float taudn = k->dsup1 * (xxc + xxdn);
  taudn = copysignf(MAX_(fabsf(taudn), dtmin), taudn);
  const float dSc = ssc * Gc + Bc * (rhodn - rhoc) / taudn;
  const float dss = ssc - ssup;
  const float dSu = dss / dt;
  const float dt3 = dt * k->third;
  float P1 = ssup + dt3 * dSu;
  float P2 = ssc - dt3 * dSc;
  const float smin = MIN_(ssup, ssc), smax = MAX_(ssup, ssc);
  P1 = MIN_(MAX_(P1, smin), smax);
  P2 = MIN_(MAX_(P2, smin), smax);
  float I = Inuup * ex + w0 * ssup + w1 * P1 + w2 * P2 + w3 * ssc;
  *Inu = MAX_(-k->imax, MIN_(k->imax, I));
  *plam = lw_select((w1 * k->third + w2 + w3) / xxc, 0.0f, k->get_rates && dt > 0 && xxc > 0);
}

void orc_long_char_hermite_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                               const float *beta, const float *pe4, const float *temp, const float *dpe_dz,
                               const float *dlnt_dz, const float *hknu, int ghost_kind, const float *gpe,
                               const float *gdpe, const float *gtemp, const float *gdlnt, float *Psi)
{
  const int nf = p->nf, nx = p->nx, ny = p->ny, nz = p->nz, dir = p->dir;
  const size_t sx = (size_t)nf, sy = sx * nx, sz = sy * ny, nxy = (size_t)nx * ny;
  lch_c k;
  const float ds = (float)p->ds[2];
  const float dsup = (float)((double)ds / fabs(p->vec[2]));
  k.rdsup = 1.0f / dsup;
  k.dsup1 = 0.5f * dsup;
  k.dsup2 = 1.0f / 12.0f * dsup * dsup;
  k.mdtau = -(float)p->maser_dtau_max;
  k.imax = (float)p->maser_Inu_max;
  k.third = (float)(1.0 / 3.0);
  k.get_rates = p->get_rates;
  const float gfac = (float)(dir * fabs(p->vec[2]));
  const int zs = dir > 0 ? 1 : nz, ze = dir > 0 ? nz : 1;
  const int upstream = ghost_kind == ORC_LCH_UPSTREAM;
#define W_(a, f, x, y, z) a[(f) + sx * (x) + sy * (y) + sz * (size_t)(z)]
#define C_(a, x, y, z) a[(x) + (size_t)nx * (y) + nxy * (size_t)((z) - 1)]
#define G_(a, x, y) a[(x) + (size_t)nx * (y)]
#define B_(f, x, y, z) beta[(f) + sx * (x) + sy * (y) + sz * (size_t)((z) - 1)]
#define P_(f, x, y, z) Psi[(f) + sx * (x) + sy * (y) + sz * (size_t)((z) - 1)]
  /* one plane z over [zup, z]; zt the parabola's third sample (herm_plane) */
#define HERM_PLANE(Z, ZT, CT)                                                                                  \
  do {                                                                                                         \
    const int z = (Z), zt = (ZT), ct = (CT), zup = z - dir;                                                    \
    const int ghost_up = z == zs, ghost_t = zt < 1 || zt > nz;                                                 \
    for (int y = 0; y < ny; y++)                                                                               \
      for (int x = 0; x < nx; x++) {                                                                           \
        const float pec = C_(pe4, x, y, z), rpec = 1.0f / pec, gpc = gfac * C_(dpe_dz, x, y, z);               \
        const float aTc = gfac * C_(dlnt_dz, x, y, z) / C_(temp, x, y, z);                                     \
        float peu, rpeu, gpu, aTu, rTup = 0.0f, rpet;                                                          \
        if (ghost_up) {                                                                                        \
          peu = G_(gpe, x, y);                                                                                 \
          rpeu = 1.0f / peu;                                                                                   \
          gpu = gfac * G_(gdpe, x, y);                                                                         \
          rTup = 1.0f / G_(gtemp, x, y);                                                                       \
          aTu = gfac * G_(gdlnt, x, y) * rTup;                                                                 \
        } else {                                                                                               \
          peu = C_(pe4, x, y, zup);                                                                            \
          rpeu = 1.0f / peu;                                                                                   \
          gpu = gfac * C_(dpe_dz, x, y, zup);                                                                  \
          aTu = gfac * C_(dlnt_dz, x, y, zup) / C_(temp, x, y, zup);                                           \
        }                                                                                                      \
        if (!ghost_t) rpet = 1.0f / C_(pe4, x, y, zt);                                                         \
        else if (upstream) rpet = 1.0f / G_(gpe, x, y);                                                        \
        else rpet = rpec;                                                                                      \
        for (int f = 0; f < nf; f++) {                                                                         \
          const float hk = hknu[f];                                                                            \
          float bu, bt;                                                                                        \
          if (ghost_up) bu = MAX_(expf(-hk * rTup), FLT_MIN_);                                                 \
          else bu = B_(f, x, y, zup);                                                                          \
          if (!ghost_t) bt = B_(f, x, y, zt);                                                                  \
          else if (upstream) bt = MAX_(expf(-hk / G_(gtemp, x, y)), FLT_MIN_);                                 \
          else bt = B_(f, x, y, z);                                                                            \
          const float dt = herm_dtau(&k, W_(chi, f, x, y, zup), W_(chi, f, x, y, z), W_(chi, f, x, y, zt),     \
                                     peu, rpeu, gpu, pec, rpec, gpc, rpet, ct);                                \
          float ex, w0, w1, w2, w3, Inu, pl;                                                                   \
          orc_bezier3_weights(dt, &ex, &w0, &w1, &w2, &w3);                                                    \
          herm_source(&k, W_(chi, f, x, y, zup), W_(chi, f, x, y, z), W_(S, f, x, y, zup), W_(S, f, x, y, z),  \
                      W_(S, f, x, y, zt), bu, B_(f, x, y, z), bt, hk, W_(I, f, x, y, zup), aTu, aTc, ct, dt,   \
                      ex, w0, w1, w2, w3, &Inu, &pl);                                                          \
          W_(I, f, x, y, z) = Inu;                                                                             \
          if (p->get_rates && Psi) P_(f, x, y, z) = pl;                                                        \
        }                                                                                                      \
      }                                                                                                        \
  } while (0)
#define FLT_MIN_ 1.17549435e-38f
  if (ghost_kind == ORC_LCH_UPSTREAM || ghost_kind == ORC_LCH_BOTTOM) {
    HERM_PLANE(zs, zs + dir, 1);
  } else {
    const int z = zs, zup = z - dir, zdn = z + dir;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const float pec = C_(pe4, x, y, z), pedn = C_(pe4, x, y, zdn);
        const float rpec = 1.0f / pec, rpedn = 1.0f / pedn;
        const float gpc = gfac * C_(dpe_dz, x, y, z);
        const float aTc = gfac * C_(dlnt_dz, x, y, z) / C_(temp, x, y, z);
        for (int f = 0; f < nf; f++) {
          float Inu, pl;
          herm_first(&k, W_(chi, f, x, y, zup), W_(chi, f, x, y, z), W_(chi, f, x, y, zdn), W_(S, f, x, y, zup),
                     W_(S, f, x, y, z), W_(S, f, x, y, zdn), B_(f, x, y, z), B_(f, x, y, zdn), hknu[f],
                     W_(I, f, x, y, zup), pec, rpec, pedn, rpedn, gpc, aTc, &Inu, &pl);
          W_(I, f, x, y, z) = Inu;
          if (p->get_rates && Psi) P_(f, x, y, z) = pl;
        }
      }
  }
  for (int zi = zs + dir; zi != ze; zi += dir) HERM_PLANE(zi, zi + dir, 1);
  HERM_PLANE(ze, ze - 2 * dir, 0);
#undef HERM_PLANE
#undef FLT_MIN_
#undef W_
#undef C_
#undef G_
#undef B_
#undef P_
}

void orc_long_char_linear_f32(const orc_sc_params *p, const float *chi, const float *S, float *I, float *Psi,
                              float *Irem, float *escape)
{
  const int nf = p->nf, nx = p->nx, ny = p->ny, nz = p->nz, dir = p->dir;
  const size_t sx = (size_t)nf, sy = sx * nx, sz = sy * ny;
  const float ds = (float)p->ds[2];
  const float dsup1 = (float)((double)(0.5f * ds) / fabs(p->vec[2]));
  const float mdt = -(float)p->maser_dtau_max, imax = (float)p->maser_Inu_max;
  const int zs = dir > 0 ? 1 : nz, ze = dir > 0 ? nz : 1;
#define W_(a, f, x, y, z) a[(f) + sx * (x) + sy * (y) + sz * (size_t)(z)]
#define O_(a, f, x, y, z) a[(f) + sx * (x) + sy * (y) + sz * (size_t)((z) - 1)]
  for (int z = zs; dir > 0 ? z <= ze : z >= ze; z += dir) {
    const int zup = z - dir;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++)
        for (int f = 0; f < nf; f++) {
          const float xxup = W_(chi, f, x, y, zup), ssup = W_(S, f, x, y, zup), Inuup = W_(I, f, x, y, zup);
          const float xx = W_(chi, f, x, y, z), ss = W_(S, f, x, y, z);
          float dt = dsup1 * (xxup + xx);
          dt = MAX_(dt, mdt);
          float E, W, Ps, Esc;
          orc_linear_weights(dt, &E, &W, &Ps, &Esc);
          float Inu = Inuup * E + W * ssup + Ps * ss;
          if (p->split_rates) {
            const int off = dt <= 0 || xx <= 0 || fabsf(Inu) > imax;
            const float Ir = Inuup * E + W * ssup;
            Inu = MAX_(-imax, MIN_(imax, Inu));
            Ps = off ? 0.0f : Ps;
            if (Irem) O_(Irem, f, x, y, z) = off ? Inu : Ir;
            if (escape) O_(escape, f, x, y, z) = off ? 1.0f : Esc;
          } else {
            Inu = MAX_(-imax, MIN_(imax, Inu));
          }
          W_(I, f, x, y, z) = Inu;
          if (p->get_rates && Psi) O_(Psi, f, x, y, z) = (dt > 0 && xx > 0) ? Ps / xx : 0.0f;
        }
  }
#undef W_
#undef O_
}

typedef struct {
  float rdsup, dsup1, dsup2, imax;
  double mdtau;
  int get_rates;
} lch2_c;

static double tau_floor64(double dt, double mdtau)
{
  return MAX_(copysign(MAX_(fabs(dt), (double)1e-20f), dt), mdtau);
}

/* herm2_step (SRC/rt_solvers_mod.f90:1768-1853); onesided: 1 the third sample is
 * real (last plane, nz >= 3), 0 it is the ghost, -1 not the last plane */
static void herm2_step(const lch2_c *k, float xxup, float xxc, float xx3, float ssup, float ssc, float ss3,
                       float Inuup, float peu, float rpeu, float gpu, float pec, float rpec, float gpc, float pe3,
                       float rpe3, float gp3, int ct, int onesided, float *Inu, float *plam)
{
  const float hh = 0.5f * k->rdsup;
  const float dchi = xxc - xxup;
  float t1 = dchi * k->rdsup;
  float sgn = copysignf(1.0f, t1);
  const float rxpu = xxup * rpeu, rxpc = xxc * rpec, rxp3 = xx3 * rpe3;
  float drxpu, drxpc;
  if (ct) {
    drxpc = (rxp3 - rxpu) * hh;
    drxpu = (4.0f * rxpc - 3.0f * rxpu - rxp3) * hh;
  } else {
    drxpu = (rxpc - rxp3) * hh;
    drxpc = (rxp3 - 4.0f * rxpu + 3.0f * rxpc) * hh;
  }
  float mup = rxpu * gpu + peu * drxpu;
  float mc = rxpc * gpc + pec * drxpc;
  mup = sgn * MAX_(0.0f, MIN_(sgn * mup, 3.0f * fabsf(t1)));
  mc = sgn * MAX_(0.0f, MIN_(sgn * mc, 3.0f * fabsf(t1)));
  const double dt = tau_floor64((double)(k->dsup1 * (xxup + xxc) + k->dsup2 * (mup - mc)), k->mdtau);
  const double rdt = 1.0 / dt;
  float ex, U0, U1, q;
  orc_parab_moments((float)dt, &ex, &U0, &U1, &q);
  const float dss = ssc - ssup;
  float dsdtau;
  if (ct) {
    float t3 = (xx3 - xxc) * k->rdsup;
    const float sgn3 = copysignf(1.0f, t3);
    const float drxp3 = (3.0f * rxp3 - 4.0f * rxpc + rxpu) * hh;
    float m3 = rxp3 * gp3 + pe3 * drxp3;
    m3 = sgn3 * MAX_(0.0f, MIN_(sgn3 * m3, 3.0f * fabsf(t3)));
    const double dt3 = tau_floor64((double)(k->dsup1 * (xxc + xx3) + k->dsup2 * (mc - m3)), k->mdtau);
    t1 = dss * (float)rdt;
    t3 = (ss3 - ssc) / (float)dt3;
    dsdtau = 0.5f * (t1 + t3);
    dsdtau = (copysignf(1.0f, t1) + copysignf(1.0f, t3)) * MIN_(MIN_(fabsf(t1), fabsf(t3)), fabsf(0.5f * dsdtau));
  } else {
    t1 = dss * (float)rdt;
    dsdtau = t1;
    if (onesided == 1 && xxc > 0.0f) {
      dsdtau = (3.0f * ssc - 4.0f * ssup + ss3) * hh / xxc;
      sgn = copysignf(1.0f, t1);
      dsdtau = sgn * MAX_(0.0f, MIN_(sgn * dsdtau, 2.0f * fabsf(t1)));
    }
  }
  const double B = (2.0 * U1 - q) * rdt, A = U0 - B, C = q - U1;
  float I = Inuup * ex + (float)A * ssup + (float)B * ssc + (float)C * dsdtau;
  *Inu = MAX_(-k->imax, MIN_(k->imax, I));
  *plam = (k->get_rates && dt > 0 && xxc > 0) ? (float)B / xxc : 0.0f;
}

/* herm2_first (SRC/rt_solvers_mod.f90:1860-1893) */
static void herm2_first(const lch2_c *k, float xxup, float xxc, float xxdn, float ssup, float ssc, float Inuup,
                        float pec, float rpec, float rpedn, float gpc, float *Inu, float *plam)
{
  const float t1 = (xxc - xxup) * k->rdsup;
  const float sgn = copysignf(1.0f, t1);
  const float rxpc = xxc * rpec;
  float mc = rxpc * gpc + pec * (xxdn * rpedn - rxpc) * k->rdsup;
  mc = sgn * MAX_(0.0f, MIN_(sgn * mc, 3.0f * fabsf(t1)));
  const double dt = tau_floor64((double)(k->dsup1 * (xxup + xxc) + k->dsup2 * (t1 - mc)), k->mdtau);
  const double rdt = 1.0 / dt;
  float ex, U0, U1, q;
  orc_parab_moments((float)dt, &ex, &U0, &U1, &q);
  const float dsdtau = (ssc - ssup) * (float)rdt;
  const double B = (2.0 * U1 - q) * rdt, A = U0 - B, C = q - U1;
  float I = Inuup * ex + (float)A * ssup + (float)B * ssc + (float)C * dsdtau;
  *Inu = MAX_(-k->imax, MIN_(k->imax, I));
  *plam = (k->get_rates && dt > 0 && xxc > 0) ? (float)B / xxc : 0.0f;
}

void orc_long_char_hermite2_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                                const float *pe4, const float *dpe_dz, int ghost_kind, const float *gpe,
                                const float *gdpe, float *Psi)
{
  const int nf = p->nf, nx = p->nx, ny = p->ny, nz = p->nz, dir = p->dir;
  const size_t sx = (size_t)nf, sy = sx * nx, sz = sy * ny, nxy = (size_t)nx * ny;
  lch2_c k;
  const float ds = (float)p->ds[2];
  const float dsup = (float)((double)ds / fabs(p->vec[2]));
  k.rdsup = 1.0f / dsup;
  k.dsup1 = 0.5f * dsup;
  k.dsup2 = 1.0f / 12.0f * dsup * dsup;
  k.mdtau = -p->maser_dtau_max;
  k.imax = (float)p->maser_Inu_max;
  k.get_rates = p->get_rates;
  const float gfac = (float)(dir * fabs(p->vec[2]));
  const int zs = dir > 0 ? 1 : nz, ze = dir > 0 ? nz : 1;
  const int ghosted = ghost_kind == ORC_LCH_UPSTREAM || ghost_kind == ORC_LCH_BOTTOM;
#define W_(a, f, x, y, z) a[(f) + sx * (x) + sy * (y) + sz * (size_t)(z)]
#define C_(a, x, y, z) a[(x) + (size_t)nx * (y) + nxy * (size_t)((z) - 1)]
#define G_(a, x, y) a[(x) + (size_t)nx * (y)]
#define P_(f, x, y, z) Psi[(f) + sx * (x) + sy * (y) + sz * (size_t)((z) - 1)]
#define STORE(Z)                                                                                               \
  do {                                                                                                         \
    W_(I, f, x, y, (Z)) = Inu;                                                                                 \
    if (p->get_rates && Psi) P_(f, x, y, (Z)) = pl;                                                            \
  } while (0)
  /* the first plane */
  {
    const int z = zs, zup = z - dir, zdn = z + dir;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const float pec = C_(pe4, x, y, z), rpec = 1.0f / pec, gpc = gfac * C_(dpe_dz, x, y, z);
        const float pedn = C_(pe4, x, y, zdn), rpedn = 1.0f / pedn;
        for (int f = 0; f < nf; f++) {
          float Inu, pl;
          if (ghosted) {
            const float peu = G_(gpe, x, y), rpeu = 1.0f / peu, gpu = gfac * G_(gdpe, x, y);
            const float gpdn = gfac * C_(dpe_dz, x, y, zdn);
            herm2_step(&k, W_(chi, f, x, y, zup), W_(chi, f, x, y, z), W_(chi, f, x, y, zdn), W_(S, f, x, y, zup),
                       W_(S, f, x, y, z), W_(S, f, x, y, zdn), W_(I, f, x, y, zup), peu, rpeu, gpu, pec, rpec, gpc,
                       pedn, rpedn, gpdn, 1, -1, &Inu, &pl);
          } else {
            herm2_first(&k, W_(chi, f, x, y, zup), W_(chi, f, x, y, z), W_(chi, f, x, y, zdn), W_(S, f, x, y, zup),
                        W_(S, f, x, y, z), W_(I, f, x, y, zup), pec, rpec, rpedn, gpc, &Inu, &pl);
          }
          STORE(z);
        }
      }
  }
  /* interior planes */
  for (int z = zs + dir; z != ze; z += dir) {
    const int zup = z - dir, zdn = z + dir;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const float peu = C_(pe4, x, y, zup), pec = C_(pe4, x, y, z), pedn = C_(pe4, x, y, zdn);
        const float rpeu = 1.0f / peu, rpec = 1.0f / pec, rpedn = 1.0f / pedn;
        const float gpu = gfac * C_(dpe_dz, x, y, zup), gpc = gfac * C_(dpe_dz, x, y, z),
                    gpdn = gfac * C_(dpe_dz, x, y, zdn);
        for (int f = 0; f < nf; f++) {
          float Inu, pl;
          herm2_step(&k, W_(chi, f, x, y, zup), W_(chi, f, x, y, z), W_(chi, f, x, y, zdn), W_(S, f, x, y, zup),
                     W_(S, f, x, y, z), W_(S, f, x, y, zdn), W_(I, f, x, y, zup), peu, rpeu, gpu, pec, rpec, gpc,
                     pedn, rpedn, gpdn, 1, -1, &Inu, &pl);
          STORE(z);
        }
      }
  }
  /* the last plane: third sample two planes upwind, or the ghost when nz = 2 */
  {
    const int z = ze, zup = z - dir, zff = z - 2 * dir;
    const int real3 = zff >= 1 && zff <= nz;
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const float peu = C_(pe4, x, y, zup), pec = C_(pe4, x, y, z);
        const float rpeu = 1.0f / peu, rpec = 1.0f / pec;
        const float gpu = gfac * C_(dpe_dz, x, y, zup), gpc = gfac * C_(dpe_dz, x, y, z);
        float rpe3;
        if (real3) rpe3 = 1.0f / C_(pe4, x, y, zff);
        else if (ghost_kind == ORC_LCH_UPSTREAM) rpe3 = 1.0f / G_(gpe, x, y);
        else rpe3 = rpec;
        for (int f = 0; f < nf; f++) {
          float Inu, pl;
          herm2_step(&k, W_(chi, f, x, y, zup), W_(chi, f, x, y, z), W_(chi, f, x, y, zff), W_(S, f, x, y, zup),
                     W_(S, f, x, y, z), W_(S, f, x, y, zff), W_(I, f, x, y, zup), peu, rpeu, gpu, pec, rpec, gpc,
                     0.0f, rpe3, 0.0f, 0, real3 ? 1 : 0, &Inu, &pl);
          STORE(z);
        }
      }
  }
#undef STORE
#undef W_
#undef C_
#undef G_
#undef P_
}
