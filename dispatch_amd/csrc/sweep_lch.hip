// Column sweeps with the cubic-Bezier source integration of long_char_hermite
// (SRC/rt_solvers_mod.f90:1052-1504) and the hybrid long_char_hermite2 (:1536-1895),
// fused with integrate_intensity
// (SRC/rt_patch_mod.f90:3399-3431).  Each column is its own long characteristic (the
// output pass applies it along columns, after rotate_array for inclined rays).
//
// One lane per (frequency, column) walks the column holding three planes of samples
// in registers: chi, S, beta of its frequency and Pe, 1/Pe, dPe/ds, dlnT/ds / T of
// the cell.  Every sample is read once; per plane the lane writes I (and Psi) and
// updates J (and PsiJ).  Lanes run along the unit-stride axis of the layout, so a
// wave's reads of a plane are one contiguous segment.  FP32 with every operation
// separately rounded, as the reference's build.
#include "m3drt_internal.hpp"
#include <cfloat>
#include <cmath>

#pragma clang fp contract(off)

namespace m3drt {

namespace {

struct LchArgs {
  const float *chi, *S, *gchi, *gS, *gI, *beta;
  const float *pe4, *temp, *dpe_dz, *dlnt_dz, *hknu;
  const float *gpe, *gdpe, *gtemp, *gdlnt;
  float *J, *PsiJ, *Iout, *Psiout, *exitI;
  int64_t cs_f, cs_x, cs_y, cs_z, gs_f, gs_x, gs_y;
  int nf, nx, ny, nz, dir, f_fast, kind, get_rates;
  float rdsup, dsup1, dsup2, mdtau, imax, wj, gfac, third;
  double mdtau64;
};

// one sample of the column: the frequency's chi, S, beta; the cell's Pe, 1/Pe,
// dPe/ds and dlnT/ds / T (ds along the ray); and the quotients every step that
// touches the sample forms from them (1/(1-beta), B, rho = S/B, 1/chi, chi/Pe), formed
// once when the sample enters the window
struct Samp {
  float chi, S, beta, pe, rpe, gp, aT;
  float r1mb, B, rho, rx, rxp;
};

__device__ __forceinline__ void derive(Samp &p) {
  p.r1mb = 1.0f / (1.0f - p.beta);
  p.B = p.beta * p.r1mb;
  p.rho = p.S / p.B;
  p.rx = 1.0f / p.chi;
  p.rxp = p.chi * p.rpe;
}

// P2 = w2/t and P3 = w3/t of the Bernstein weights on |t| <= 3, minimax of degree
// 10 and 9 (the reference's fits, SRC/rt_solvers_mod.f90:1327-1352)
__device__ __forceinline__ float b3_p2(float t) {
  return 2.500000000e-01f + t * (-1.000000238e-01f + t * (2.499995567e-02f + t * (-4.761847202e-03f +
         t * (7.440865738e-04f + t * (-9.924791084e-05f + t * (1.156294547e-05f + t * (-1.190877015e-06f +
         t * (1.137804801e-07f + t * (-1.100420821e-08f + t * 7.658396139e-10f)))))))));
}
__device__ __forceinline__ float b3_p3(float t) {
  return 2.500000000e-01f + t * (-4.999998957e-02f + t * (8.333289996e-03f + t * (-1.190488110e-03f +
         t * (1.488493726e-04f + t * (-1.653099389e-05f + t * (1.640762775e-06f + t * (-1.503805294e-07f +
         t * (1.414234507e-08f + t * -1.025371787e-09f))))))));
}

// bezier3_weights (SRC/rt_solvers_mod.f90:1301-1323): w_k = t int_0^1 b_k(u) e^-tu du
// for the Bernstein basis u^3, 3u^2(1-u), 3u(1-u)^2, (1-u)^3 (S_up, the two control
// points, S_target).  Below t = 3 the series (w0, w1 by the reflection u -> 1-u);
// above, the moments M_n = n M_(n-1) - t^n E from M0 = 1 - E.  E and M0 from
// dev::exp_neg.
__device__ __forceinline__ void bezier3(float t, float &E, float &w0, float &w1, float &w2, float &w3) {
  float M0;
  dev::exp_neg(t, E, M0);
  const float tb = fmaxf(t, 3.0f), tb2 = tb * tb, tb3 = tb2 * tb, r3 = 1.0f / tb3;
  const float M1 = M0 - tb * E;
  const float M2 = 2.0f * M1 - tb2 * E;
  const float M3 = 3.0f * M2 - tb3 * E;
  const float b0 = M3 * r3;
  const float b1 = 3.0f * (M2 * tb - M3) * r3;
  const float b2 = 3.0f * (M1 * tb2 - 2.0f * M2 * tb + M3) * r3;
  if (t <= 3.0f) {
    const float tE = t * E;
    w0 = tE * b3_p3(-t);
    w1 = tE * b3_p2(-t);
    w2 = t * b3_p2(t);
    w3 = t * b3_p3(t);
  } else {
    w0 = b0;
    w1 = b1;
    w2 = b2;
    w3 = M0 - b0 - b1 - b2;
  }
}

__device__ __forceinline__ float tau_floor(float dt, float mdtau) {
  return fmaxf(copysignf(fmaxf(fabsf(dt), 1e-20f), dt), mdtau);
}

// herm_dtau + bezier3_weights + herm_source (SRC/rt_solvers_mod.f90:1365-1459): the
// interval [up, c] with the parabola's third sample t3 downwind (center) or two
// planes upwind (the last plane)
__device__ __forceinline__ void herm_step(const LchArgs &a, const Samp &u, const Samp &c, const Samp &t3,
                                          bool center, float hk, float Iup, float &Inu, float &plam) {
  const float hh = 0.5f * a.rdsup;
  const float t1 = (c.chi - u.chi) * a.rdsup;
  const float sgn = copysignf(1.0f, t1);
  const float rxpu = u.rxp, rxpc = c.rxp, rxp3 = t3.rxp;
  float drxpu, drxpc;
  if (center) {
    drxpc = (rxp3 - rxpu) * hh;
    drxpu = (4.0f * rxpc - 3.0f * rxpu - rxp3) * hh;
  } else {
    drxpu = (rxpc - rxp3) * hh;
    drxpc = (rxp3 - 4.0f * rxpu + 3.0f * rxpc) * hh;
  }
  float mup = rxpu * u.gp + u.pe * drxpu;
  float mc = rxpc * c.gp + c.pe * drxpc;
  mup = sgn * fmaxf(0.0f, fminf(sgn * mup, 2.0f * fabsf(t1)));
  mc = sgn * fmaxf(0.0f, fminf(sgn * mc, 2.0f * fabsf(t1)));
  const float dt = tau_floor(a.dsup1 * (u.chi + c.chi) + a.dsup2 * (mup - mc), a.mdtau);
  float ex, w0, w1, w2, w3;
  bezier3(dt, ex, w0, w1, w2, w3);
  // rho = S/B at the three samples and its parabola's slopes at both ends
  const float r1mbu = u.r1mb, r1mbc = c.r1mb;
  const float Bu = u.B, Bc = c.B;
  const float rhou = u.rho, rhoc = c.rho, rho3 = t3.rho;
  float drhu, drhc;
  if (center) {
    drhc = (rho3 - rhou) * hh;
    drhu = (4.0f * rhoc - 3.0f * rhou - rho3) * hh;
  } else {
    drhu = (rhoc - rho3) * hh;
    drhc = (rho3 - 4.0f * rhou + 3.0f * rhoc) * hh;
  }
  const float rxxu = u.rx, rxxc = c.rx;
  const float gu = hk * r1mbu * u.aT * rxxu;   // (dB/dtau)/B at up
  const float gc = hk * r1mbc * c.aT * rxxc;   // and at c
  const float dt3 = dt * a.third;
  float P1 = u.S + dt3 * (u.S * gu + Bu * drhu * rxxu);
  float P2 = c.S - dt3 * (c.S * gc + Bc * drhc * rxxc);
  const float smin = fminf(u.S, c.S), smax = fmaxf(u.S, c.S);
  P1 = fminf(fmaxf(P1, smin), smax);
  P2 = fminf(fmaxf(P2, smin), smax);
  const float I = Iup * ex + w0 * u.S + w1 * P1 + w2 * P2 + w3 * c.S;
  Inu = fmaxf(-a.imax, fminf(a.imax, I));
  const float d1 = fminf(fmaxf(Bu * a.third / Bc, 0.0f), 1.0f);
  const float d2 = fminf(fmaxf(2.0f * a.third - dt3 * gc, 0.0f), 1.0f);
  plam = (a.get_rates && dt > 0.0f && c.chi > 0.0f) ? (w1 * d1 + w2 * d2 + w3) * rxxc : 0.0f;
}

// herm_first (SRC/rt_solvers_mod.f90:1465-1520): [ghost, first plane] below a zeroed
// top ghost, which has no thermodynamic state: the upwind-end slopes are secants
__device__ __forceinline__ void herm_first(const LchArgs &a, float xxup, float ssup, const Samp &c, const Samp &d,
                                           float hk, float Iup, float &Inu, float &plam) {
  const float t1 = (c.chi - xxup) * a.rdsup;
  const float sgn = copysignf(1.0f, t1);
  const float rxpc = c.rxp;
  float mc = rxpc * c.gp + c.pe * (d.rxp - rxpc) * a.rdsup;
  mc = sgn * fmaxf(0.0f, fminf(sgn * mc, 2.0f * fabsf(t1)));
  const float dt = tau_floor(a.dsup1 * (xxup + c.chi) + a.dsup2 * (t1 - mc), a.mdtau);
  float ex, w0, w1, w2, w3;
  bezier3(dt, ex, w0, w1, w2, w3);
  const float r1mbc = c.r1mb;
  const float Bc = c.B;
  const float Gc = hk * r1mbc * c.aT / c.chi;
  const float rhoc = c.rho, rhodn = d.rho;
  float taudn = a.dsup1 * (c.chi + d.chi);
  taudn = copysignf(fmaxf(fabsf(taudn), 1e-20f), taudn);
  const float dSc = c.S * Gc + Bc * (rhodn - rhoc) / taudn;
  const float dSu = (c.S - ssup) / dt;
  const float dt3 = dt * a.third;
  float P1 = ssup + dt3 * dSu;
  float P2 = c.S - dt3 * dSc;
  const float smin = fminf(ssup, c.S), smax = fmaxf(ssup, c.S);
  P1 = fminf(fmaxf(P1, smin), smax);
  P2 = fminf(fmaxf(P2, smin), smax);
  const float I = Iup * ex + w0 * ssup + w1 * P1 + w2 * P2 + w3 * c.S;
  Inu = fmaxf(-a.imax, fminf(a.imax, I));
  plam = (a.get_rates && dt > 0.0f && c.chi > 0.0f) ? (w1 * a.third + w2 + w3) / c.chi : 0.0f;
}

__global__ void __launch_bounds__(256) lch_sweep(const LchArgs a) {
  const int64_t nxy = (int64_t)a.nx * a.ny;
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= nxy * a.nf) return;
  int f, col;
  if (a.f_fast) {
    f = (int)(id % a.nf);
    col = (int)(id / a.nf);
  } else {
    col = (int)(id % nxy);
    f = (int)(id / nxy);
  }
  const int x = col % a.nx, y = col / a.nx;
  const int64_t o = f * a.cs_f + x * a.cs_x + y * a.cs_y;
  const int64_t go = f * a.gs_f + x * a.gs_x + y * a.gs_y;
  const float hk = a.hknu[f];
  // step s along the ray -> storage plane
  auto zof = [&](int s) { return a.dir > 0 ? s : a.nz - 1 - s; };
  auto load = [&](int s) {
    const int z = zof(s);
    const int64_t oz = o + (int64_t)z * a.cs_z;
    const int64_t cz = col + (int64_t)z * nxy;
    Samp p;
    p.chi = a.chi[oz];
    p.S = a.S[oz];
    p.beta = a.beta[oz];
    p.pe = a.pe4[cz];
    p.rpe = 1.0f / p.pe;
    p.gp = a.gfac * a.dpe_dz[cz];
    p.aT = a.gfac * a.dlnt_dz[cz] / a.temp[cz];
    derive(p);
    return p;
  };
  auto emit = [&](int s, float Inu, float plam) {
    const int64_t oz = o + (int64_t)zof(s) * a.cs_z;
    if (a.J) a.J[oz] = a.J[oz] + a.wj * Inu;
    if (a.PsiJ) a.PsiJ[oz] = a.PsiJ[oz] + a.wj * plam;
    if (a.Iout) a.Iout[oz] = Inu;
    if (a.Psiout) a.Psiout[oz] = plam;
  };
  // the ghost plane below the first step; with a thermodynamic state unless the top
  Samp g;
  g.chi = a.gchi[go];
  g.S = a.gS[go];
  g.beta = g.pe = g.rpe = g.gp = g.aT = 0.0f;
  g.r1mb = g.B = g.rho = g.rx = g.rxp = 0.0f;
  if (a.kind != M3DRT_LCH_TOP) {
    g.pe = a.gpe[col];
    g.rpe = 1.0f / g.pe;
    g.gp = a.gfac * a.gdpe[col];
    const float rT = 1.0f / a.gtemp[col];
    g.aT = a.gfac * a.gdlnt[col] * rT;
    g.beta = fmaxf(expf(-hk * rT), FLT_MIN);
    derive(g);
  }
  Samp pm = g, p0 = load(0), p1 = load(1);
  float Inu, plam;
  if (a.kind == M3DRT_LCH_TOP)
    herm_first(a, g.chi, g.S, p0, p1, hk, a.gI[go], Inu, plam);
  else
    herm_step(a, g, p0, p1, true, hk, a.gI[go], Inu, plam);
  emit(0, Inu, plam);
  for (int s = 1; s < a.nz - 1; ++s) {
    pm = p0;
    p0 = p1;
    p1 = load(s + 1);
    herm_step(a, pm, p0, p1, true, hk, Inu, Inu, plam);
    emit(s, Inu, plam);
  }
  // the last plane: the parabola centred at its upwind point, third sample two
  // planes upwind (the ghost when nz = 2: its own state for a neighbour's plane,
  // else the target's Pe and beta)
  if (a.nz == 2) {
    if (a.kind == M3DRT_LCH_UPSTREAM) {
      pm.beta = fmaxf(expf(-hk / a.gtemp[col]), FLT_MIN);
    } else {
      pm.rpe = p1.rpe;
      pm.beta = p1.beta;
    }
    derive(pm);
  }
  herm_step(a, p0, p1, pm, false, hk, Inu, Inu, plam);
  emit(a.nz - 1, Inu, plam);
  if (a.exitI) a.exitI[go] = Inu;
}

// long_char_hermite2 (SRC/rt_solvers_mod.f90:1536-1895): long_char_hermite's optical
// depth (product-rule chi slopes, here clamped to 3x the secant) with
// long_char_quadratic's source treatment, the MC-limited dS/dtau at the target from
// the secants of the two intervals beside it (the second interval's optical depth
// from the same parabola of chi/Pe) and the parab_moments weights, A, B, C in FP64
// as the reference keeps them.  onesided: the last plane's third sample is real
// (backward one-sided source slope) or the ghost (secant).
__device__ __forceinline__ double tau_floor64(double dt, double mdtau) {
  return fmax(copysign(fmax(fabs(dt), (double)1e-20f), dt), mdtau);
}

__device__ __forceinline__ void herm2_step(const LchArgs &a, const Samp &u, const Samp &c, const Samp &t3,
                                           bool center, bool onesided, float Iup, float &Inu, float &plam) {
  const float hh = 0.5f * a.rdsup;
  float t1 = (c.chi - u.chi) * a.rdsup;
  float sgn = copysignf(1.0f, t1);
  const float rxpu = u.chi * u.rpe, rxpc = c.chi * c.rpe, rxp3 = t3.chi * t3.rpe;
  float drxpu, drxpc;
  if (center) {
    drxpc = (rxp3 - rxpu) * hh;
    drxpu = (4.0f * rxpc - 3.0f * rxpu - rxp3) * hh;
  } else {
    drxpu = (rxpc - rxp3) * hh;
    drxpc = (rxp3 - 4.0f * rxpu + 3.0f * rxpc) * hh;
  }
  float mup = rxpu * u.gp + u.pe * drxpu;
  float mc = rxpc * c.gp + c.pe * drxpc;
  mup = sgn * fmaxf(0.0f, fminf(sgn * mup, 3.0f * fabsf(t1)));
  mc = sgn * fmaxf(0.0f, fminf(sgn * mc, 3.0f * fabsf(t1)));
  const double mdt = a.mdtau64;
  const double dt = tau_floor64((double)(a.dsup1 * (u.chi + c.chi) + a.dsup2 * (mup - mc)), mdt);
  const double rdt = 1.0 / dt;
  float ex, U0, U1, q;
  dev::parab_moments((float)dt, ex, U0, U1, q);
  const float dss = c.S - u.S;
  float dsdtau;
  if (center) {
    float t3s = (t3.chi - c.chi) * a.rdsup;
    const float sgn3 = copysignf(1.0f, t3s);
    const float drxp3 = (3.0f * rxp3 - 4.0f * rxpc + rxpu) * hh;
    float m3 = rxp3 * t3.gp + t3.pe * drxp3;
    m3 = sgn3 * fmaxf(0.0f, fminf(sgn3 * m3, 3.0f * fabsf(t3s)));
    const double dt3 = tau_floor64((double)(a.dsup1 * (c.chi + t3.chi) + a.dsup2 * (mc - m3)), mdt);
    t1 = dss * (float)rdt;
    t3s = (t3.S - c.S) / (float)dt3;
    dsdtau = 0.5f * (t1 + t3s);
    dsdtau = (copysignf(1.0f, t1) + copysignf(1.0f, t3s)) * fminf(fminf(fabsf(t1), fabsf(t3s)), fabsf(0.5f * dsdtau));
  } else {
    t1 = dss * (float)rdt;
    dsdtau = t1;
    if (onesided && c.chi > 0.0f) {
      dsdtau = (3.0f * c.S - 4.0f * u.S + t3.S) * hh / c.chi;
      sgn = copysignf(1.0f, t1);
      dsdtau = sgn * fmaxf(0.0f, fminf(sgn * dsdtau, 2.0f * fabsf(t1)));
    }
  }
  const double B = (2.0 * (double)U1 - (double)q) * rdt, A = (double)U0 - B, C = (double)q - (double)U1;
  const float I = Iup * ex + (float)A * u.S + (float)B * c.S + (float)C * dsdtau;
  Inu = fmaxf(-a.imax, fminf(a.imax, I));
  plam = (a.get_rates && dt > 0.0 && c.chi > 0.0f) ? (float)B / c.chi : 0.0f;
}

// herm2_first (SRC/rt_solvers_mod.f90:1860-1893): below the zeroed top ghost
__device__ __forceinline__ void herm2_first(const LchArgs &a, float xxup, float ssup, const Samp &c, const Samp &d,
                                            float Iup, float &Inu, float &plam) {
  const float t1 = (c.chi - xxup) * a.rdsup;
  const float sgn = copysignf(1.0f, t1);
  const float rxpc = c.chi * c.rpe;
  float mc = rxpc * c.gp + c.pe * (d.chi * d.rpe - rxpc) * a.rdsup;
  mc = sgn * fmaxf(0.0f, fminf(sgn * mc, 3.0f * fabsf(t1)));
  const double dt = tau_floor64((double)(a.dsup1 * (xxup + c.chi) + a.dsup2 * (t1 - mc)), a.mdtau64);
  const double rdt = 1.0 / dt;
  float ex, U0, U1, q;
  dev::parab_moments((float)dt, ex, U0, U1, q);
  const float dsdtau = (c.S - ssup) * (float)rdt;
  const double B = (2.0 * (double)U1 - (double)q) * rdt, A = (double)U0 - B, C = (double)q - (double)U1;
  const float I = Iup * ex + (float)A * ssup + (float)B * c.S + (float)C * dsdtau;
  Inu = fmaxf(-a.imax, fminf(a.imax, I));
  plam = (a.get_rates && dt > 0.0 && c.chi > 0.0f) ? (float)B / c.chi : 0.0f;
}

__global__ void __launch_bounds__(256) lch2_sweep(const LchArgs a) {
  const int64_t nxy = (int64_t)a.nx * a.ny;
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= nxy * a.nf) return;
  int f, col;
  if (a.f_fast) {
    f = (int)(id % a.nf);
    col = (int)(id / a.nf);
  } else {
    col = (int)(id % nxy);
    f = (int)(id / nxy);
  }
  const int x = col % a.nx, y = col / a.nx;
  const int64_t o = f * a.cs_f + x * a.cs_x + y * a.cs_y;
  const int64_t go = f * a.gs_f + x * a.gs_x + y * a.gs_y;
  auto zof = [&](int s) { return a.dir > 0 ? s : a.nz - 1 - s; };
  auto load = [&](int s) {
    const int z = zof(s);
    const int64_t oz = o + (int64_t)z * a.cs_z;
    const int64_t cz = col + (int64_t)z * nxy;
    Samp p;
    p.chi = a.chi[oz];
    p.S = a.S[oz];
    p.beta = p.aT = 0.0f;
    p.pe = a.pe4[cz];
    p.rpe = 1.0f / p.pe;
    p.gp = a.gfac * a.dpe_dz[cz];
    return p;
  };
  auto emit = [&](int s, float Inu, float plam) {
    const int64_t oz = o + (int64_t)zof(s) * a.cs_z;
    if (a.J) a.J[oz] = a.J[oz] + a.wj * Inu;
    if (a.PsiJ) a.PsiJ[oz] = a.PsiJ[oz] + a.wj * plam;
    if (a.Iout) a.Iout[oz] = Inu;
    if (a.Psiout) a.Psiout[oz] = plam;
  };
  Samp g;
  g.chi = a.gchi[go];
  g.S = a.gS[go];
  g.beta = g.aT = g.pe = g.rpe = g.gp = 0.0f;
  if (a.kind != M3DRT_LCH_TOP) {
    g.pe = a.gpe[col];
    g.rpe = 1.0f / g.pe;
    g.gp = a.gfac * a.gdpe[col];
  }
  Samp pm = g, p0 = load(0), p1 = load(1);
  float Inu, plam;
  if (a.kind == M3DRT_LCH_TOP)
    herm2_first(a, g.chi, g.S, p0, p1, a.gI[go], Inu, plam);
  else
    herm2_step(a, g, p0, p1, true, false, a.gI[go], Inu, plam);
  emit(0, Inu, plam);
  for (int s = 1; s < a.nz - 1; ++s) {
    pm = p0;
    p0 = p1;
    p1 = load(s + 1);
    herm2_step(a, pm, p0, p1, true, false, Inu, Inu, plam);
    emit(s, Inu, plam);
  }
  // the last plane: third sample two planes upwind; at nz = 2 the ghost, with the
  // neighbour's Pe or else the target's
  if (a.nz == 2 && a.kind != M3DRT_LCH_UPSTREAM) pm.rpe = p1.rpe;
  herm2_step(a, p0, p1, pm, false, a.nz >= 3, Inu, Inu, plam);
  emit(a.nz - 1, Inu, plam);
  if (a.exitI) a.exitI[go] = Inu;
}

}  // namespace

int sweep_lch(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b, const m3drt_hermite_fields *h,
              const m3drt_lc_ghost *g, bool hermite2) {
  LchArgs a;
  a.chi = static_cast<const float *>(b->chi);
  a.S = static_cast<const float *>(b->S);
  a.gchi = static_cast<const float *>(b->ghost_chi);
  a.gS = static_cast<const float *>(b->ghost_S);
  a.gI = static_cast<const float *>(b->ghost_I);
  a.beta = h->beta;
  a.pe4 = h->pe4;
  a.temp = h->temp;
  a.dpe_dz = h->dpe_dz;
  a.dlnt_dz = h->dlnt_dz;
  a.hknu = h->hknu;
  a.kind = g->kind;
  a.gpe = g->pe;
  a.gdpe = g->dpe_dz;
  a.gtemp = g->temp;
  a.gdlnt = g->dlnt_dz;
  a.get_rates = (d->flags & M3DRT_GET_RATES) ? 1 : 0;
  a.J = static_cast<float *>(b->J);
  a.PsiJ = a.get_rates ? static_cast<float *>(b->PsiJ) : nullptr;
  a.Iout = static_cast<float *>(b->I_out);
  a.Psiout = a.get_rates ? static_cast<float *>(b->Psi_out) : nullptr;
  a.exitI = static_cast<float *>(b->exit_I);
  a.cs_f = d->cs_f; a.cs_x = d->cs_x; a.cs_y = d->cs_y; a.cs_z = d->cs_z;
  a.gs_f = d->gs_f; a.gs_x = d->gs_x; a.gs_y = d->gs_y;
  a.nf = d->nf; a.nx = d->nx; a.ny = d->ny; a.nz = d->nz; a.dir = d->dir > 0 ? 1 : -1;
  a.f_fast = d->cs_f <= d->cs_x ? 1 : 0;
  // SRC/rt_solvers_mod.f90:1081-1090: the path per plane in real(4); gradients per cm
  // along +z turned into d/ds along the ray
  const float ds = (float)d->ds[2];
  const float dsup = (float)((double)ds / fabs(d->vec[2]));
  a.rdsup = 1.0f / dsup;
  a.dsup1 = 0.5f * dsup;
  a.dsup2 = 1.0f / 12.0f * dsup * dsup;
  a.gfac = (float)(a.dir * fabs(d->vec[2]));
  a.third = (float)(1.0 / 3.0);
  a.mdtau = -(float)d->maser_dtau_max;
  a.mdtau64 = -d->maser_dtau_max;
  a.imax = (float)d->maser_Inu_max;
  a.wj = (float)d->wj;
  const int64_t n = (int64_t)d->nx * d->ny * d->nf;
  if (hermite2)
    hipLaunchKernelGGL(lch2_sweep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL(lch_sweep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
