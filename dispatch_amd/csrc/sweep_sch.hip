// 3D short characteristics with quadratic (parabolic-moment) weights:
// short_char_hermite (SRC/rt_solvers_mod.f90:1917-2248), fused with
// integrate_intensity (SRC/rt_patch_mod.f90:3399-3431).
//
// Upwind chi, S, I and, past the slab's first plane, chi/Pe and S/B from the 4x4
// Keys stencil, each clamped from below by the inner 2x2 minimum; the step length
// is the trapezoid plus the Euler-Maclaurin term with the MC-limited target slope of
// chi (product rule through Pe); dS/dtau = S (dB/dtau)/B + B d(S/B)/dtau, clamped to
// [0, 2x] the secant; I = I_up e^-dt + A S_up + B S + C dS/dtau with A, B, C in FP64
// as the reference keeps them.  The slab's first plane takes the linear step.
//
// One workgroup per frequency owns the periodic plane: the upwind plane's I is
// double-buffered in LDS; chi, S, beta and Pe of the upwind plane come through the
// cache.  FP32 (FP64 for the step length and the weights A, B, C), as the reference.
#include "sweep_sc2.hpp"
#include <cfloat>
#include <cmath>
#include <algorithm>

#pragma clang fp contract(off)

namespace m3drt {

namespace {

struct SchArgs {
  const float *chi, *S, *gchi, *gS, *gI;
  const float *beta, *pe4, *temp, *dpe_dx, *dpe_dy, *dpe_dz, *dlnt_dx, *dlnt_dy, *dlnt_dz, *hknu;
  float *J, *PsiJ, *Iout, *Psiout, *Irem, *esc, *exitI;
  int64_t cs_f, cs_x, cs_y, cs_z, gs_f, gs_x, gs_y;
  int nf, nx, ny, nz, dir, xo, yo, split, get_rates;
  float kx[4], ky[4];
  float dsup05, rpath, dsc2, v1, v2, v3, mdtau4, imax;
  double mdtau;
  float wj;
};

__global__ void __launch_bounds__(1024) sch_sweep(const SchArgs a) {
  using namespace dev;
  extern __shared__ float sI[];  // two planes
  const int nx = a.nx, ny = a.ny, nxy = nx * ny;
  const int f = blockIdx.x;
  const float *gchi = a.gchi + f * a.gs_f, *gS = a.gS + f * a.gs_f, *gI = a.gI + f * a.gs_f;
  for (int c = threadIdx.x; c < nxy; c += blockDim.x) sI[c] = gI[(c % nx) * a.gs_x + (c / nx) * a.gs_y];
  __syncthreads();
  const float hk = a.hknu[f];
  int cur = 0;
  for (int step = 0; step < a.nz; ++step) {
    const int z = a.dir > 0 ? step : a.nz - 1 - step;
    const bool chain = step > 0;  // the ghost plane has no beta / Pe
    const int zu = z - a.dir;
    const int64_t zb = (int64_t)z * a.cs_z + (int64_t)f * a.cs_f;
    const int64_t ub = (int64_t)zu * a.cs_z + (int64_t)f * a.cs_f;
    const float *Iup = sI + cur * nxy;
    float *Inew = sI + (1 - cur) * nxy;
    for (int c = threadIdx.x; c < nxy; c += blockDim.x) {
      const int x = c % nx, y = c / nx;
      int xs[4], ys[4];
      for (int i = 0; i < 4; ++i) {
        xs[i] = wrap(x + a.xo + i, nx);
        ys[i] = wrap(y + a.yo + i, ny);
      }
      // the five upwind fields at one stencil point
      auto fields = [&](int xi, int yi, float v[5]) {
        if (!chain) {
          const int64_t g = xi * a.gs_x + yi * a.gs_y;
          v[0] = gchi[g];
          v[1] = gS[g];
          v[3] = v[4] = 0.0f;
        } else {
          const int64_t o = ub + xi * a.cs_x + yi * a.cs_y;
          v[0] = a.chi[o];
          v[1] = a.S[o];
          const float rpe = 1.0f / a.pe4[(int64_t)zu * nxy + yi * nx + xi];
          const float bet = a.beta[o];
          v[3] = v[0] * rpe;
          v[4] = v[1] * (1.0f - bet) / bet;
        }
        v[2] = Iup[yi * nx + xi];
      };
      float up[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      float mn[5] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
      for (int j = 0; j < 4; ++j) {
        float rr[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        for (int i = 0; i < 4; ++i) {
          float v[5];
          fields(xs[i], ys[j], v);
          for (int k = 0; k < 5; ++k) rr[k] = rr[k] + a.kx[i] * v[k];
          if (j >= 1 && j <= 2 && i >= 1 && i <= 2)
            for (int k = 0; k < 5; ++k) mn[k] = fminf(mn[k], v[k]);
        }
        for (int k = 0; k < 5; ++k) up[k] = up[k] + a.ky[j] * rr[k];
      }
      for (int k = 0; k < 5; ++k) up[k] = fmaxf(up[k], mn[k]);
      const float xxup = up[0], ssup = up[1], Inuup = up[2];
      const int64_t o = zb + x * a.cs_x + y * a.cs_y;
      const float xx = a.chi[o], ss = a.S[o];
      float Inu, psi_out, rem = 0.0f, escv = 1.0f;
      bool write_split = a.split;
      if (chain) {
        const int64_t cc = (int64_t)z * nxy + y * nx + x;
        const float gpe = a.v1 * a.dpe_dx[cc] + a.v2 * a.dpe_dy[cc] + a.v3 * a.dpe_dz[cc];
        const float aT = (a.v1 * a.dlnt_dx[cc] + a.v2 * a.dlnt_dy[cc] + a.v3 * a.dlnt_dz[cc]) / a.temp[cc];
        const float pec = a.pe4[cc], rpec = 1.0f / pec;
        const float rxpup = up[3], rhou = up[4], betc = a.beta[o];
        const float dchi = xx - xxup, t1 = dchi * a.rpath, sgn = copysignf(1.0f, t1);
        const float rxpc = xx * rpec;
        float mc = rxpc * gpe + pec * (rxpc - rxpup) * a.rpath;
        mc = sgn * fmaxf(0.0f, fminf(sgn * mc, 2.0f * fabsf(t1)));
        double dt = (double)(a.dsup05 * (xxup + xx) + a.dsc2 * (t1 - mc));
        dt = fmax(copysign(fmax(fabs(dt), (double)1e-20f), dt), a.mdtau);
        const double rdt = 1.0 / dt;
        const float rdt4 = (float)rdt;
        float ex, U0, U1, q;
        dev::parab_moments((float)dt, ex, U0, U1, q);
        const float r1mbc = 1.0f / (1.0f - betc), Bc = betc * r1mbc;
        const float gc = hk * r1mbc * aT / xx;
        const float rhoc = ss / Bc, t1s = (ss - ssup) * rdt4, sgns = copysignf(1.0f, t1s);
        float dsdtau = ss * gc + Bc * (rhoc - rhou) * rdt4, kf, srem;
        if (sgns * dsdtau <= 0.0f) {
          dsdtau = 0.0f; kf = 0.0f; srem = 0.0f;
        } else if (sgns * dsdtau >= 2.0f * fabsf(t1s)) {
          dsdtau = 2.0f * t1s; kf = 2.0f * rdt4; srem = -2.0f * rdt4 * ssup;
        } else {
          kf = gc + rdt4; srem = -Bc * rhou * rdt4;
        }
        const double B = (2.0 * U1 - q) * rdt, A = U0 - B, C = q - U1;
        Inu = Inuup * ex + (float)A * ssup + (float)B * ss + (float)C * dsdtau;
        Inu = fmaxf(-a.imax, fminf(a.imax, Inu));
        const float Psi_loc = (float)(B + C * (double)kf);
        if (dt > 0 && xx > 0.0f && Psi_loc >= 0.0f && Psi_loc <= 1.0f && fabsf(Inu) < a.imax) {
          psi_out = Psi_loc / xx;
          rem = Inuup * ex + (float)A * ssup + (float)C * srem;
          escv = (float)(1.0 - (B + C * (double)kf));
        } else {
          psi_out = 0.0f;
          rem = Inu;
          escv = 1.0f;
        }
        write_split = a.split && a.get_rates;
      } else {
        float dtl = fmaxf(a.dsup05 * (xxup + xx), a.mdtau4);
        float El, Wl, Ps4, Es4;
        lin_weights(dtl, El, Wl, Ps4, Es4);
        Inu = Inuup * El + Wl * ssup + Ps4 * ss;
        const bool off = dtl <= 0.0f || xx <= 0.0f || fabsf(Inu) > a.imax;
        const float Ir = Inuup * El + Wl * ssup;
        Inu = fmaxf(-a.imax, fminf(a.imax, Inu));
        if (a.split) {
          Ps4 = off ? 0.0f : Ps4;
          rem = off ? Inu : Ir;
          escv = off ? 1.0f : Es4;
        }
        psi_out = (dtl > 0.0f && xx > 0.0f) ? Ps4 / xx : 0.0f;
      }
      Inew[c] = Inu;
      if (a.J) a.J[o] = a.J[o] + a.wj * Inu;
      if (a.PsiJ) a.PsiJ[o] = a.PsiJ[o] + a.wj * psi_out;
      if (a.Iout) a.Iout[o] = Inu;
      if (a.Psiout) a.Psiout[o] = psi_out;
      if (write_split) {
        a.Irem[o] = rem;
        a.esc[o] = escv;
      }
      if (a.exitI && step == a.nz - 1) a.exitI[f * a.gs_f + x * a.gs_x + y * a.gs_y] = Inu;
    }
    __syncthreads();
    cur = 1 - cur;
  }
}

}  // namespace

int sweep_sch(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b, const m3drt_hermite_fields *h) {
  if (d->prec != M3DRT_F32) return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep_hermite: FP32 (as the reference)");
  if (b->eta_row) return set_error(M3DRT_EINVAL, "m3drt_sc_sweep_hermite: no eta flags");
  const size_t lds = 2 * (size_t)d->nx * d->ny * sizeof(float);
  if (lds > (size_t)ctx->lds_max) return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep_hermite: two I planes exceed the LDS");
  const Sc2Geom g = sc2_geometry(d);
  SchArgs a;
  a.chi = static_cast<const float *>(b->chi);
  a.S = static_cast<const float *>(b->S);
  a.gchi = static_cast<const float *>(b->ghost_chi);
  a.gS = static_cast<const float *>(b->ghost_S);
  a.gI = static_cast<const float *>(b->ghost_I);
  a.beta = h->beta; a.pe4 = h->pe4; a.temp = h->temp;
  a.dpe_dx = h->dpe_dx; a.dpe_dy = h->dpe_dy; a.dpe_dz = h->dpe_dz;
  a.dlnt_dx = h->dlnt_dx; a.dlnt_dy = h->dlnt_dy; a.dlnt_dz = h->dlnt_dz; a.hknu = h->hknu;
  a.get_rates = (d->flags & M3DRT_GET_RATES) ? 1 : 0;
  a.J = static_cast<float *>(b->J);
  a.PsiJ = a.get_rates ? static_cast<float *>(b->PsiJ) : nullptr;
  a.Iout = static_cast<float *>(b->I_out);
  a.Psiout = a.get_rates ? static_cast<float *>(b->Psi_out) : nullptr;
  a.Irem = static_cast<float *>(b->Irem_out);
  a.esc = static_cast<float *>(b->escape_out);
  a.exitI = static_cast<float *>(b->exit_I);
  a.cs_f = d->cs_f; a.cs_x = d->cs_x; a.cs_y = d->cs_y; a.cs_z = d->cs_z;
  a.gs_f = d->gs_f; a.gs_x = d->gs_x; a.gs_y = d->gs_y;
  a.nf = d->nf; a.nx = d->nx; a.ny = d->ny; a.nz = d->nz; a.dir = d->dir > 0 ? 1 : -1;
  a.xo = g.xo; a.yo = g.yo;
  a.split = (d->flags & M3DRT_SPLIT_RATES) ? 1 : 0;
  for (int i = 0; i < 4; ++i) {
    a.kx[i] = (float)g.kx[i];
    a.ky[i] = (float)g.ky[i];
  }
  a.dsup05 = 0.5f * (float)g.dsup;
  a.rpath = 0.5f / a.dsup05;
  a.dsc2 = a.dsup05 * a.dsup05 / 3.0f;
  const double sg = d->vec[2] >= 0 ? 1.0 : -1.0, dr = d->dir > 0 ? 1.0 : -1.0;
  a.v1 = (float)(d->vec[0] * dr * sg);
  a.v2 = (float)(d->vec[1] * dr * sg);
  a.v3 = (float)(d->vec[2] * dr * sg);
  a.mdtau = -d->maser_dtau_max;
  a.mdtau4 = -(float)d->maser_dtau_max;
  a.imax = (float)d->maser_Inu_max;
  a.wj = (float)d->wj;
  const int nt = (int)std::min<int64_t>(1024, ((int64_t)d->nx * d->ny + 63) / 64 * 64);
  M3DRT_HIP(hipFuncSetAttribute((const void *)sch_sweep, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(sch_sweep, dim3(d->nf), dim3(nt), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
