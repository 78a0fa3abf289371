// 3D short characteristics with the 'linear' Multi3D-comparison step of
// short_char (SRC/rt_solvers_mod.f90:19-189), fused with integrate_intensity
// (SRC/rt_patch_mod.f90:3399-3431): chi, S and I of the upwind point all from the
// 4x4 Keys stencil, each clamped from below by the inner 2x2 minimum
// (SRC/rt_solvers_mod.f90:89-134), dtau = dsup (chi_up + chi) / 2 and the linear step.
//
// One workgroup per frequency owns the periodic plane; the upwind plane's I is
// double-buffered in LDS (the new plane is written while the old one is read), chi
// and S of the upwind plane come through the cache from the cube itself (or the
// ghost plane).  A comparison solver, kept simple: a grid-stride loop over the
// plane's cells per plane.  FP32, as the reference.
#include "sweep_sc2.hpp"
#include <cmath>
#include <algorithm>

#pragma clang fp contract(off)

namespace m3drt {

namespace {

struct Sc1Args {
  const float *chi, *S, *gchi, *gS, *gI;
  float *J, *PsiJ, *Iout, *Psiout, *Irem, *esc, *exitI;
  int64_t cs_f, cs_x, cs_y, cs_z, gs_f, gs_x, gs_y;
  int nf, nx, ny, nz, dir, xo, yo, split;
  float kx[4], ky[4];
  float dsup05, mdtau, imax, wj;
};

__global__ void __launch_bounds__(1024) sc1_sweep(const Sc1Args a) {
  using namespace dev;
  extern __shared__ float sI[];  // two planes
  const int nx = a.nx, ny = a.ny, nxy = nx * ny;
  const int f = blockIdx.x;
  const float *gchi = a.gchi + f * a.gs_f, *gS = a.gS + f * a.gs_f, *gI = a.gI + f * a.gs_f;
  for (int c = threadIdx.x; c < nxy; c += blockDim.x) {
    const int x = c % nx, y = c / nx;
    sI[c] = gI[x * a.gs_x + y * a.gs_y];
  }
  __syncthreads();
  int cur = 0;
  for (int step = 0; step < a.nz; ++step) {
    const int z = a.dir > 0 ? step : a.nz - 1 - step;
    const bool ghost = step == 0;
    const int64_t zb = (int64_t)z * a.cs_z + (int64_t)f * a.cs_f;
    const int64_t ub = (int64_t)(z - a.dir) * a.cs_z + (int64_t)f * a.cs_f;
    const float *Iup = sI + cur * nxy;
    float *Inew = sI + (1 - cur) * nxy;
    for (int c = threadIdx.x; c < nxy; c += blockDim.x) {
      const int x = c % nx, y = c / nx;
      int xs[4], ys[4];
      for (int i = 0; i < 4; ++i) {
        xs[i] = wrap(x + a.xo + i, nx);
        ys[i] = wrap(y + a.yo + i, ny);
      }
      auto chiup = [&](int xi, int yi) {
        return ghost ? gchi[xi * a.gs_x + yi * a.gs_y] : a.chi[ub + xi * a.cs_x + yi * a.cs_y];
      };
      auto sup = [&](int xi, int yi) {
        return ghost ? gS[xi * a.gs_x + yi * a.gs_y] : a.S[ub + xi * a.cs_x + yi * a.cs_y];
      };
      // row pass in x, then the y pass, as the reference accumulates them
      float wc = 0.0f, ws = 0.0f, wi = 0.0f;
      for (int j = 0; j < 4; ++j) {
        float rc = 0.0f, rs = 0.0f, ri = 0.0f;
        for (int i = 0; i < 4; ++i) {
          rc = rc + a.kx[i] * chiup(xs[i], ys[j]);
          rs = rs + a.kx[i] * sup(xs[i], ys[j]);
          ri = ri + a.kx[i] * Iup[ys[j] * nx + xs[i]];
        }
        wc = wc + a.ky[j] * rc;
        ws = ws + a.ky[j] * rs;
        wi = wi + a.ky[j] * ri;
      }
      float mc = 3.402823466e38f, ms = 3.402823466e38f, mi = 3.402823466e38f;
      for (int j = 1; j <= 2; ++j)
        for (int i = 1; i <= 2; ++i) {
          mc = fminf(mc, chiup(xs[i], ys[j]));
          ms = fminf(ms, sup(xs[i], ys[j]));
          mi = fminf(mi, Iup[ys[j] * nx + xs[i]]);
        }
      const float xxup = fmaxf(wc, mc), ssup = fmaxf(ws, ms), Inuup = fmaxf(wi, mi);
      const int64_t o = zb + x * a.cs_x + y * a.cs_y;
      const float xx = a.chi[o], ss = a.S[o];
      float dt = a.dsup05 * (xxup + xx);
      dt = fmaxf(dt, a.mdtau);
      float E, W, Ps, Es;
      lin_weights(dt, E, W, Ps, Es);
      float Inu = Inuup * E + W * ssup + Ps * ss;
      if (a.split) {
        const bool off = dt <= 0.0f || xx <= 0.0f || fabsf(Inu) > a.imax;
        const float Ir = Inuup * E + W * ssup;
        Inu = fmaxf(-a.imax, fminf(a.imax, Inu));
        Ps = off ? 0.0f : Ps;
        a.Irem[o] = off ? Inu : Ir;
        a.esc[o] = off ? 1.0f : Es;
      } else {
        Inu = fmaxf(-a.imax, fminf(a.imax, Inu));
      }
      const float psi = (dt > 0.0f && xx > 0.0f) ? Ps / xx : 0.0f;
      Inew[c] = Inu;
      if (a.J) a.J[o] = a.J[o] + a.wj * Inu;
      if (a.PsiJ) a.PsiJ[o] = a.PsiJ[o] + a.wj * psi;
      if (a.Iout) a.Iout[o] = Inu;
      if (a.Psiout) a.Psiout[o] = psi;
      if (a.exitI && step == a.nz - 1) a.exitI[f * a.gs_f + x * a.gs_x + y * a.gs_y] = Inu;
    }
    __syncthreads();
    cur = 1 - cur;
  }
}

}  // namespace

int sweep_sc1(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  if (d->prec != M3DRT_F32) return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep: short_char is FP32 (as the reference)");
  if (b->eta_row) return set_error(M3DRT_EINVAL, "m3drt_sc_sweep: short_char takes no eta flags");
  const size_t lds = 2 * (size_t)d->nx * d->ny * sizeof(float);
  if (lds > (size_t)ctx->lds_max) return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep: short_char: two I planes exceed the LDS");
  const Sc2Geom g = sc2_geometry(d);
  Sc1Args a;
  a.chi = static_cast<const float *>(b->chi);
  a.S = static_cast<const float *>(b->S);
  a.gchi = static_cast<const float *>(b->ghost_chi);
  a.gS = static_cast<const float *>(b->ghost_S);
  a.gI = static_cast<const float *>(b->ghost_I);
  a.J = static_cast<float *>(b->J);
  a.PsiJ = (d->flags & M3DRT_GET_RATES) ? static_cast<float *>(b->PsiJ) : nullptr;
  a.Iout = static_cast<float *>(b->I_out);
  a.Psiout = (d->flags & M3DRT_GET_RATES) ? static_cast<float *>(b->Psi_out) : nullptr;
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
  a.mdtau = -(float)d->maser_dtau_max;
  a.imax = (float)d->maser_Inu_max;
  a.wj = (float)d->wj;
  const int nt = (int)std::min<int64_t>(1024, ((int64_t)d->nx * d->ny + 63) / 64 * 64);
  M3DRT_HIP(hipFuncSetAttribute((const void *)sc1_sweep, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(sc1_sweep, dim3(d->nf), dim3(nt), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
