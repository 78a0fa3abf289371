// Column sweep with the quadratic (Bezier-type) source integration of
// long_char_quadratic + quad_step (SRC/rt_solvers_mod.f90:891-1020), fused with
// integrate_intensity (SRC/rt_patch_mod.f90:3399-3431).  Each column is its own long
// characteristic (the reference applies it along columns, after rotate_array for
// inclined rays).  One thread per (frequency, column) walks the column holding the
// three-plane window of chi and S in registers; lanes run along the unit-stride
// axis of the layout.  FP32, as the reference.  The same file holds the column form
// of long_char_linear (SRC/rt_solvers_mod.f90:666-757).
#include "m3drt_internal.hpp"
#include <cmath>
#include <algorithm>

#pragma clang fp contract(off)

namespace m3drt {

namespace {

struct LcqArgs {
  const float *chi, *S, *gchi, *gS, *gI;
  float *J, *PsiJ, *Iout, *Psiout, *exitI;
  int64_t cs_f, cs_x, cs_y, cs_z, gs_f, gs_x, gs_y;
  int nf, nx, ny, nz, dir, f_fast;
  float rdsup, dsup1, dsup2, mdtau, imax, wj;
};

// planes of chi and S a lane of the linear column sweep holds in flight ahead of the
// one it steps onto.  The linear step is short enough for the load latency to show
// (1.58 -> 1.46 ms per 848M-point ray); the quadratic and hermite steps are bound by
// their arithmetic, and the same ring made them slower (3.21 -> 3.37 ms quadratic,
// on the same box).
constexpr int kLcPF = 4;

// the monotonised (MC-limited) centred slope of two one-sided differences
__device__ __forceinline__ float mc_slope(float t1, float t3) {
  const float t2 = 0.5f * (t1 + t3);
  return (copysignf(1.0f, t1) + copysignf(1.0f, t3)) * fminf(fminf(fabsf(t1), fabsf(t3)), fabsf(0.5f * t2));
}

// quad_step (SRC/rt_solvers_mod.f90:961-1020): the step onto the centre point
// (at_target) or, on the last plane, onto the downwind point of the stencil
__device__ __forceinline__ void quad_step(float xxup, float xxc, float xxdn, float ssup, float ssc, float ssdn,
                                          float Inuup, bool at_target, const LcqArgs &a, float &Inu,
                                          float &plam) {
  const float dtmin = 1e-20f;
  const float t1 = (xxc - xxup) * a.rdsup, t3 = (xxdn - xxc) * a.rdsup;
  const float t2 = mc_slope(t1, t3);
  float dtup = a.dsup1 * (xxup + xxc) + a.dsup2 * (t1 - t2);
  float dtdn = a.dsup1 * (xxdn + xxc) + a.dsup2 * (t2 - t3);
  dtup = fmaxf(copysignf(fmaxf(fabsf(dtup), dtmin), dtup), a.mdtau);
  dtdn = fmaxf(copysignf(fmaxf(fabsf(dtdn), dtmin), dtdn), a.mdtau);
  const float dt = at_target ? dtup : dtdn;
  const float rdt = 1.0f / dt;
  float ex, U0, U1, q;
  dev::parab_moments(dt, ex, U0, U1, q);
  float s1, s3;
  if (at_target) {
    s1 = (ssc - ssup) * rdt;
    s3 = (ssdn - ssc) / dtdn;
  } else {
    s1 = (ssc - ssup) / dtup;
    s3 = (ssdn - ssc) * rdt;
  }
  const float dsdtau = mc_slope(s1, s3);
  plam = 0.0f;
  if (at_target) {
    const float B = (2.0f * U1 - q) * rdt, A = U0 - B, C = q - U1;
    Inu = Inuup * ex + A * ssup + B * ssc + C * dsdtau;
    if (dt > 0 && xxc > 0) plam = B / xxc;
  } else {
    const float A = q * rdt, B = U0 - A, C = U1 - q;
    Inu = Inuup * ex + A * ssdn + B * ssc + C * dsdtau;
    if (dt > 0 && xxdn > 0) plam = A / xxdn;
  }
  Inu = fmaxf(-a.imax, fminf(a.imax, Inu));
}

__global__ void __launch_bounds__(256) lcq_sweep(const LcqArgs a) {
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
  auto plane = [&](int step) { return o + (int64_t)(a.dir > 0 ? step : a.nz - 1 - step) * a.cs_z; };
  auto emit = [&](int64_t oz, float Inu, float plam) {
    if (a.J) a.J[oz] = a.J[oz] + a.wj * Inu;
    if (a.PsiJ) a.PsiJ[oz] = a.PsiJ[oz] + a.wj * plam;
    if (a.Iout) a.Iout[oz] = Inu;
    if (a.Psiout) a.Psiout[oz] = plam;
  };
  // the window (up, centre, down) starts at (ghost, first plane, second plane)
  float cu = a.gchi[go], su = a.gS[go], Iup = a.gI[go];
  int64_t oc = plane(0), od = plane(1);
  float cc = a.chi[oc], sc = a.S[oc], cd = a.chi[od], sd = a.S[od];
  for (int step = 0; step < a.nz - 1; ++step) {
    float Inu, plam;
    quad_step(cu, cc, cd, su, sc, sd, Iup, true, a, Inu, plam);
    emit(oc, Inu, plam);
    Iup = Inu;
    if (step < a.nz - 2) {
      cu = cc; su = sc;
      cc = cd; sc = sd; oc = od;
      od = plane(step + 2);
      cd = a.chi[od];
      sd = a.S[od];
    }
  }
  // the last plane: the same three points, stepping onto the downwind one
  float Inu, plam;
  quad_step(cu, cc, cd, su, sc, sd, Iup, false, a, Inu, plam);
  emit(od, Inu, plam);
  if (a.exitI) a.exitI[go] = Inu;
}

// long_char_linear (SRC/rt_solvers_mod.f90:666-757): the trapezoid step with the linear
// weights along each column, one lane per (frequency, column) carrying the upwind
// chi, S and I in registers; with split rates the remainder and the escape factor
struct LclArgs {
  const float *chi, *S, *gchi, *gS, *gI;
  float *J, *PsiJ, *Iout, *Psiout, *Irem, *esc, *exitI;
  int64_t cs_f, cs_x, cs_y, cs_z, gs_f, gs_x, gs_y;
  int nf, nx, ny, nz, dir, f_fast, split;
  float dsup1, mdtau, imax, wj;
};

__global__ void __launch_bounds__(256) lcl_sweep(const LclArgs a) {
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
  auto plane = [&](int step) { return o + (int64_t)(a.dir > 0 ? step : a.nz - 1 - step) * a.cs_z; };
  float xxup = a.gchi[go], ssup = a.gS[go], Iup = a.gI[go];
  // chi and S of the planes ahead, loaded kLcPF steps before their use in a register
  // ring (planes past the last repeat it and are never used)
  float qc[kLcPF], qs[kLcPF];
#pragma unroll
  for (int k = 0; k < kLcPF; ++k) {
    const int64_t op = plane(min(k, a.nz - 1));
    qc[k] = a.chi[op];
    qs[k] = a.S[op];
  }
  for (int s = 0; s < a.nz; ++s) {
    const int64_t oz = plane(s);
    const float xx = qc[0], ss = qs[0];
#pragma unroll
    for (int k = 0; k + 1 < kLcPF; ++k) {
      qc[k] = qc[k + 1];
      qs[k] = qs[k + 1];
    }
    {
      const int64_t op = plane(min(s + kLcPF, a.nz - 1));
      qc[kLcPF - 1] = a.chi[op];
      qs[kLcPF - 1] = a.S[op];
    }
    const float dt = fmaxf(a.dsup1 * (xxup + xx), a.mdtau);
    float E, W, Ps, Esc;
    dev::lin_weights(dt, E, W, Ps, Esc);
    float Inu = Iup * E + W * ssup + Ps * ss;
    const bool off = dt <= 0.0f || xx <= 0.0f || fabsf(Inu) > a.imax;
    const float rem = Iup * E + W * ssup;
    Inu = fmaxf(-a.imax, fminf(a.imax, Inu));
    if (a.split) {
      Ps = off ? 0.0f : Ps;
      a.Irem[oz] = off ? Inu : rem;
      a.esc[oz] = off ? 1.0f : Esc;
    }
    const float plam = (dt > 0.0f && xx > 0.0f) ? Ps / xx : 0.0f;
    if (a.J) a.J[oz] = a.J[oz] + a.wj * Inu;
    if (a.PsiJ) a.PsiJ[oz] = a.PsiJ[oz] + a.wj * plam;
    if (a.Iout) a.Iout[oz] = Inu;
    if (a.Psiout) a.Psiout[oz] = plam;
    xxup = xx;
    ssup = ss;
    Iup = Inu;
  }
  if (a.exitI) a.exitI[go] = Iup;
}

}  // namespace

int sweep_lcl(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  if (d->prec != M3DRT_F32)
    return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep: the long-characteristics linear solver is FP32 (as the reference)");
  if (b->eta_row) return set_error(M3DRT_EINVAL, "m3drt_sc_sweep: the long-characteristics linear solver takes no eta flags");
  LclArgs a;
  a.chi = static_cast<const float *>(b->chi);
  a.S = static_cast<const float *>(b->S);
  a.gchi = static_cast<const float *>(b->ghost_chi);
  a.gS = static_cast<const float *>(b->ghost_S);
  a.gI = static_cast<const float *>(b->ghost_I);
  const bool rates = (d->flags & M3DRT_GET_RATES) != 0;
  a.split = (d->flags & M3DRT_SPLIT_RATES) ? 1 : 0;
  a.J = static_cast<float *>(b->J);
  a.PsiJ = rates ? static_cast<float *>(b->PsiJ) : nullptr;
  a.Iout = static_cast<float *>(b->I_out);
  a.Psiout = rates ? static_cast<float *>(b->Psi_out) : nullptr;
  a.Irem = static_cast<float *>(b->Irem_out);
  a.esc = static_cast<float *>(b->escape_out);
  a.exitI = static_cast<float *>(b->exit_I);
  a.cs_f = d->cs_f; a.cs_x = d->cs_x; a.cs_y = d->cs_y; a.cs_z = d->cs_z;
  a.gs_f = d->gs_f; a.gs_x = d->gs_x; a.gs_y = d->gs_y;
  a.nf = d->nf; a.nx = d->nx; a.ny = d->ny; a.nz = d->nz; a.dir = d->dir > 0 ? 1 : -1;
  a.f_fast = d->cs_f <= d->cs_x ? 1 : 0;
  // SRC/rt_solvers_mod.f90:681-682: half the path per plane, 0.5*ds in real(4) over |mu|
  const float ds = (float)d->ds[2];
  a.dsup1 = (float)((double)(0.5f * ds) / fabs(d->vec[2]));
  a.mdtau = -(float)d->maser_dtau_max;
  a.imax = (float)d->maser_Inu_max;
  a.wj = (float)d->wj;
  const int64_t n = (int64_t)d->nx * d->ny * d->nf;
  hipLaunchKernelGGL(lcl_sweep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

int sweep_lcq(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  if (d->prec != M3DRT_F32)
    return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep: the quadratic solver is FP32 (as the reference)");
  if (d->nz < 2) return set_error(M3DRT_EINVAL, "m3drt_sc_sweep: the quadratic solver needs nz >= 2");
  if ((d->flags & M3DRT_SPLIT_RATES) || b->eta_row)
    return set_error(M3DRT_EINVAL, "m3drt_sc_sweep: the quadratic solver takes neither split rates nor eta flags");
  LcqArgs a;
  a.chi = static_cast<const float *>(b->chi);
  a.S = static_cast<const float *>(b->S);
  a.gchi = static_cast<const float *>(b->ghost_chi);
  a.gS = static_cast<const float *>(b->ghost_S);
  a.gI = static_cast<const float *>(b->ghost_I);
  a.J = static_cast<float *>(b->J);
  a.PsiJ = (d->flags & M3DRT_GET_RATES) ? static_cast<float *>(b->PsiJ) : nullptr;
  a.Iout = static_cast<float *>(b->I_out);
  a.Psiout = (d->flags & M3DRT_GET_RATES) ? static_cast<float *>(b->Psi_out) : nullptr;
  a.exitI = static_cast<float *>(b->exit_I);
  a.cs_f = d->cs_f; a.cs_x = d->cs_x; a.cs_y = d->cs_y; a.cs_z = d->cs_z;
  a.gs_f = d->gs_f; a.gs_x = d->gs_x; a.gs_y = d->gs_y;
  a.nf = d->nf; a.nx = d->nx; a.ny = d->ny; a.nz = d->nz; a.dir = d->dir > 0 ? 1 : -1;
  a.f_fast = d->cs_f <= d->cs_x ? 1 : 0;
  // SRC/rt_solvers_mod.f90:905-912: the path per plane in real(4)
  const float ds = (float)d->ds[2];
  const float dsup = (float)((double)ds / fabs(d->vec[2]));
  a.rdsup = 1.0f / dsup;
  a.dsup1 = 0.5f * dsup;
  a.dsup2 = 1.0f / 12.0f * dsup * dsup;
  a.mdtau = -(float)d->maser_dtau_max;
  a.imax = (float)d->maser_Inu_max;
  a.wj = (float)d->wj;
  const int64_t n = (int64_t)d->nx * d->ny * d->nf;
  hipLaunchKernelGGL(lcq_sweep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
