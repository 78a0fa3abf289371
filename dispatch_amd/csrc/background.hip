// The LTE background factors and the boundary planes of a directional sweep:
//   m3drt_lte_background          beta = exp(-h nu/kT) (clamped at the smallest normal
//                                 FP32 number) and Bnu = beta/(1 - beta) per
//                                 (frequency, cell): lte_background_opacities,
//                                 SRC/rt_patch_mod.f90:2004-2023
//   m3drt_set_upstream_boundary   the ghost plane a sweep starts from:
//                                 set_upstream_boundary, SRC/rt_patch_mod.f90:3270-3374
//   m3drt_set_downstream_boundary the exit plane handed to the next slab:
//                                 set_downstream_boundary, :3471-3511
// One lane per (frequency, cell) of a plane or spectrum, cells fastest; HBM-bound
// elementwise passes.  The exponential is the FP64 one rounded to FP32 (the
// reference takes the FP32 one: the two differ by at most an ulp where the FP32
// library rounds off the exact value).
#include "m3drt_internal.hpp"
#include <cfloat>

namespace m3drt {
namespace {

constexpr double kHK = 6.626176e-27 / 1.380662e-16;  // h/k, HOP/const_mod.f90:39,43,71

__device__ __forceinline__ float planck_beta(double nu, float h_kT) {
  const float hv = (float)(nu * (double)h_kT);
  const float b = (float)exp(-(double)hv);
  return b >= FLT_MIN ? b : FLT_MIN;
}

__global__ void __launch_bounds__(256) lte_background_kernel(int nf, int64_t ncell, const double *__restrict__ nu,
                                                             const double *__restrict__ temp, float *__restrict__ beta,
                                                             float *__restrict__ Bnu, int64_t sf, int64_t sc) {
#pragma clang fp contract(off)
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (c >= ncell || f >= nf) return;
  const float h_kT = (float)(kHK / (double)(float)temp[c]);
  const float b = planck_beta(nu[f], h_kT);
  const int64_t o = f * sf + c * sc;
  beta[o] = b;
  Bnu[o] = b / (1.0f - b);
}

struct GhostArgs {
  int kind, nf, nx, ny;
  int64_t ps_f, ps_x, ps_y;
  const float *chi1, *S1, *chi2, *S2, *g500, *c500, *T;
  const double *nu;
  const float *mchi, *mS, *mI;
  float *gchi, *gS, *gI;
  float x1, dx1;
};

__global__ void __launch_bounds__(256) ghost_kernel(const GhostArgs a) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;  // x + nx*y
  const int f = blockIdx.y;
  if (c >= a.nx * a.ny || f >= a.nf) return;
  const int x = c % a.nx, y = c / a.nx;
  const int64_t o = f * a.ps_f + x * a.ps_x + y * a.ps_y;
  float chi = 0.0f, S = 0.0f, I = 0.0f;
  if (a.kind == M3DRT_GHOST_LTE) {
    const float r500 = a.g500[c] / a.c500[c];
    const float h_kT = (float)(kHK / (double)a.T[c]);
    const float b = planck_beta(a.nu[f], h_kT);
    chi = a.chi1[o] * r500;
    S = b / (1.0f - b);
    I = S;
  } else if (a.kind == M3DRT_GHOST_PCHIP) {
    // pchip_mono at one plane below the first three: the first interval's cubic at
    // t = -1 with the slopes 2 dy1/dx1 (node 1) and 0 (node 2, never formed)
    auto extrap = [&](float y1, float y2) {
      const float m2 = 0.0f;
      const float m1 = 2.0f * (y2 - y1) / a.dx1 - m2;
      const float t1 = (0.0f - a.x1) / a.dx1;
      const float t2 = t1 * t1, t3 = t2 * t1;
      const float p = 3.0f * t2 - 2.0f * t3;
      const float v = (1.0f - p) * y1 + p * y2 + a.dx1 * ((t3 - 2.0f * t2 + t1) * m1 + (t3 - t2) * m2);
      return fmaxf(v, 0.0f);
    };
    chi = extrap(a.chi1[o], a.chi2[o]);
    S = extrap(a.S1[o], a.S2[o]);
    I = S;
  } else if (a.kind == M3DRT_GHOST_UPSTREAM) {
    chi = a.mchi[o];
    S = a.mS[o];
    I = a.mI[o];
  }
  a.gchi[o] = chi;
  a.gS[o] = S;
  a.gI[o] = I;
}

__global__ void __launch_bounds__(256) copy_plane_kernel(int nf, int nxy, int nx, int64_t ps_f, int64_t ps_x,
                                                         int64_t ps_y, const float *__restrict__ a,
                                                         const float *__restrict__ b, const float *__restrict__ c,
                                                         float *__restrict__ ma, float *__restrict__ mb,
                                                         float *__restrict__ mc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (i >= nxy || f >= nf) return;
  const int64_t o = f * ps_f + (i % nx) * ps_x + (i / nx) * ps_y;
  ma[o] = a[o];
  mb[o] = b[o];
  mc[o] = c[o];
}

}  // namespace
}  // namespace m3drt

using m3drt::set_error;

extern "C" int m3drt_lte_background(m3drt_ctx *ctx, const m3drt_background_lte_args *a) {
  if (!ctx || !a || !a->nu || !a->temp || !a->beta || !a->Bnu)
    return set_error(M3DRT_EINVAL, "m3drt_lte_background: NULL argument");
  if (a->nf < 0 || a->ncell < 0 || a->nf > 65535)
    return set_error(M3DRT_EINVAL, "m3drt_lte_background: 0 <= nf <= 65535 and ncell >= 0");
  if (a->nf == 0 || a->ncell == 0) return M3DRT_OK;
  const int64_t sf = a->sf ? a->sf : a->ncell, sc = a->sc ? a->sc : 1;
  M3DRT_HIP(hipSetDevice(ctx->device));
  const dim3 grid((unsigned)((a->ncell + 255) / 256), (unsigned)a->nf);
  hipLaunchKernelGGL(m3drt::lte_background_kernel, grid, dim3(256), 0, ctx->stream, a->nf, a->ncell, a->nu, a->temp,
                     a->beta, a->Bnu, sf, sc);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

extern "C" int m3drt_set_upstream_boundary(m3drt_ctx *ctx, const m3drt_boundary_args *b) {
  if (!ctx || !b) return set_error(M3DRT_EINVAL, "m3drt_set_upstream_boundary: NULL argument");
  if (b->nf < 1 || b->nx < 1 || b->ny < 1 || b->nf > 65535 || (int64_t)b->nx * b->ny > (1 << 30))
    return set_error(M3DRT_EINVAL, "m3drt_set_upstream_boundary: bad plane size");
  if (!b->ghost_chi || !b->ghost_S || !b->ghost_I)
    return set_error(M3DRT_EINVAL, "m3drt_set_upstream_boundary: missing ghost plane");
  switch (b->kind) {
    case M3DRT_GHOST_TOP:
      break;
    case M3DRT_GHOST_LTE:
      if (!b->chi1 || !b->g500 || !b->chi500_1 || !b->T_ghost || !b->nu)
        return set_error(M3DRT_EINVAL, "m3drt_set_upstream_boundary: the LTE ghost needs chi1, g500, chi500_1, "
                                       "T_ghost and nu");
      break;
    case M3DRT_GHOST_PCHIP:
      if (!b->chi1 || !b->S1 || !b->chi2 || !b->S2 || !(b->ds3 > 0))
        return set_error(M3DRT_EINVAL, "m3drt_set_upstream_boundary: the PCHIP ghost needs the first two planes' "
                                       "chi and S and ds3 > 0");
      break;
    case M3DRT_GHOST_UPSTREAM:
      if (!b->mem_chi || !b->mem_S || !b->mem_I)
        return set_error(M3DRT_EINVAL, "m3drt_set_upstream_boundary: the upstream ghost needs mem_chi, mem_S, mem_I");
      break;
    default:
      return set_error(M3DRT_EINVAL, "m3drt_set_upstream_boundary: unknown kind");
  }
  m3drt::GhostArgs a;
  a.kind = b->kind;
  a.nf = b->nf;
  a.nx = b->nx;
  a.ny = b->ny;
  a.ps_f = b->ps_f ? b->ps_f : (int64_t)b->nx * b->ny;
  a.ps_x = b->ps_x ? b->ps_x : 1;
  a.ps_y = b->ps_y ? b->ps_y : b->nx;
  a.chi1 = b->chi1;
  a.S1 = b->S1;
  a.chi2 = b->chi2;
  a.S2 = b->S2;
  a.g500 = b->g500;
  a.c500 = b->chi500_1;
  a.T = b->T_ghost;
  a.nu = b->nu;
  a.mchi = b->mem_chi;
  a.mS = b->mem_S;
  a.mI = b->mem_I;
  a.gchi = b->ghost_chi;
  a.gS = b->ghost_S;
  a.gI = b->ghost_I;
  a.x1 = (float)(1 * b->ds3);
  a.dx1 = (float)(2 * b->ds3) - a.x1;
  M3DRT_HIP(hipSetDevice(ctx->device));
  const int nxy = b->nx * b->ny;
  hipLaunchKernelGGL(m3drt::ghost_kernel, dim3((nxy + 255) / 256, b->nf), dim3(256), 0, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

extern "C" int m3drt_set_downstream_boundary(m3drt_ctx *ctx, const m3drt_boundary_args *b) {
  if (!ctx || !b) return set_error(M3DRT_EINVAL, "m3drt_set_downstream_boundary: NULL argument");
  if (b->nf < 1 || b->nx < 1 || b->ny < 1 || b->nf > 65535 || (int64_t)b->nx * b->ny > (1 << 30))
    return set_error(M3DRT_EINVAL, "m3drt_set_downstream_boundary: bad plane size");
  if (!b->chi1 || !b->S1 || !b->exit_I || !b->mem_chi || !b->mem_S || !b->mem_I)
    return set_error(M3DRT_EINVAL, "m3drt_set_downstream_boundary: needs the exit plane (chi1, S1, exit_I) and mem");
  const int64_t ps_f = b->ps_f ? b->ps_f : (int64_t)b->nx * b->ny;
  const int64_t ps_x = b->ps_x ? b->ps_x : 1, ps_y = b->ps_y ? b->ps_y : b->nx;
  M3DRT_HIP(hipSetDevice(ctx->device));
  const int nxy = b->nx * b->ny;
  hipLaunchKernelGGL(m3drt::copy_plane_kernel, dim3((nxy + 255) / 256, b->nf), dim3(256), 0, ctx->stream, b->nf, nxy,
                     b->nx, ps_f, ps_x, ps_y, b->chi1, b->S1, b->exit_I, b->mem_chi, b->mem_S, b->mem_I);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}
