// Elementwise conversions between the working pair (xhat, ehat) of the opacity
// accumulation and the (chi, S) pair the formal solvers take, over the (f, cell)
// points of a window in any common layout:
//   convert_to_work_units   SRC/rt_patch_mod.f90:3518-3536
//   work_to_transfer_units  SRC/rt_patch_mod.f90:3571-3651 (maser floor on |chi|,
//                           line-scattering source, the split-tracked eta)
//   lte_background_emission SRC/rt_patch_mod.f90:2127-2328 (the continuum-scattering
//                           ALI step in FP64, J - B, the background emissivity)
// HBM-bound streaming kernels (a grid-stride loop; 5-7 floats per point).
#include "m3drt_internal.hpp"
#include <cfloat>
#include <algorithm>

#pragma clang fp contract(off)

namespace m3drt {

namespace {

struct UnitsArgs {
  int64_t n;
  int split, scat, suppress;
  float floor_frac;
  float *w1, *w2;
  const float *beta, *so1, *so2, *chi_scat, *dJnu;
  unsigned long long *n_floored;
};

__global__ void __launch_bounds__(256) work_to_transfer_kernel(const UnitsArgs a) {
  unsigned long long mine = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x) {
    const float bt = a.beta[i];
    const float xhat = a.w1[i];
    float eta, chi;
    if (a.split) {
      const float eta_rest = bt * a.so2[i];
      eta = bt * a.so1[i] + eta_rest;
      chi = xhat - eta;
      if (a.scat) eta = bt * a.so1[i] + (eta_rest + a.chi_scat[i] * a.dJnu[i]);
    } else {
      eta = a.w2[i] * bt;
      chi = xhat - eta;
      if (a.scat) eta = eta + a.chi_scat[i] * a.dJnu[i];
    }
    const float floor = fmaxf(a.floor_frac * fabsf(xhat), FLT_MIN);  // max(., tiny(1.0))
    if (a.suppress) {
      chi = fmaxf(chi, floor);
    } else if (fabsf(chi) < floor) {
      chi = copysignf(floor, chi);
      ++mine;
    }
    a.w1[i] = chi;
    a.w2[i] = eta / chi;
  }
  if (a.n_floored && mine) atomicAdd(a.n_floored, mine);
}

__global__ void __launch_bounds__(256) convert_to_work_kernel(int64_t n, const float *__restrict__ cont,
                                                              float *__restrict__ xhat, float *__restrict__ ehat,
                                                              const float *__restrict__ beta) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float e = ehat[i];
    xhat[i] = cont[i] + e;
    ehat[i] = e / beta[i];
  }
}

struct BgArgs {
  int64_t n;
  int scat_ali, j_to_b;
  float wscat;
  double scat_op_max;
  float *J, *Jali, *eta, *dJnu;
  const float *B, *chi, *sigma, *PsiJ, *Jold;
  unsigned long long *n_neg, *diag;  // diag: op_raw, op_max, amp_max as FP64 bit patterns (>= 0)
};

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// lte_background_emission, SRC/rt_patch_mod.f90:2127-2328
__global__ void __launch_bounds__(256) background_kernel(const BgArgs a) {
  unsigned long long neg = 0;
  double op_raw = 0.0, op_max = 0.0, amp_max = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x) {
    float j = a.J[i];
    if (a.scat_ali) {
      double op = (double)a.sigma[i] * (double)a.PsiJ[i];
      op_raw = fmax(op_raw, op);
      op = fmax(0.0, fmin(op, a.scat_op_max));
      op_max = fmax(op_max, op);
      const double Jfs = j;
      double Jn = (Jfs - op * (double)a.Jold[i]) / (1.0 - op);
      if (Jn < 0.0) {
        ++neg;
        Jn = Jfs;
      }
      if (Jfs > 0.0) amp_max = fmax(amp_max, Jn / Jfs);
      j = (float)Jn;
      if (a.Jali) a.Jali[i] = j;
    }
    const float b = a.B[i];
    if (a.j_to_b) j = b;
    j = j - b;
    a.J[i] = j;
    a.eta[i] = a.chi[i] * b + a.wscat * a.sigma[i] * j;
    if (a.dJnu) a.dJnu[i] = a.wscat * j;
  }
  if (a.scat_ali) {
    if (a.n_neg && neg) atomicAdd(a.n_neg, neg);
    op_raw = wave_max(op_raw);
    op_max = wave_max(op_max);
    amp_max = wave_max(amp_max);
    if (a.diag && (threadIdx.x & 63) == 0) {
      atomicMax(&a.diag[0], (unsigned long long)__double_as_longlong(op_raw));
      atomicMax(&a.diag[1], (unsigned long long)__double_as_longlong(op_max));
      atomicMax(&a.diag[2], (unsigned long long)__double_as_longlong(amp_max));
    }
  }
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

}  // namespace

int work_to_transfer(m3drt_ctx *ctx, const m3drt_units_args *r) {
  UnitsArgs a;
  a.n = r->n;
  a.split = (r->flags & M3DRT_UNITS_SPLIT) ? 1 : 0;
  a.scat = (r->flags & M3DRT_UNITS_SCAT) ? 1 : 0;
  a.suppress = (r->flags & M3DRT_UNITS_SUPPRESS_MASERS) ? 1 : 0;
  a.floor_frac = (float)r->maser_chi_floor;
  a.w1 = r->work1; a.w2 = r->work2;
  a.beta = r->beta; a.so1 = r->split_opac1; a.so2 = r->split_opac2; a.chi_scat = r->chi_scat; a.dJnu = r->dJnu;
  a.n_floored = reinterpret_cast<unsigned long long *>(r->n_floored);
  hipLaunchKernelGGL(work_to_transfer_kernel, dim3(grid_for(r->n)), dim3(256), 0, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

int background_emission(m3drt_ctx *ctx, const m3drt_background_args *r) {
  BgArgs a;
  a.n = r->n;
  a.scat_ali = (r->flags & M3DRT_BG_SCAT_ALI) ? 1 : 0;
  a.j_to_b = (r->flags & M3DRT_BG_J_TO_B) ? 1 : 0;
  a.wscat = r->wscat;
  a.scat_op_max = r->scat_op_max;
  a.J = r->J; a.Jali = r->Jali; a.eta = r->eta; a.dJnu = r->dJnu;
  a.B = r->B; a.chi = r->chi_cont; a.sigma = r->sigma; a.PsiJ = r->PsiJ; a.Jold = r->Jali_old;
  a.n_neg = reinterpret_cast<unsigned long long *>(r->n_neg);
  a.diag = reinterpret_cast<unsigned long long *>(r->diag);
  hipLaunchKernelGGL(background_kernel, dim3(grid_for(r->n)), dim3(256), 0, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

int convert_to_work_units(m3drt_ctx *ctx, int64_t n, const float *cont, float *xhat, float *ehat,
                          const float *beta) {
  hipLaunchKernelGGL(convert_to_work_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, cont, xhat, ehat,
                     beta);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
