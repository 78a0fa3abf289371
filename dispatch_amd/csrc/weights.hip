// The step weights the sweep kernels use, evaluated elementwise over an array of
// optical depths: linear_weights (SRC/rt_solvers_mod.f90:785-797) and
// parab_moments (SRC/rt_solvers_mod.f90:862-877), both built on exp_neg
// (HOP/exp_neg.inc:20-37).  The kernels inline dev::lin_weights and
// dev::parab_moments; these entry points run the very same functions so that
// their accuracy can be measured against the closed forms.
#include "m3drt_internal.hpp"

namespace m3drt {
namespace {

__global__ void __launch_bounds__(256) linear_weights_kernel(int64_t n, const float *__restrict__ t,
                                                              float *__restrict__ E, float *__restrict__ W,
                                                              float *__restrict__ Psi, float *__restrict__ Esc) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float e, w, p, s;
    dev::lin_weights(t[i], e, w, p, s);
    E[i] = e;
    W[i] = w;
    Psi[i] = p;
    Esc[i] = s;
  }
}

__global__ void __launch_bounds__(256) parab_moments_kernel(int64_t n, const float *__restrict__ t,
                                                             float *__restrict__ E, float *__restrict__ U0,
                                                             float *__restrict__ U1, float *__restrict__ q) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float e, u0, u1, qq;
    dev::parab_moments(t[i], e, u0, u1, qq);
    E[i] = e;
    U0[i] = u0;
    U1[i] = u1;
    q[i] = qq;
  }
}

int grid_of(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 16384); }

}  // namespace

int step_weights(m3drt_ctx *ctx, int kind, int64_t n, const float *t, float *o1, float *o2, float *o3,
                 float *o4) {
  if (n < 0 || (n > 0 && !(t && o1 && o2 && o3 && o4))) return set_error(M3DRT_EINVAL, "weights: null array");
  if (n == 0) return M3DRT_OK;
  if (kind == 0)
    hipLaunchKernelGGL(linear_weights_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, t, o1, o2, o3, o4);
  else
    hipLaunchKernelGGL(parab_moments_kernel, dim3(grid_of(n)), dim3(256), 0, ctx->stream, n, t, o1, o2, o3, o4);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
