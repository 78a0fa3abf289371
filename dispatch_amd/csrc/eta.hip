// update_eta_mask (SRC/rt_patch_mod.f90:3671-3802) on dense per-(f, cell) flags in
// the native layout (z, f, y, x): which pairs take short_char2's eta step.
//   A   fire = min(chi, |chi| - q (chi + (1-beta) chi S)) <= 0; new = fire & !src;
//       src |= fire (and the same on the upstream ghost plane with 1-beta = 1)
//   B1  c = new(z) | new(z - dir) [| ghost new on the slab's first plane]
//   B2  OR over x +- rx (periodic)           B3  mask |= OR over y +- ry (periodic)
// The box OR of a union is the union of the box ORs, so the sources' spread to
// their own plane and the next one downwind is one separable stencil.  Streaming,
// byte-wide passes; the scratch holds two byte cubes.
#include "m3drt_internal.hpp"
#include <algorithm>

#pragma clang fp contract(off)

namespace m3drt {

namespace {

struct EtaArgs {
  const float *chi, *S, *beta, *gchi, *gS;
  uint8_t *src, *mask, *gsrc, *nw, *gnw, *tmp;
  int nf, nx, ny, nz, dir, rx, ry, ghost;
};

__device__ __forceinline__ bool fires(float chi, float one_m_b, float S) {
  const float t = fabsf(chi) - 1e-2f * (chi + one_m_b * chi * S);
  return fminf(chi, t) <= 0.0f;
}

__global__ void __launch_bounds__(256) eta_fire(const EtaArgs a, int64_t n, int64_t npl) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n + npl; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < n) {
      const bool f = fires(a.chi[i], 1.0f - a.beta[i], a.S[i]);
      const uint8_t s = a.src[i];
      a.nw[i] = (uint8_t)(f && !s);
      if (f && !s) a.src[i] = 1;
    } else if (a.ghost) {
      const int64_t g = i - n;
      const bool f = fires(a.gchi[g], 1.0f, a.gS[g]);
      const uint8_t s = a.gsrc[g];
      a.gnw[g] = (uint8_t)(f && !s);
      if (f && !s) a.gsrc[g] = 1;
    }
  }
}

__global__ void __launch_bounds__(256) eta_union(const EtaArgs a, int64_t n, int64_t npl) {
  const int z1 = a.dir > 0 ? 0 : a.nz - 1;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i / npl);
    const int64_t r = i - (int64_t)z * npl;
    const int zu = z - a.dir;
    uint8_t c = a.nw[i];
    if (zu >= 0 && zu < a.nz) c |= a.nw[(int64_t)zu * npl + r];
    if (a.ghost && z == z1) c |= a.gnw[r];
    a.tmp[i] = c;
  }
}

// OR over a periodic window of +-w along x (step 1, period nx) or y (step nx, period ny)
__global__ void __launch_bounds__(256) eta_window(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                  int64_t n, int nx, int ny, int w, int along_y, int into) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % nx);
    const int y = (int)((i / nx) % ny);
    const int64_t row0 = i - (along_y ? (int64_t)y * nx : x);
    const int p = along_y ? ny : nx, at = along_y ? y : x;
    const int64_t step = along_y ? nx : 1;
    uint8_t m = 0;
    for (int d = -w; d <= w && !m; ++d) m |= in[row0 + (int64_t)dev::wrap(at + d, p) * step];
    if (into)
      out[i] |= m;
    else
      out[i] = m;
  }
}

int grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

}  // namespace

int update_eta_mask(m3drt_ctx *ctx, const m3drt_eta_args *r) {
  const int64_t npl = (int64_t)r->nf * r->nx * r->ny, n = npl * r->nz;
  int rc = ensure_scratch(ctx, (size_t)(2 * n + npl));
  if (rc) return rc;
  EtaArgs a;
  a.chi = r->chi; a.S = r->S; a.beta = r->beta; a.gchi = r->ghost_chi; a.gS = r->ghost_S;
  a.src = r->src; a.mask = r->mask; a.gsrc = r->ghost_src;
  a.nw = static_cast<uint8_t *>(ctx->scratch);
  a.tmp = a.nw + n;
  a.gnw = a.tmp + n;
  a.nf = r->nf; a.nx = r->nx; a.ny = r->ny; a.nz = r->nz; a.dir = r->dir > 0 ? 1 : -1;
  a.rx = r->rx; a.ry = r->ry; a.ghost = (r->flags & M3DRT_ETA_GHOST) ? 1 : 0;
  hipLaunchKernelGGL(eta_fire, dim3(grid(n + npl)), dim3(256), 0, ctx->stream, a, n, npl);
  M3DRT_HIP(hipGetLastError());
  hipLaunchKernelGGL(eta_union, dim3(grid(n)), dim3(256), 0, ctx->stream, a, n, npl);
  M3DRT_HIP(hipGetLastError());
  // x window: tmp -> nw (reused), y window: nw -> mask (OR-ed in)
  hipLaunchKernelGGL(eta_window, dim3(grid(n)), dim3(256), 0, ctx->stream, a.tmp, a.nw, n, a.nx, a.ny, a.rx, 0, 0);
  M3DRT_HIP(hipGetLastError());
  hipLaunchKernelGGL(eta_window, dim3(grid(n)), dim3(256), 0, ctx->stream, a.nw, a.mask, n, a.nx, a.ny, a.ry, 1, 1);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
