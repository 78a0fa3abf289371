// The sampled background-opacity table of HOP (HOP/optab_mod.f90): a 3D table over
// (wavelength, PC axis 1, PC axis 2) of the continuum extinction and its
// scattering part (/chi500), with monotone PCHIP node slopes along every axis
// (sample_build_derivatives, :127-166), interpolated per (frequency, cell) by a
// separable tensor-product cubic Hermite (sample_interpolate, :181-287): the step
// before the formal solution that fills chi_cont and sigma
// (lte_background_opacities, SRC/rt_patch_mod.f90:2115-2117).
//
// A plan holds the table on the device with its slopes, interleaved per node as
// (value, slope along wavelength, along PC axis 1, along PC axis 2), so that one
// corner of a component is one 16-byte read.  The slopes are built on the device,
// one lane per table line (FP64 along wavelength as pchip_slopes8; FP32 quotients
// with an FP64 harmonic mean along the PC axes as pchip_slopes4).
//
// Interpolation: one lane per cell, a workgroup row per block of frequencies; the
// per-cell PCA coordinates and Hermite coefficients (FP64 logarithms and rotation,
// FP32 coefficients) once per lane and block, the per-frequency wavelength bracket
// and coefficients from a small table read at wave-uniform addresses.  Every FP32
// product and sum rounded apart in the reference's order (fp contract off), so the
// results equal the reference arithmetic's.
#include "m3drt_internal.hpp"
#include <cmath>

struct m3drt_optab_plan {
  m3drt_ctx *ctx = nullptr;
  int nnu = 0, nt = 0, ne = 0;
  double mu[2], rot[4], c0[2], dc[2];
  double *lam = nullptr;      // (nnu)
  float *slopes = nullptr;    // 6 x (nnu, nt, ne): dl, dt, de, sl, st, se
  float4 *node = nullptr;     // 2 x (nne, nt, nnu): (value, dl, dt, de) per node, val then scat
};

namespace m3drt {
namespace {

__global__ void __launch_bounds__(256) slopes_lam_kernel(int nnu, int nlines, const double *__restrict__ lam,
                                                         const float *__restrict__ f, float *__restrict__ out) {
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= nlines) return;
  const float *y = f + (size_t)line * nnu;
  float *d = out + (size_t)line * nnu;
  // pchip_slopes8 (HOP/interpol_mod.f90:243-283) with imin = 1, imax = n
  double dprev = 0.0;  // slope of node i-1 (node 0's is written last)
  double first = 0.0;
  for (int i = 0; i < nnu; ++i) d[i] = 0.0f;
  if (nnu == 2) {
    const double s = ((double)y[1] - (double)y[0]) / (lam[1] - lam[0]);
    d[0] = (float)s;
    d[1] = (float)s;
    return;
  }
  double d1s = 0.0;
  for (int i = 1; i < nnu - 1; ++i) {
    const double dx0 = lam[i] - lam[i - 1], dx1 = lam[i + 1] - lam[i];
    const double dy0 = (double)y[i] - (double)y[i - 1], dy1 = (double)y[i + 1] - (double)y[i];
    double s = 0.0;
    if (dy0 != 0 && dy1 != 0) {
      const double q0 = dy0 / dx0, q1 = dy1 / dx1;
      if ((q0 < 0) == (q1 < 0)) {
        const double alpha = (1.0 + dx0 / (dx1 + dx0)) / 3.0;
        s = 1.0 / ((1.0 - alpha) / q0 + alpha / q1);
      }
    }
    d[i] = (float)s;
    if (i == 1) d1s = s;
    dprev = s;
  }
  (void)first;
  const double dxa = lam[1] - lam[0], dya = (double)y[1] - (double)y[0];
  d[0] = (float)(2.0 * dya / dxa - d1s);
  const double dxb = lam[nnu - 1] - lam[nnu - 2], dyb = (double)y[nnu - 1] - (double)y[nnu - 2];
  d[nnu - 1] = (float)(2.0 * dyb / dxb - dprev);
}

// pchip_slopes4 along a PC axis: lines of n nodes at stride s; line l starts at
// base(l) = (l % nfast) + (l / nfast) * outer
__global__ void __launch_bounds__(256) slopes_axis_kernel(int n, int nlines, int nfast, int64_t outer, int64_t s,
                                                          double c0, double dc, const float *__restrict__ f,
                                                          float *__restrict__ out) {
#pragma clang fp contract(off)
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= nlines) return;
  const int64_t b = (int64_t)(line % nfast) + (int64_t)(line / nfast) * outer;
  const float *y = f + b;
  float *d = out + b;
  auto ax = [&](int i) { return (float)(c0 + i * dc); };
  for (int i = 0; i < n; ++i) d[i * s] = 0.0f;
  if (n == 2) {
    const float q = (y[s] - y[0]) / (ax(1) - ax(0));
    d[0] = q;
    d[s] = q;
    return;
  }
  float d1s = 0.0f, dprev = 0.0f;
  for (int i = 1; i < n - 1; ++i) {
    const float dx0 = ax(i) - ax(i - 1), dx1 = ax(i + 1) - ax(i);
    const float dy0 = y[i * s] - y[(i - 1) * s], dy1 = y[(i + 1) * s] - y[i * s];
    float r = 0.0f;
    if (dy0 != 0 && dy1 != 0) {
      const double q0 = (double)(dy0 / dx0), q1 = (double)(dy1 / dx1);
      if ((q0 < 0) == (q1 < 0)) {
        const double alpha = (double)((1.0f + dx0 / (dx1 + dx0)) / 3.0f);
        r = (float)(1.0 / ((1.0 - alpha) / q0 + alpha / q1));
      }
    }
    d[i * s] = r;
    if (i == 1) d1s = r;
    dprev = r;
  }
  d[0] = 2.0f * (y[s] - y[0]) / (ax(1) - ax(0)) - d1s;
  d[(n - 1) * s] = 2.0f * (y[(n - 1) * s] - y[(n - 2) * s]) / (ax(n - 1) - ax(n - 2)) - dprev;
}

__global__ void __launch_bounds__(256) interleave_kernel(int64_t n3, const float *__restrict__ val,
                                                         const float *__restrict__ scat,
                                                         const float *__restrict__ sl, float4 *__restrict__ node) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
    node[i] = make_float4(val[i], sl[i], sl[n3 + i], sl[2 * n3 + i]);
    node[n3 + i] = make_float4(scat[i], sl[3 * n3 + i], sl[4 * n3 + i], sl[5 * n3 + i]);
  }
}

struct LamCoef {
  int k;  // 0-based lower node of the bracket
  float pl, wc0, wc1, wcm0, wcm1;
};

__device__ __forceinline__ void hermite(float t, float h, float &p, float &c0, float &c1, float &cm0, float &cm1) {
#pragma clang fp contract(off)
  const float t2 = t * t, t3 = t2 * t;
  p = 3.0f * t2 - 2.0f * t3;
  c0 = 1.0f - p;
  c1 = p;
  cm0 = h * (t3 - 2.0f * t2 + t);
  cm1 = h * (t3 - t2);
}

__global__ void __launch_bounds__(256) lam_coef_kernel(int nf, int nnu, const double *__restrict__ lam,
                                                       const double *__restrict__ lq, LamCoef *__restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double x = lq[f];
  // searchsorted8 (HOP/op_math_mod.f90:122-199), 1-based, then clamped to 1..nnu-1
  int k;
  if (x < lam[0]) {
    k = 0;
  } else if (x > lam[nnu - 1]) {
    k = nnu;
  } else {
    int l = 1, r = nnu;
    while (r - l > 1) {
      const int m = l + (r - l) / 2;
      if (lam[m - 1] >= x) r = m;
      else l = m;
    }
    k = l;
  }
  k = k < 1 ? 1 : (k > nnu - 1 ? nnu - 1 : k);
  LamCoef c;
  c.k = k - 1;
  const double span = lam[k] - lam[k - 1];
  hermite((float)((x - lam[k - 1]) / span), (float)span, c.pl, c.wc0, c.wc1, c.wcm0, c.wcm1);
  out[f] = c;
}

struct InterpArgs {
  int nnu, nt, ne, nf, fchunk;
  int64_t ncell, sf, sc;
  double mu0, mu1, r11, r21, r12, r22, c00, c01, rd1, rd2;
  float dc1, dc2;
  const float4 *node;
  const LamCoef *lc;
  const double *temp, *nein;
  const float *chi500;
  float *back, *scat;
};

__global__ void __launch_bounds__(256) optab_interp_kernel(const InterpArgs a) {
#pragma clang fp contract(off)
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  const double d1 = log(a.temp[c]) - a.mu0, d2 = log(a.nein[c]) - a.mu1;
  const double p1 = a.r11 * d1 + a.r21 * d2, p2 = a.r12 * d1 + a.r22 * d2;
  const double ft = (p1 - a.c00) * a.rd1 + 1.0, fe = (p2 - a.c01) * a.rd2 + 1.0;
  int it = (int)floor(ft), je = (int)floor(fe);
  it = it < 1 ? 1 : (it > a.nt - 1 ? a.nt - 1 : it);
  je = je < 1 ? 1 : (je > a.ne - 1 ? a.ne - 1 : je);
  float pt, tc0, tc1, tcm0, tcm1, pe, nc0, nc1, ncm0, ncm1;
  hermite((float)(ft - it), a.dc1, pt, tc0, tc1, tcm0, tcm1);
  hermite((float)(fe - je), a.dc2, pe, nc0, nc1, ncm0, ncm1);
  (void)pe;
  const float c500 = a.chi500[c];
  const int64_t n3 = (int64_t)a.nnu * a.nt * a.ne;
  // the four (PC1, PC2) corners' node rows: (it, je), (it+1, je), (it, je+1), (it+1, je+1)
  const int64_t r00 = ((int64_t)(je - 1) * a.nt + (it - 1)) * a.nnu;
  const int64_t r10 = r00 + a.nnu, r01 = r00 + (int64_t)a.nt * a.nnu, r11 = r01 + a.nnu;
  const dev::cptr<LamCoef> lc = dev::ro(a.lc);
  const int f0 = blockIdx.y * a.fchunk, f1 = min(a.nf, f0 + a.fchunk);
  for (int f = f0; f < f1; ++f) {
    const LamCoef L = lc[f];
    float kap[2];
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      const float4 *nd = a.node + comp * n3 + L.k;
      const float4 A00 = nd[r00], B00 = nd[r00 + 1], A10 = nd[r10], B10 = nd[r10 + 1];
      const float4 A01 = nd[r01], B01 = nd[r01 + 1], A11 = nd[r11], B11 = nd[r11 + 1];
      auto vcol = [&](const float4 &A, const float4 &B) {
        return L.wc0 * A.x + L.wc1 * B.x + L.wcm0 * A.y + L.wcm1 * B.y;
      };
      const float q = 1.0f - L.pl, p = L.pl;
      const float v00 = vcol(A00, B00), v10 = vcol(A10, B10), v01 = vcol(A01, B01), v11 = vcol(A11, B11);
      const float gt00 = q * A00.z + p * B00.z, gt10 = q * A10.z + p * B10.z;
      const float gt01 = q * A01.z + p * B01.z, gt11 = q * A11.z + p * B11.z;
      const float ge00 = q * A00.w + p * B00.w, ge10 = q * A10.w + p * B10.w;
      const float ge01 = q * A01.w + p * B01.w, ge11 = q * A11.w + p * B11.w;
      const float fa = tc0 * v00 + tc1 * v10 + tcm0 * gt00 + tcm1 * gt10;
      const float fb = tc0 * v01 + tc1 * v11 + tcm0 * gt01 + tcm1 * gt11;
      const float g0 = (1.0f - pt) * ge00 + pt * ge10;
      const float g1 = (1.0f - pt) * ge01 + pt * ge11;
      kap[comp] = nc0 * fa + nc1 * fb + ncm0 * g0 + ncm1 * g1;
    }
    const int64_t o = f * a.sf + c * a.sc;
    a.back[o] = (kap[0] + kap[1]) * c500;
    a.scat[o] = kap[1] * c500;
  }
}

int grid1(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 1 << 20); }

}  // namespace
}  // namespace m3drt

using m3drt::set_error;

extern "C" int m3drt_optab_plan_create(m3drt_ctx *ctx, const m3drt_optab_table *t, m3drt_optab_plan **plan) {
  if (!ctx || !t || !plan) return set_error(M3DRT_EINVAL, "m3drt_optab_plan_create: NULL argument");
  *plan = nullptr;
  if (t->nnu < 2 || t->ntemp < 2 || t->nne < 2)
    return set_error(M3DRT_EINVAL, "m3drt_optab_plan_create: a sampled table needs at least two nodes per axis");
  if (!t->lam || !t->val || !t->scat) return set_error(M3DRT_EINVAL, "m3drt_optab_plan_create: missing table array");
  if (!(t->dc[0] > 0 && t->dc[1] > 0)) return set_error(M3DRT_EINVAL, "m3drt_optab_plan_create: dc must be > 0");
  const int64_t n3 = (int64_t)t->nnu * t->ntemp * t->nne;
  if (n3 > (int64_t)1 << 31) return set_error(M3DRT_ELIMIT, "m3drt_optab_plan_create: table beyond 2^31 nodes");
  M3DRT_HIP(hipSetDevice(ctx->device));
  m3drt_optab_plan *p = new m3drt_optab_plan();
  p->ctx = ctx;
  p->nnu = t->nnu;
  p->nt = t->ntemp;
  p->ne = t->nne;
  for (int i = 0; i < 2; ++i) {
    p->mu[i] = t->mu[i];
    p->c0[i] = t->c0[i];
    p->dc[i] = t->dc[i];
  }
  for (int i = 0; i < 4; ++i) p->rot[i] = t->rot[i];
  auto fail = [&](hipError_t e, const char *w) {
    (void)hipFree(p->lam);
    (void)hipFree(p->slopes);
    (void)hipFree(p->node);
    delete p;
    return m3drt::hip_check(e, w);
  };
  hipError_t e;
  if ((e = hipMalloc(&p->lam, sizeof(double) * t->nnu)) != hipSuccess) return fail(e, "hipMalloc(lam)");
  if ((e = hipMalloc(&p->slopes, sizeof(float) * 6 * n3)) != hipSuccess) return fail(e, "hipMalloc(slopes)");
  if ((e = hipMalloc(&p->node, sizeof(float4) * 2 * n3)) != hipSuccess) return fail(e, "hipMalloc(node)");
  hipStream_t s = ctx->stream;
  if ((e = hipMemcpyAsync(p->lam, t->lam, sizeof(double) * t->nnu, hipMemcpyDeviceToDevice, s)) != hipSuccess)
    return fail(e, "hipMemcpyAsync(lam)");
  const int nnu = t->nnu, nt = t->ntemp, ne = t->nne;
  for (int comp = 0; comp < 2; ++comp) {
    const float *f = comp == 0 ? t->val : t->scat;
    float *sl = p->slopes + (int64_t)comp * 3 * n3;
    hipLaunchKernelGGL(m3drt::slopes_lam_kernel, dim3(m3drt::grid1(nt * ne)), dim3(256), 0, s, nnu, nt * ne, p->lam, f,
                       sl);
    // along PC axis 1: lines over (k, je), nodes nnu apart; along PC axis 2: lines over (k, it)
    hipLaunchKernelGGL(m3drt::slopes_axis_kernel, dim3(m3drt::grid1((int64_t)nnu * ne)), dim3(256), 0, s, nt, nnu * ne,
                       nnu, (int64_t)nnu * nt, (int64_t)nnu, t->c0[0], t->dc[0], f, sl + n3);
    hipLaunchKernelGGL(m3drt::slopes_axis_kernel, dim3(m3drt::grid1((int64_t)nnu * nt)), dim3(256), 0, s, ne, nnu * nt,
                       nnu * nt, (int64_t)0, (int64_t)nnu * nt, t->c0[1], t->dc[1], f, sl + 2 * n3);
  }
  hipLaunchKernelGGL(m3drt::interleave_kernel, dim3(m3drt::grid1(n3)), dim3(256), 0, s, n3, t->val, t->scat, p->slopes,
                     p->node);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e, "optab slopes");
  *plan = p;
  return M3DRT_OK;
}

extern "C" int m3drt_optab_slopes(m3drt_optab_plan *p, float *dl, float *dt, float *de, float *sl, float *st,
                                  float *se) {
  if (!p) return set_error(M3DRT_EINVAL, "m3drt_optab_slopes: NULL plan");
  const int64_t n3 = (int64_t)p->nnu * p->nt * p->ne;
  float *out[6] = {dl, dt, de, sl, st, se};
  M3DRT_HIP(hipSetDevice(p->ctx->device));
  for (int i = 0; i < 6; ++i)
    if (out[i])
      M3DRT_HIP(hipMemcpyAsync(out[i], p->slopes + i * n3, sizeof(float) * n3, hipMemcpyDeviceToDevice, p->ctx->stream));
  return M3DRT_OK;
}

extern "C" int m3drt_optab_interpolate(m3drt_optab_plan *p, const m3drt_optab_query *q) {
  if (!p || !q) return set_error(M3DRT_EINVAL, "m3drt_optab_interpolate: NULL argument");
  if (q->ncell < 0 || q->nf < 0) return set_error(M3DRT_EINVAL, "m3drt_optab_interpolate: negative size");
  if (q->ncell == 0 || q->nf == 0) return M3DRT_OK;
  if (!q->temp || !q->ne || !q->chi500 || !q->lam_q || !q->back || !q->scat)
    return set_error(M3DRT_EINVAL, "m3drt_optab_interpolate: missing array");
  m3drt_ctx *ctx = p->ctx;
  M3DRT_HIP(hipSetDevice(ctx->device));
  int rc = m3drt::ensure_scratch(ctx, sizeof(m3drt::LamCoef) * (size_t)q->nf);
  if (rc) return rc;
  m3drt::LamCoef *lc = static_cast<m3drt::LamCoef *>(ctx->scratch);
  hipLaunchKernelGGL(m3drt::lam_coef_kernel, dim3((q->nf + 255) / 256), dim3(256), 0, ctx->stream, q->nf, p->nnu,
                     p->lam, q->lam_q, lc);
  m3drt::InterpArgs a;
  a.nnu = p->nnu;
  a.nt = p->nt;
  a.ne = p->ne;
  a.nf = q->nf;
  a.fchunk = 32;
  a.ncell = q->ncell;
  a.sf = q->sf ? q->sf : q->ncell;
  a.sc = q->sc ? q->sc : 1;
  a.mu0 = p->mu[0];
  a.mu1 = p->mu[1];
  a.r11 = p->rot[0];
  a.r21 = p->rot[1];
  a.r12 = p->rot[2];
  a.r22 = p->rot[3];
  a.c00 = p->c0[0];
  a.c01 = p->c0[1];
  a.rd1 = 1.0 / p->dc[0];
  a.rd2 = 1.0 / p->dc[1];
  a.dc1 = (float)p->dc[0];
  a.dc2 = (float)p->dc[1];
  a.node = p->node;
  a.lc = lc;
  a.temp = q->temp;
  a.nein = q->ne;
  a.chi500 = q->chi500;
  a.back = q->back;
  a.scat = q->scat;
  const int nfb = (q->nf + a.fchunk - 1) / a.fchunk;
  if ((q->ncell + 255) / 256 > 0x7fffffff || nfb > 65535)
    return set_error(M3DRT_ELIMIT, "m3drt_optab_interpolate: grid beyond the launch limits");
  hipLaunchKernelGGL(m3drt::optab_interp_kernel, dim3((unsigned)((q->ncell + 255) / 256), nfb), dim3(256), 0,
                     ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

extern "C" int m3drt_optab_plan_destroy(m3drt_optab_plan *p) {
  if (!p) return M3DRT_OK;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  (void)hipFree(p->lam);
  (void)hipFree(p->slopes);
  (void)hipFree(p->node);
  delete p;
  return M3DRT_OK;
}
