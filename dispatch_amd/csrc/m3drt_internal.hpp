// Internal declarations of libm3drt: context, error state, device helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include "m3drt.h"

struct m3drt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void *scratch = nullptr;   // per-call scratch (dt500, rc per cell, ...); grown on demand
  size_t scratch_bytes = 0;
  bool scratch_captured = false;  // a stream capture recorded launches that use `scratch`
  bool scratch_fixed = false;     // `scratch` is a caller-fixed buffer for this call: never grown
  std::vector<void *> retired;    // outgrown scratch buffers captured graphs may still use
  int lds_max = 0;           // dynamic LDS bytes available per workgroup
  float *qlg_tab = nullptr;  // QLG-bot's two tables on the device (2 x 9020 floats)
};

namespace m3drt {

// error state of the calling thread (m3drt_last_error)
int set_error(int code, const std::string &msg);
int hip_check(hipError_t e, const char *what);
// Makes ctx->scratch hold at least `bytes`.  Launches recorded by a stream capture
// keep the scratch address, so a buffer that has been captured is never freed when
// it is outgrown (it is retired and freed by m3drt_destroy), and growth inside a
// capture is refused.  With a fixed buffer (m3drt_sc_sweep_noalloc) a larger need
// is refused with M3DRT_ELIMIT.
int ensure_scratch(m3drt_ctx *ctx, size_t bytes);
// the sweep entry point's checks and dispatch (m3drt_sc_sweep, m3drt_sc_sweep_noalloc)
int sc_sweep_checked(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b, const char *who);
// the fixed device scratch of m3drt_sc_sweep_noalloc: its address on `device`
// (resolved once per device by noalloc_setup, which m3drt_create calls)
int noalloc_setup(int device);
void *noalloc_scratch(int device, size_t *bytes);

#define M3DRT_HIP(call)                                         \
  do {                                                          \
    hipError_t e_ = (call);                                     \
    if (e_ != hipSuccess) return ::m3drt::hip_check(e_, #call); \
  } while (0)

int sweep_sc2(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b);
int sweep_lc2(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b);
int qlg_tables(m3drt_ctx *ctx);
// lref_from_cont: the tolerance reference logref is ln(tol/voigt_peak) + ln cont,
// formed at the nodes the lines need instead of read from o->logref (the line-list
// composition, whose cont is the working opacity the reference forms logref from).
// direct: o->chi and o->eta are that working opacity (the same array as o->cont), for
// a list of at most 4 lines alone in its composition and without a scattering split:
// each point's line sum, formed from zero, is added to it once, as the composition's
// zeroed increments would be.
int line_opacity(m3drt_ctx *ctx, const m3drt_linelist *ll, const m3drt_opacity_args *o, bool general,
                 bool lref_from_cont = false, bool direct = false);
int rate_solve(m3drt_ctx *ctx, const m3drt_rate_solve_args *r, m3drt_rate_solve_result *res);
int normalize_rates(m3drt_ctx *ctx, const m3drt_rate_solve_args *r, const float *wphi);
int bb_rates(m3drt_ctx *ctx, const m3drt_bb_rates_args *r);
int sweep_lcq(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b);
int sweep_lcl(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b);
int sweep_sc1(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b);
int sweep_sch(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b, const m3drt_hermite_fields *h);
int sweep_lch(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b, const m3drt_hermite_fields *h,
              const m3drt_lc_ghost *g, bool hermite2);
int sweep_adjoint(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_adjoint_bufs *b);
int b1d_synth_plain(m3drt_ctx *ctx, const m3drt_b1d_batch *b, const m3drt_b1d_tile *tile);
int b1d_plan_create(m3drt_ctx *ctx, const m3drt_b1d_batch *b, const m3drt_b1d_tile *tile, m3drt_b1d_plan **out);
int b1d_plan_run(m3drt_b1d_plan *p);
int b1d_plan_info(const m3drt_b1d_plan *p, m3drt_b1d_tile *tile, double *times, int32_t *ntimes);
int b1d_plan_destroy(m3drt_b1d_plan *p);
int update_eta_mask(m3drt_ctx *ctx, const m3drt_eta_args *r);
int work_to_transfer(m3drt_ctx *ctx, const m3drt_units_args *r);
int background_emission(m3drt_ctx *ctx, const m3drt_background_args *r);
int convert_to_work_units(m3drt_ctx *ctx, int64_t n, const float *cont, float *xhat, float *ehat,
                          const float *beta);
int bb_opacity(m3drt_ctx *ctx, const m3drt_bb_opacity_args *r);
int window_opacity(m3drt_ctx *ctx, const m3drt_bb_opacity_args *r, const m3drt_window_args *w);
int rotate_array(m3drt_ctx *ctx, const m3drt_rotate_args *r);
int bf_rates(m3drt_ctx *ctx, const m3drt_bf_rates_args *r);
int step_weights(m3drt_ctx *ctx, int kind, int64_t n, const float *t, float *o1, float *o2, float *o3, float *o4);

}  // namespace m3drt

// ------------------------------------------------------------------ device helpers
#ifdef __HIPCC__
namespace m3drt {
namespace dev {

// Read-only data the kernel never writes, read at wave-uniform addresses: through the
// constant address space the compiler issues scalar loads (one per wave, into SGPRs,
// through the scalar data cache) instead of 64-lane vector loads of one address.
// (The host pass of the compiler parses device code too; the qualifier is for the
// device pass.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
#else
template <typename T>
using cptr = const T *;
#endif
template <typename T>
__device__ __forceinline__ cptr<T> ro(const T *p) {
  return (cptr<T>)(p);
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }

// E = exp(-t) and u0 = 1 - exp(-t) in FP32, neither formed from the other (the
// quantities of exp_neg, HOP/exp_neg.inc:20-37), with t clamped at 87 as there.
// Cody-Waite reduction t = n ln2 + r, |r| <= ln2/2, ln2 split so that n*LN2_HI is
// exact for |n| < 2^15; m = 1 - e^-r = r*M(r) by the Taylor series of
// (1 - e^-r)/r to r^7 (truncation below 1e-9 relative on |r| <= ln2/2); s = 2^-n by
// ldexp.  Then E = s*(1 - m) and u0 = (1 - s) + s*m: within 1.2 and 1.6 ulp of the
// exact values over -2 <= t <= 87, where a single hardware exponential of
// -t*log2 e would carry the rounding of that product, |t| ulp and more.
__device__ __forceinline__ void exp_neg(float t, float &E, float &u0) {
  const float tc = fminf(t, 87.0f);
  const float n = __builtin_rintf(tc * 1.44269504088896341f);
  float r = __builtin_fmaf(-n, 0.693359375f, tc);
  r = __builtin_fmaf(-n, -2.12194440e-4f, r);
  const float M = 1.0f + r * (-0.5f + r * (1.66666672e-01f + r * (-4.16666679e-02f + r * (8.33333377e-03f +
                  r * (-1.38888892e-03f + r * (1.98412701e-04f + r * -2.48015876e-05f))))));
  const float m = r * M;
  const float s = __builtin_ldexpf(1.0f, -(int)n);
  E = s * (1.0f - m);
  u0 = __builtin_fmaf(s, m, 1.0f - s);
}

// Weights of the linear step over optical depth t (the quantities of
// linear_weights, SRC/rt_solvers_mod.f90:785-797): E = exp(-t), Psi = 1 - u0/t,
// W = u0 - Psi, Esc = 1 - Psi with u0 = 1 - E.
//
// FP32 form for gfx950: E and u0 from exp_neg above.  For |t| <= 1, where 1 - u0/t
// cancels, Psi = t*P(t) with the reference's degree-7 minimax P (lw_series,
// SRC/rt_solvers_mod.f90:803-809) and W = u0 - Psi; above, u0/t by the hardware
// reciprocal, Psi = 1 - u0/t, W = u0/t - E.  Both forms are evaluated and one is
// selected, so the wavefront never diverges.  Every output is within the worst
// error of the reference's own FP32 arithmetic over -2 <= t <= 90 (E to 87), plus
// at most one ulp.
__device__ __forceinline__ void lin_weights(float t, float &E, float &W, float &Psi, float &Esc) {
  const float ps = t * (5.000000000e-01f + t * (-1.666666716e-01f + t * (4.166666791e-02f +
                   t * (-8.333330974e-03f + t * (1.388740959e-03f + t * (-1.983609400e-04f +
                   t * (2.522676368e-05f + t * (-2.834696488e-06f))))))));
  float u0;
  exp_neg(t, E, u0);
  const float ut = u0 * rcp(t);
  const bool thin = fabsf(t) <= 1.0f;
  Psi = thin ? ps : 1.0f - ut;
  Esc = thin ? 1.0f - ps : ut;
  W = thin ? u0 - ps : ut - E;
}

// parab_moments (SRC/rt_solvers_mod.f90:862-877): E = e^-t, U0 = 1 - E, U1 = t - U0
// (t * Psi(t) below |t| = 1, Psi the degree-7 series of linear_weights), q = U2/t
// (degree-8 minimax below |t| = 2).  E and U0 from exp_neg, as the linear weights.
__device__ __forceinline__ void parab_moments(float t, float &E, float &U0, float &U1, float &q) {
#pragma clang fp contract(off)
  const float ps = t * (5.000000000e-01f + t * (-1.666666716e-01f + t * (4.166666791e-02f +
                   t * (-8.333330974e-03f + t * (1.388740959e-03f + t * (-1.983609400e-04f +
                   t * (2.522676368e-05f + t * (-2.834696488e-06f))))))));
  exp_neg(t, E, U0);
  const bool thin = fabsf(t) <= 1.0f;
  U1 = thin ? t * ps : t - U0;
  const float qs = t * t * (3.333333433e-01f + t * (-8.333329856e-02f + t * (1.666668802e-02f +
                   t * (-2.777901944e-03f + t * (3.968053206e-04f + t * (-4.948913192e-05f +
                   t * (5.514686109e-06f + t * (-5.894584660e-07f + t * 5.153443894e-08f))))))));
  q = fabsf(t) <= 2.0f ? qs : t - 2.0f * (U1 / t);
}

// FP64 form: exact weights, the plain FP64 reference of the sweep.
__device__ __forceinline__ void lin_weights(double t, double &E, double &W, double &Psi, double &Esc) {
  E = exp(-t);
  const double u0 = -expm1(-t);
  if (fabs(t) < 1e-2) {
    const double ps = t * (0.5 + t * (-1.0 / 6 + t * (1.0 / 24 + t * (-1.0 / 120 + t * (1.0 / 720 +
                      t * (-1.0 / 5040 + t * (1.0 / 40320)))))));
    Psi = ps;
    Esc = 1.0 - ps;
    W = u0 - ps;
  } else {
    const double ut = u0 / t;
    Psi = 1.0 - ut;
    Esc = ut;
    W = ut - E;
  }
}

// Waits for this wave's LDS traffic and meets the workgroup at a barrier; global
// loads issued earlier stay in flight across it (only LDS is shared here).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int wrap(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

// periodic index without a division for -n <= v < 2n; planes narrower than the
// 4-point stencil (n < 4) take the modulo
__device__ __forceinline__ int wrap_near(int v, int n) {
  if (n < 4) return wrap(v, n);
  return v < 0 ? v + n : (v >= n ? v - n : v);
}

}  // namespace dev
}  // namespace m3drt
#endif
