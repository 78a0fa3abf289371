// Batched synthesis of plane-parallel LTE atmospheres of uneven size (m3drt_b1d_*).
//
// Per (atmosphere, frequency) one lane walks the depth points from the bottom up,
// holding the intensity of every upward mu in registers.  At each point it sums the
// atmosphere's lines onto the continuum -- the QLG-bot profile K(a, v) of
// HOP/qlg_mod.f90:538-627 (qlg_bot_line_eval on the two 9020-entry tables, held in
// LDS), each line only where the lane is inside the line's window at that depth (the
// windowing of add_opacity_equi, HOP/linelist_mod.f90:853-906: lines below the
// tolerance against the continuum are skipped) -- forms S = B_nu(T), and takes the
// linear step of long_char2 (SRC/rt_solvers_mod.f90:527-663) in tau500 for each mu:
// dt = (x_up + x) / 2 * dtau500 / mu with x = chi / chi500.  Nothing per (f, depth) is
// stored: the opacity and source cubes of the reference's 1D run never reach HBM, so
// the kernel is bound by the VALU work of the line sums, not by memory.
//
// Mapping: a wavefront takes one item = (atmosphere, 64 FPL consecutive frequencies);
// lane l holds frequencies f0 + l + 64 j, j < FPL.  Per depth point the wave-uniform
// values (step, temperature, Doppler speed, continuum) are scalar loads shared by the
// 64 lanes; the lines that can reach the item are found with one vector load of 64
// lines' frequency ranges and a ballot, and only their profile records are read.  The persistent kernel's waves
// take items from a counter until the list (sorted by decreasing cost) is exhausted;
// the plain kernel gives each atmosphere its own launch over a static grid of items.
#include "m3drt_internal.hpp"
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

struct m3drt_b1d_plan {
  m3drt_ctx *ctx = nullptr;
  m3drt_b1d_batch batch{};
  m3drt_b1d_tile tile{1, 256};
  double times[6] = {-1, -1, -1, -1, -1, -1};
  bool tuned = false;
  int64_t nrec = 0;
  float *coef = nullptr;        // (line, depth) profile records, 8 floats each
  int2 *win = nullptr;          // (line, depth) conservative frequency range of the window
  int2 *items = nullptr;        // (atmosphere, first frequency)
  int32_t nitems = 0;
  int32_t *counter = nullptr;
  int grid = 0;
  hipStream_t stream = nullptr;  // the graph's own stream (a context may run on the null stream)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace m3drt {
namespace {

constexpr int QLG_NX = 9019;
constexpr int kTab = 2 * (QLG_NX + 1);

struct B1dArgs {
  const int32_t *nd, *nf, *nmu, *nline;
  const int64_t *d_off, *l_off, *ld_off, *mu_off, *f_off, *i_off;
  const double *lam0, *dlam;
  const float *dtau, *hckT, *ivth, *cont_a, *cont_b;
  const float *line_off, *line_cl;
  const float4 *coef;  // two float4 per (line, depth): (ih c/(lam_l v_D), bL, gG, cZH), (cL2, s, reach, lam_l - lam0)
  const int2 *win;     // per (line, depth): first and last frequency the window can reach
  const float *mu, *wflux;
  float *I_top, *flux;
  const int2 *items;
  int32_t nitems;
  int32_t *counter;
  float ilam_ref;
  const float *tab;
};

// The per-(line, depth) setup of qlg_bot_line (HOP/qlg_mod.f90:230-240, 616-627):
// half-width ih, Lorentz and Gauss weights and the two correction coefficients, as
// the model-atom rates form them (struct Qlg of rates.hip), with the Doppler scale
// c / (lam_l v_D) folded into ih, the strength, the window half-width and the line's
// offset: one 32-byte record, read by one scalar load per line in the sweep.  One
// workgroup per atmosphere; its records are depth-major.
__global__ void __launch_bounds__(256) b1d_coef(const int32_t *__restrict__ nd_, const int32_t *__restrict__ nl_,
                                                const int64_t *__restrict__ d_off, const int64_t *__restrict__ l_off,
                                                const int64_t *__restrict__ ld_off, const float *__restrict__ ivth,
                                                const float *__restrict__ line_off, const float *__restrict__ line_cl,
                                                const double *__restrict__ dlam_, const float4 *__restrict__ rec,
                                                float4 *__restrict__ coef, int2 *__restrict__ win) {
  const int m = blockIdx.x;
  const double idl = 1.0 / dlam_[m];
  const int nd = nd_[m], nl = nl_[m];
  const int64_t d0 = d_off[m], l0 = l_off[m], r0 = ld_off[m];
  for (int i = threadIdx.x; i < nd * nl; i += blockDim.x) {
    const int d = i / nl, l = i - d * nl;
    const float4 r = rec[r0 + i];
    const float ad = r.x;
    // ih rounded step by step as bot_setup forms it (HOP/qlg_mod.f90:524-540, no
    // fused multiply-add; IEEE division), and the Doppler scale as two products:
    // the profile's table index follows from them
    float ih;
    {
#pragma clang fp contract(off)
      const float dd = 1.0f + ad * 0.60602209212341318f;
      ih = dd / (ad * dd + 0.8325546112f);
    }
    const float rr = ad * ih, qq = rr * (1.0f - rr), scl = 1.7724538509f * ih, qs = qq * scl;
    const float x = 2.0f * rr - 1.0f;
    const float bL = (0.3183098862f * rr) * scl;
    const float gG = (0.4697186f * (1.0f - rr)) * scl + qs * (-0.064924393191013452f + x * (-0.11019326278794868f));
    const float cZH = qs * (0.38494195930078284f + x * (-0.10567135279610407f));
    const float cL2 = qs * (0.04054415746208978f + x * (0.16241029253712097f));
    const float xsc = (line_cl[l0 + l] * ivth[d0 + d]) * ih;
    coef[2 * (r0 + i)] = make_float4(xsc, bL, gG, cZH);
    coef[2 * (r0 + i) + 1] = make_float4(cL2, r.y, r.z, line_off[l0 + l]);
    // a frequency range that holds every f of the window (one more each side than
    // the window's edges in FP64): a wave skips a line whose range misses its tile
    // on a scalar test, and the lanes decide on the FP32 distance
    int2 w = make_int2(0x7fffffff, -1);
    if (r.z >= 0.0f) {
      const double lo = ((double)line_off[l0 + l] - (double)r.z) * idl, hi = ((double)line_off[l0 + l] + (double)r.z) * idl;
      w.x = (int)fmax(floor(lo) - 1.0, -1.0);
      w.y = (int)fmin(ceil(hi) + 1.0, 2147483000.0);
    }
    win[r0 + i] = w;
  }
}

__global__ void b1d_reset(int32_t *counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

using dev::cptr;
using dev::ro;

// One wave's item: atmosphere m, frequencies f0 + lane + 64 j.
template <int FPL>
__device__ __forceinline__ void b1d_item(const B1dArgs &a, const float *TGX, const float *TH2X, int m, int f0) {
  const int lane = threadIdx.x & 63;
  const int nd = ro(a.nd)[m], nf = ro(a.nf)[m], nmu = ro(a.nmu)[m], nl = ro(a.nline)[m];
  const int64_t d0 = ro(a.d_off)[m], l0 = ro(a.l_off)[m], r0 = ro(a.ld_off)[m], m0 = ro(a.mu_off)[m];
  const double lam0 = ro(a.lam0)[m], dlam = ro(a.dlam)[m];
  float xf[FPL], rl[FPL], r3[FPL];
  bool on[FPL];
#pragma unroll
  for (int j = 0; j < FPL; ++j) {
    const int f = f0 + lane + 64 * j;
    on[j] = f < nf;
    const double lam = lam0 + (double)f * dlam;
    xf[j] = (float)((double)f * dlam);  // offset from lam0, the lines' frame
    rl[j] = (float)(1.0 / lam);
    const float q = (float)lam * a.ilam_ref;
    r3[j] = q * q * q;
  }
  float imu[M3DRT_B1D_MAXMU];
#pragma unroll
  for (int k = 0; k < M3DRT_B1D_MAXMU; ++k) imu[k] = k < nmu ? 1.0f / ro(a.mu)[m0 + k] : 0.0f;

  float I[M3DRT_B1D_MAXMU][FPL], xu[FPL], Su[FPL];
  // opacity ratio and source function at depth point d
  const cptr<float> cont_a = ro(a.cont_a), cont_b = ro(a.cont_b), hckT = ro(a.hckT);
  const int flo = f0, fhi = f0 + 64 * FPL - 1;
  auto point = [&](int d, float *x, float *S) {
    const int64_t dd = d0 + d;
    const float ca = cont_a[dd], cb = cont_b[dd], hc = hckT[dd];
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      x[j] = ca + cb * r3[j];
      const float e = __builtin_amdgcn_exp2f(-1.44269504088896341f * hc * rl[j]);
      S[j] = e / (1.0f - e);
    }
    // the lines whose window can reach the wave's frequencies: 64 lines' ranges per
    // vector load, one lane each, gathered into a wave mask; then only those lines'
    // records, each read (scalar) while the previous one is summed, and the lanes'
    // own FP32 test
    const cptr<float4> c = ro(a.coef) + 2 * (r0 + (int64_t)d * nl);
    const int2 *wn = a.win + (r0 + (int64_t)d * nl);
    for (int base = 0; base < nl; base += 64) {
      bool hit = false;
      if (base + lane < nl) {
        const int2 w = wn[base + lane];
        hit = !(w.y < flo || w.x > fhi);
      }
      uint64_t m = __ballot(hit);
      if (!m) continue;
      int cur = base + (int)__builtin_ctzll(m);
      m &= m - 1;
      float4 p = c[2 * cur], q = c[2 * cur + 1];
      for (;;) {
        const int nxt = m ? base + (int)__builtin_ctzll(m) : cur;
        const float4 pn = c[2 * nxt], qn = c[2 * nxt + 1];
#pragma unroll
        for (int j = 0; j < FPL; ++j) {
          const float dl = fabsf(xf[j] - q.w);
          if (dl <= q.z) {
            // qlg_bot_line_eval, HOP/qlg_mod.f90:556-571
            const float xs = dl * p.x, u = xs * xs, L = dev::rcp(1.0f + u);
            const int ig = (int)roundf(fminf(xs * 800.0f, (float)QLG_NX));
            const float K = TGX[ig] * p.z + L * (p.y + L * (q.x + p.w * TH2X[ig]));
            x[j] += q.y * K;
          }
        }
        if (!m) break;
        m &= m - 1;
        p = pn;
        q = qn;
      }
    }
  };

  point(nd - 1, xu, Su);
#pragma unroll
  for (int k = 0; k < M3DRT_B1D_MAXMU; ++k)
#pragma unroll
    for (int j = 0; j < FPL; ++j) I[k][j] = Su[j];
  for (int d = nd - 2; d >= 0; --d) {
    float x[FPL], S[FPL];
    point(d, x, S);
    const float dt500 = ro(a.dtau)[d0 + d];
#pragma unroll
    for (int k = 0; k < M3DRT_B1D_MAXMU; ++k) {
      if (k >= nmu) break;
#pragma unroll
      for (int j = 0; j < FPL; ++j) {
        const float dt = 0.5f * (xu[j] + x[j]) * (dt500 * imu[k]);
        float E, W, Ps, Es;
        dev::lin_weights(dt, E, W, Ps, Es);
        I[k][j] = I[k][j] * E + W * Su[j] + Ps * S[j];
      }
    }
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      xu[j] = x[j];
      Su[j] = S[j];
    }
  }
  const int64_t io = ro(a.i_off)[m], fo = ro(a.f_off)[m];
#pragma unroll
  for (int j = 0; j < FPL; ++j) {
    if (!on[j]) continue;
    const int f = f0 + lane + 64 * j;
    float F = 0.0f;
#pragma unroll
    for (int k = 0; k < M3DRT_B1D_MAXMU; ++k) {
      if (k >= nmu) break;
      a.I_top[io + (int64_t)k * nf + f] = I[k][j];
      F += ro(a.wflux)[m0 + k] * I[k][j];
    }
    a.flux[fo + f] = F;
  }
}

__device__ __forceinline__ void load_tab(float *tab, const float *src) {
  for (int i = threadIdx.x; i < kTab; i += blockDim.x) tab[i] = src[i];
  __syncthreads();
}

// Persistent form: every wave takes items until the counter passes the list's end,
// so every wave reaches the exit and the grid drains.
template <int FPL, int NT>
__global__ void __launch_bounds__(NT) b1d_persistent(const B1dArgs a) {
  extern __shared__ float tab[];
  load_tab(tab, a.tab);
  const float *TGX = tab, *TH2X = tab + QLG_NX + 1;
  for (;;) {
    int it = 0;
    if ((threadIdx.x & 63) == 0) it = atomicAdd(a.counter, 1);
    it = __shfl(it, 0, 64);
    if (it >= a.nitems) break;
    const int2 w = ro(a.items)[uniform(it)];
    b1d_item<FPL>(a, TGX, TH2X, uniform(w.x), uniform(w.y));
  }
}

// Plain form: one atmosphere per launch, wave w of the grid takes frequencies
// from 64 FPL w.
template <int FPL, int NT>
__global__ void __launch_bounds__(NT) b1d_plain(const B1dArgs a, int m, int ntile) {
  extern __shared__ float tab[];
  load_tab(tab, a.tab);
  const float *TGX = tab, *TH2X = tab + QLG_NX + 1;
  const int w = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  if (w >= ntile) return;
  b1d_item<FPL>(a, TGX, TH2X, m, w * 64 * FPL);
}

// ------------------------------------------------------------------ host

struct Host {
  std::vector<int32_t> nd, nf, nmu, nl;
  std::vector<int64_t> ld_off;
};

int fetch_host(const m3drt_b1d_batch *b, Host &h) {
  const size_t n = (size_t)b->natm;
  h.nd.resize(n); h.nf.resize(n); h.nmu.resize(n); h.nl.resize(n); h.ld_off.resize(n);
  M3DRT_HIP(hipMemcpy(h.nd.data(), b->nd, n * 4, hipMemcpyDeviceToHost));
  M3DRT_HIP(hipMemcpy(h.nf.data(), b->nf, n * 4, hipMemcpyDeviceToHost));
  M3DRT_HIP(hipMemcpy(h.nmu.data(), b->nmu, n * 4, hipMemcpyDeviceToHost));
  M3DRT_HIP(hipMemcpy(h.nl.data(), b->nline, n * 4, hipMemcpyDeviceToHost));
  M3DRT_HIP(hipMemcpy(h.ld_off.data(), b->ld_off, n * 8, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) {
    if (h.nd[i] < 2 || h.nf[i] < 1 || h.nmu[i] < 1 || h.nmu[i] > M3DRT_B1D_MAXMU || h.nl[i] < 0)
      return set_error(M3DRT_EINVAL, "m3drt_b1d: atmosphere " + std::to_string(i) +
                                         " needs nd >= 2, nf >= 1, 1 <= nmu <= 8, nline >= 0");
  }
  return M3DRT_OK;
}

int64_t records(const Host &h) {
  int64_t n = 0;
  for (size_t i = 0; i < h.nd.size(); ++i) n = std::max<int64_t>(n, h.ld_off[i] + (int64_t)h.nd[i] * h.nl[i]);
  return n;
}

B1dArgs make_args(const m3drt_b1d_batch *b, const float *coef, const int2 *win, const float *tab) {
  B1dArgs a;
  a.nd = b->nd; a.nf = b->nf; a.nmu = b->nmu; a.nline = b->nline;
  a.d_off = b->d_off; a.l_off = b->l_off; a.ld_off = b->ld_off; a.mu_off = b->mu_off;
  a.f_off = b->f_off; a.i_off = b->i_off;
  a.lam0 = b->lam0; a.dlam = b->dlam;
  a.dtau = b->dtau; a.hckT = b->hckT; a.ivth = b->ivth; a.cont_a = b->cont_a; a.cont_b = b->cont_b;
  a.line_off = b->line_off; a.line_cl = b->line_cl;
  a.coef = reinterpret_cast<const float4 *>(coef);
  a.win = win;
  a.mu = b->mu; a.wflux = b->wflux;
  a.I_top = b->I_top; a.flux = b->flux;
  a.items = nullptr; a.nitems = 0; a.counter = nullptr;
  a.ilam_ref = (float)(1.0 / b->lam_ref);
  a.tab = tab;
  return a;
}

int run_coef(const m3drt_b1d_batch *b, float *coef, int2 *win, int64_t nrec, hipStream_t s) {
  if (nrec == 0) return M3DRT_OK;
  hipLaunchKernelGGL(b1d_coef, dim3(b->natm), dim3(256), 0, s, b->nd, b->nline, b->d_off, b->l_off, b->ld_off,
                     b->ivth, b->line_off, b->line_cl, b->dlam, reinterpret_cast<const float4 *>(b->line_rec),
                     reinterpret_cast<float4 *>(coef), win);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

bool valid_tile(const m3drt_b1d_tile &t) {
  return (t.fpl == 1 || t.fpl == 2) && (t.threads == 256 || t.threads == 512 || t.threads == 1024);
}

template <int FPL, int NT>
const void *persistent_fn() { return (const void *)b1d_persistent<FPL, NT>; }

const void *persistent_kernel(const m3drt_b1d_tile &t) {
  if (t.fpl == 1) {
    if (t.threads == 256) return persistent_fn<1, 256>();
    if (t.threads == 512) return persistent_fn<1, 512>();
    return persistent_fn<1, 1024>();
  }
  if (t.threads == 256) return persistent_fn<2, 256>();
  if (t.threads == 512) return persistent_fn<2, 512>();
  return persistent_fn<2, 1024>();
}

template <int FPL, int NT>
int launch_plain_t(m3drt_ctx *ctx, const B1dArgs &a, int m, int nf) {
  const int ntile = (nf + 64 * FPL - 1) / (64 * FPL);
  const int blocks = (ntile + NT / 64 - 1) / (NT / 64);
  auto k = b1d_plain<FPL, NT>;
  M3DRT_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, kTab * 4));
  hipLaunchKernelGGL(k, dim3(blocks), dim3(NT), kTab * 4, ctx->stream, a, m, ntile);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

int launch_plain(m3drt_ctx *ctx, const m3drt_b1d_tile &t, const B1dArgs &a, int m, int nf) {
  if (t.fpl == 1) {
    if (t.threads == 256) return launch_plain_t<1, 256>(ctx, a, m, nf);
    if (t.threads == 512) return launch_plain_t<1, 512>(ctx, a, m, nf);
    return launch_plain_t<1, 1024>(ctx, a, m, nf);
  }
  if (t.threads == 256) return launch_plain_t<2, 256>(ctx, a, m, nf);
  if (t.threads == 512) return launch_plain_t<2, 512>(ctx, a, m, nf);
  return launch_plain_t<2, 1024>(ctx, a, m, nf);
}

// Work list for a tile width: every (atmosphere, 64 fpl frequencies), the costliest
// first (depth points x (lines + 4 mu) per lane), so the list's tail is short items.
std::vector<int2> work_list(const Host &h, int fpl) {
  std::vector<std::pair<int64_t, int2>> v;
  const int w = 64 * fpl;
  for (size_t m = 0; m < h.nd.size(); ++m) {
    const int64_t cost = (int64_t)h.nd[m] * (h.nl[m] + 4 * h.nmu[m]);
    for (int f0 = 0; f0 < h.nf[m]; f0 += w) v.push_back({cost, make_int2((int)m, f0)});
  }
  std::stable_sort(v.begin(), v.end(), [](const auto &p, const auto &q) { return p.first > q.first; });
  std::vector<int2> out(v.size());
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i].second;
  return out;
}

int persistent_grid(m3drt_ctx *ctx, const m3drt_b1d_tile &t, int *grid) {
  const void *k = persistent_kernel(t);
  M3DRT_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kTab * 4));
  int per_cu = 0;
  M3DRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, t.threads, kTab * 4));
  int ncu = 0;
  M3DRT_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  if (per_cu < 1) return set_error(M3DRT_ELIMIT, "m3drt_b1d: the tile shape does not fit a compute unit");
  *grid = per_cu * ncu;
  return M3DRT_OK;
}

int launch_persistent(const m3drt_b1d_tile &t, int grid, const B1dArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(b1d_reset, dim3(1), dim3(64), 0, s, a.counter);
  M3DRT_HIP(hipGetLastError());
  B1dArgs arg = a;
  void *params[] = {&arg};
  M3DRT_HIP(hipLaunchKernel(persistent_kernel(t), dim3(grid), dim3(t.threads), params, kTab * 4, s));
  return M3DRT_OK;
}

void free_plan(m3drt_b1d_plan *p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->ev_in) (void)hipEventDestroy(p->ev_in);
  if (p->ev_out) (void)hipEventDestroy(p->ev_out);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  if (p->coef) (void)hipFree(p->coef);
  if (p->win) (void)hipFree(p->win);
  if (p->items) (void)hipFree(p->items);
  if (p->counter) (void)hipFree(p->counter);
  delete p;
}

}  // namespace

int b1d_synth_plain(m3drt_ctx *ctx, const m3drt_b1d_batch *b, const m3drt_b1d_tile *tile) {
  const m3drt_b1d_tile t = tile ? *tile : m3drt_b1d_tile{1, 256};
  if (!valid_tile(t)) return set_error(M3DRT_EINVAL, "m3drt_b1d_synth_plain: fpl must be 1 or 2, threads 256/512/1024");
  Host h;
  int rc = fetch_host(b, h);
  if (rc) return rc;
  rc = qlg_tables(ctx);
  if (rc) return rc;
  const int64_t nrec = records(h);
  rc = ensure_scratch(ctx, (size_t)std::max<int64_t>(nrec, 1) * 40);
  if (rc) return rc;
  float *coef = static_cast<float *>(ctx->scratch);
  int2 *win = reinterpret_cast<int2 *>(coef + 8 * std::max<int64_t>(nrec, 1));
  rc = run_coef(b, coef, win, nrec, ctx->stream);
  if (rc) return rc;
  const B1dArgs a = make_args(b, coef, win, ctx->qlg_tab);
  for (int m = 0; m < b->natm; ++m) {
    rc = launch_plain(ctx, t, a, m, h.nf[m]);
    if (rc) return rc;
  }
  return M3DRT_OK;
}

int b1d_plan_create(m3drt_ctx *ctx, const m3drt_b1d_batch *b, const m3drt_b1d_tile *tile, m3drt_b1d_plan **out) {
  if (tile && !valid_tile(*tile))
    return set_error(M3DRT_EINVAL, "m3drt_b1d_plan_create: fpl must be 1 or 2, threads 256/512/1024");
  Host h;
  int rc = fetch_host(b, h);
  if (rc) return rc;
  rc = qlg_tables(ctx);
  if (rc) return rc;
  m3drt_b1d_plan *p = new m3drt_b1d_plan();
  p->ctx = ctx;
  p->batch = *b;
  p->nrec = records(h);
  auto fail = [&](int code) {
    free_plan(p);
    return code;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess) return fail(hip_check(e, "hipStreamCreate"));
  if ((e = hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming)) != hipSuccess) return fail(hip_check(e, "hipEventCreate"));
  if ((e = hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming)) != hipSuccess) return fail(hip_check(e, "hipEventCreate"));
  if ((e = hipMalloc(&p->coef, (size_t)std::max<int64_t>(p->nrec, 1) * 32)) != hipSuccess)
    return fail(set_error(M3DRT_ENOMEM, "m3drt_b1d_plan_create: profile coefficients"));
  if ((e = hipMalloc(&p->win, (size_t)std::max<int64_t>(p->nrec, 1) * 8)) != hipSuccess)
    return fail(set_error(M3DRT_ENOMEM, "m3drt_b1d_plan_create: line windows"));
  if ((e = hipMalloc(&p->counter, 64)) != hipSuccess) return fail(set_error(M3DRT_ENOMEM, "m3drt_b1d_plan_create: counter"));
  // the caller's earlier work (the batch's inputs) before anything of the plan's
  if ((e = hipEventRecord(p->ev_in, ctx->stream)) != hipSuccess) return fail(hip_check(e, "hipEventRecord"));
  if ((e = hipStreamWaitEvent(p->stream, p->ev_in, 0)) != hipSuccess) return fail(hip_check(e, "hipStreamWaitEvent"));
  if ((rc = run_coef(b, p->coef, p->win, p->nrec, p->stream))) return fail(rc);
  B1dArgs a = make_args(b, p->coef, p->win, ctx->qlg_tab);
  a.counter = p->counter;

  static const m3drt_b1d_tile cands[6] = {{1, 256}, {1, 512}, {1, 1024}, {2, 256}, {2, 512}, {2, 1024}};
  if (tile) {
    p->tile = *tile;
  } else {
    // autotune: each candidate on a strided sample of its own work list (every 8th
    // item: the same mix of costs as the whole), one untimed run, then the best of two
    hipEvent_t t0, t1;
    if ((e = hipEventCreate(&t0)) != hipSuccess) return fail(hip_check(e, "hipEventCreate"));
    if ((e = hipEventCreate(&t1)) != hipSuccess) { (void)hipEventDestroy(t0); return fail(hip_check(e, "hipEventCreate")); }
    int best = -1;
    for (int c = 0; c < 6 && rc == 0; ++c) {
      const std::vector<int2> all = work_list(h, cands[c].fpl);
      std::vector<int2> sample;
      for (size_t i = 0; i < all.size(); i += 8) sample.push_back(all[i]);
      int2 *d_items = nullptr;
      if (hipMalloc(&d_items, std::max<size_t>(sample.size(), 1) * sizeof(int2)) != hipSuccess) {
        rc = set_error(M3DRT_ENOMEM, "m3drt_b1d_plan_create: autotune work list");
        break;
      }
      if ((e = hipMemcpy(d_items, sample.data(), sample.size() * sizeof(int2), hipMemcpyHostToDevice)) != hipSuccess) {
        rc = hip_check(e, "hipMemcpy");
      }
      int grid = 0;
      if (!rc) rc = persistent_grid(ctx, cands[c], &grid);
      B1dArgs s = a;
      s.items = d_items;
      s.nitems = (int32_t)sample.size();
      double tbest = -1.0;
      for (int r = 0; r < 3 && !rc; ++r) {
        if ((e = hipEventRecord(t0, p->stream)) != hipSuccess) { rc = hip_check(e, "hipEventRecord"); break; }
        if ((rc = launch_persistent(cands[c], grid, s, p->stream))) break;
        if ((e = hipEventRecord(t1, p->stream)) != hipSuccess) { rc = hip_check(e, "hipEventRecord"); break; }
        if ((e = hipEventSynchronize(t1)) != hipSuccess) { rc = hip_check(e, "hipEventSynchronize"); break; }
        float ms = 0.0f;
        if ((e = hipEventElapsedTime(&ms, t0, t1)) != hipSuccess) { rc = hip_check(e, "hipEventElapsedTime"); break; }
        if (r > 0 && (tbest < 0 || ms < tbest)) tbest = ms;
      }
      (void)hipFree(d_items);
      if (rc) break;
      p->times[c] = tbest;
      if (best < 0 || tbest < p->times[best]) best = c;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return fail(rc);
    p->tile = cands[best];
    p->tuned = true;
  }
  const std::vector<int2> list = work_list(h, p->tile.fpl);
  p->nitems = (int32_t)list.size();
  if ((e = hipMalloc(&p->items, std::max<size_t>(list.size(), 1) * sizeof(int2))) != hipSuccess)
    return fail(set_error(M3DRT_ENOMEM, "m3drt_b1d_plan_create: work list"));
  if ((e = hipMemcpy(p->items, list.data(), list.size() * sizeof(int2), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(hip_check(e, "hipMemcpy"));
  if ((rc = persistent_grid(ctx, p->tile, &p->grid))) return fail(rc);
  a.items = p->items;
  a.nitems = p->nitems;
  // the graph: counter reset (a kernel, not a memset node) and the persistent kernel
  if ((e = hipStreamSynchronize(p->stream)) != hipSuccess) return fail(hip_check(e, "hipStreamSynchronize"));
  if ((e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
    return fail(hip_check(e, "hipStreamBeginCapture"));
  rc = launch_persistent(p->tile, p->grid, a, p->stream);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(p->stream, &g);
  p->graph = g;
  if (rc) return fail(rc);
  if (e != hipSuccess) return fail(hip_check(e, "hipStreamEndCapture"));
  if ((e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0)) != hipSuccess)
    return fail(hip_check(e, "hipGraphInstantiate"));
  *out = p;
  return M3DRT_OK;
}

int b1d_plan_run(m3drt_b1d_plan *p) {
  m3drt_ctx *ctx = p->ctx;
  M3DRT_HIP(hipEventRecord(p->ev_in, ctx->stream));
  M3DRT_HIP(hipStreamWaitEvent(p->stream, p->ev_in, 0));
  int rc = run_coef(&p->batch, p->coef, p->win, p->nrec, p->stream);
  if (rc) return rc;
  M3DRT_HIP(hipGraphLaunch(p->exec, p->stream));
  M3DRT_HIP(hipEventRecord(p->ev_out, p->stream));
  M3DRT_HIP(hipStreamWaitEvent(ctx->stream, p->ev_out, 0));
  return M3DRT_OK;
}

int b1d_plan_info(const m3drt_b1d_plan *p, m3drt_b1d_tile *tile, double *times, int32_t *ntimes) {
  if (tile) *tile = p->tile;
  if (times)
    for (int i = 0; i < 6; ++i) times[i] = p->times[i];
  if (ntimes) *ntimes = 6;
  return M3DRT_OK;
}

int b1d_plan_destroy(m3drt_b1d_plan *p) {
  if (p && p->stream) (void)hipStreamSynchronize(p->stream);
  free_plan(p);
  return M3DRT_OK;
}

}  // namespace m3drt
