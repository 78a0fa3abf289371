// C ABI of libm3drt (include/m3drt.h): context, error state, memory helpers and
// argument checking in front of the kernels.  Status conventions follow the
// reference's HOP C ABI (HOP/standalone/hop_c_api.f90:1-18): int status, message
// by m3drt_last_error, no process exit in library code.
#include "m3drt_internal.hpp"
#include <cstring>
#include <cstdio>
#include <string>

namespace m3drt {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_check(hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return set_error(e == hipErrorOutOfMemory ? M3DRT_ENOMEM : M3DRT_EDEVICE, m);
}

int ensure_scratch(m3drt_ctx *ctx, size_t bytes) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  M3DRT_HIP(hipStreamIsCapturing(ctx->stream, &cs));
  const bool capturing = cs != hipStreamCaptureStatusNone;
  if (bytes <= ctx->scratch_bytes) {
    if (capturing && !ctx->scratch_fixed) ctx->scratch_captured = true;
    return M3DRT_OK;
  }
  if (ctx->scratch_fixed)
    return set_error(M3DRT_ELIMIT, "this call needs " + std::to_string(bytes) + " bytes of scratch, beyond the fixed " +
                                       std::to_string(ctx->scratch_bytes) + " bytes of m3drt_sc_sweep_noalloc");
  if (capturing)
    return set_error(M3DRT_ELIMIT, "the scratch buffer would grow inside a stream capture: run the call once "
                                   "before capturing it");
  if (ctx->scratch) {
    if (ctx->scratch_captured) {
      ctx->retired.push_back(ctx->scratch);  // captured graphs still replay into it
    } else {
      M3DRT_HIP(hipStreamSynchronize(ctx->stream));
      M3DRT_HIP(hipFree(ctx->scratch));
    }
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    ctx->scratch_captured = false;
  }
  const size_t rounded = (bytes + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
  M3DRT_HIP(hipMalloc(&ctx->scratch, rounded));
  ctx->scratch_bytes = rounded;
  return M3DRT_OK;
}

}  // namespace m3drt

using m3drt::set_error;

extern "C" {

int m3drt_last_error(char *buf, int buflen) {
  if (!buf || buflen <= 0) return M3DRT_EINVAL;
  const std::string &m = m3drt::g_last_error;
  const size_t n = std::min<size_t>(m.size(), (size_t)buflen - 1);
  std::memcpy(buf, m.data(), n);
  buf[n] = '\0';
  return M3DRT_OK;
}

int m3drt_create(m3drt_ctx **ctx, int device, int flags) {
  (void)flags;
  if (!ctx) return set_error(M3DRT_EINVAL, "m3drt_create: ctx is NULL");
  *ctx = nullptr;
  int n = 0;
  M3DRT_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_error(M3DRT_EINVAL, "m3drt_create: no such device");
  M3DRT_HIP(hipSetDevice(device));
  {
    const int rc = m3drt::noalloc_setup(device);
    if (rc) return rc;
  }
  hipDeviceProp_t prop;
  M3DRT_HIP(hipGetDeviceProperties(&prop, device));
  m3drt_ctx *c = new m3drt_ctx();
  c->device = device;
  c->stream = nullptr;
  c->lds_max = (int)prop.sharedMemPerBlockOptin > 0 ? (int)prop.sharedMemPerBlockOptin : (int)prop.maxSharedMemoryPerMultiProcessor;
  if (c->lds_max <= 0 || c->lds_max > 160 * 1024) c->lds_max = (int)prop.sharedMemPerBlock;
  *ctx = c;
  return M3DRT_OK;
}

int m3drt_destroy(m3drt_ctx *ctx) {
  if (!ctx) return M3DRT_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->scratch || ctx->qlg_tab || !ctx->retired.empty()) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  for (void *p : ctx->retired) (void)hipFree(p);
  if (ctx->qlg_tab) (void)hipFree(ctx->qlg_tab);
  delete ctx;
  return M3DRT_OK;
}

int m3drt_set_stream(m3drt_ctx *ctx, void *hip_stream) {
  if (!ctx) return set_error(M3DRT_EINVAL, "m3drt_set_stream: ctx is NULL");
  ctx->stream = static_cast<hipStream_t>(hip_stream);
  return M3DRT_OK;
}

int m3drt_synchronize(m3drt_ctx *ctx) {
  if (!ctx) return set_error(M3DRT_EINVAL, "m3drt_synchronize: ctx is NULL");
  M3DRT_HIP(hipSetDevice(ctx->device));
  M3DRT_HIP(hipStreamSynchronize(ctx->stream));
  return M3DRT_OK;
}

int m3drt_device_alloc(m3drt_ctx *ctx, size_t bytes, void **ptr) {
  if (!ctx || !ptr) return set_error(M3DRT_EINVAL, "m3drt_device_alloc: NULL argument");
  M3DRT_HIP(hipSetDevice(ctx->device));
  M3DRT_HIP(hipMalloc(ptr, bytes));
  return M3DRT_OK;
}

int m3drt_device_free(m3drt_ctx *ctx, void *ptr) {
  if (!ctx) return set_error(M3DRT_EINVAL, "m3drt_device_free: ctx is NULL");
  M3DRT_HIP(hipSetDevice(ctx->device));
  M3DRT_HIP(hipFree(ptr));
  return M3DRT_OK;
}

int m3drt_copy_to_device(m3drt_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return set_error(M3DRT_EINVAL, "m3drt_copy_to_device: NULL argument");
  M3DRT_HIP(hipSetDevice(ctx->device));
  M3DRT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  M3DRT_HIP(hipStreamSynchronize(ctx->stream));
  return M3DRT_OK;
}

int m3drt_copy_to_host(m3drt_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return set_error(M3DRT_EINVAL, "m3drt_copy_to_host: NULL argument");
  M3DRT_HIP(hipSetDevice(ctx->device));
  M3DRT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  M3DRT_HIP(hipStreamSynchronize(ctx->stream));
  return M3DRT_OK;
}

int m3drt_native_strides(m3drt_sweep_desc *d) {
  if (!d) return set_error(M3DRT_EINVAL, "m3drt_native_strides: desc is NULL");
  const int64_t nx = d->nx, ny = d->ny, nf = d->nf;
  d->cs_x = 1;
  d->cs_y = nx;
  d->cs_f = nx * ny;
  d->cs_z = nx * ny * nf;
  d->gs_x = 1;
  d->gs_y = nx;
  d->gs_f = nx * ny;
  if (d->eta_ld < nf) d->eta_ld = nf;
  return M3DRT_OK;
}

int m3drt_reference_strides(m3drt_sweep_desc *d) {
  if (!d) return set_error(M3DRT_EINVAL, "m3drt_reference_strides: desc is NULL");
  const int64_t nx = d->nx, ny = d->ny, nf = d->nf;
  d->cs_f = 1;
  d->cs_x = nf;
  d->cs_y = nf * nx;
  d->cs_z = nf * nx * ny;
  d->gs_f = 1;
  d->gs_x = nf;
  d->gs_y = nf * nx;
  if (d->eta_ld < nf) d->eta_ld = nf;
  return M3DRT_OK;
}

}  // extern "C"

namespace m3drt {

int sc_sweep_checked(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b, const char *who) {
  const std::string w(who);
  if (!ctx || !d || !b) return set_error(M3DRT_EINVAL, w + ": NULL argument");
  if (d->nf < 1 || d->nx < 1 || d->ny < 1 || d->nz < 1)
    return set_error(M3DRT_EINVAL, w + ": empty grid");
  if (d->dir != 1 && d->dir != -1) return set_error(M3DRT_EINVAL, w + ": dir must be +1 or -1");
  if (d->prec != M3DRT_F32 && d->prec != M3DRT_F64) return set_error(M3DRT_EINVAL, w + ": bad prec");
  if (d->solver != M3DRT_LINEAR2 && d->solver != M3DRT_QUADRATIC && d->solver != M3DRT_LINEAR &&
      d->solver != M3DRT_LONG_LINEAR)
    return set_error(M3DRT_EINVAL,
                     w + ": solver must be M3DRT_LINEAR2, M3DRT_QUADRATIC, M3DRT_LINEAR or M3DRT_LONG_LINEAR");
  if (!(d->ds[0] > 0 && d->ds[1] > 0 && d->ds[2] > 0)) return set_error(M3DRT_EINVAL, w + ": ds must be > 0");
  if (d->vec[2] == 0.0) return set_error(M3DRT_EINVAL, w + ": horizontal ray (vec(3) = 0)");
  if ((d->vec[2] > 0) != (d->dir > 0))
    return set_error(M3DRT_EINVAL, w + ": dir must follow the sign of vec(3)");
  if (!b->chi || !b->S || !b->ghost_chi || !b->ghost_S || !b->ghost_I)
    return set_error(M3DRT_EINVAL, w + ": missing input array");
  if (d->solver != M3DRT_LINEAR2) {
    if (d->cs_f < 1 || d->cs_x < 1 || d->cs_y < 1 || d->cs_z < 1 || d->gs_f < 1 || d->gs_x < 1 || d->gs_y < 1)
      return set_error(M3DRT_EINVAL, w + ": strides must be >= 1 (see m3drt_native_strides)");
    if ((d->flags & M3DRT_SPLIT_RATES) && (!b->Irem_out || !b->escape_out))
      return set_error(M3DRT_EINVAL, w + ": M3DRT_SPLIT_RATES needs Irem_out and escape_out");
    M3DRT_HIP(hipSetDevice(ctx->device));
    if (d->solver == M3DRT_QUADRATIC) return sweep_lcq(ctx, d, b);
    if (d->solver == M3DRT_LONG_LINEAR) return sweep_lcl(ctx, d, b);
    return sweep_sc1(ctx, d, b);
  }
  if (!b->chi500 || !b->g500) return set_error(M3DRT_EINVAL, w + ": missing chi500 / g500");
  if (d->nx * (int64_t)d->ny > 1 && !b->chi500_up)
    return set_error(M3DRT_EINVAL, w + ": short_char2: chi500 at the upwind points is missing");
  if (b->eta_row && (!b->eta_mask || d->eta_ld < d->nf))
    return set_error(M3DRT_EINVAL, w + ": eta_row without eta_mask / eta_ld < nf");
  if ((d->flags & M3DRT_SPLIT_RATES) && (!b->Irem_out || !b->escape_out))
    return set_error(M3DRT_EINVAL, w + ": M3DRT_SPLIT_RATES needs Irem_out and escape_out");
  if (d->cs_f < 1 || d->cs_x < 1 || d->cs_y < 1 || d->cs_z < 1 || d->gs_f < 1 || d->gs_x < 1 || d->gs_y < 1)
    return set_error(M3DRT_EINVAL, w + ": strides must be >= 1 (see m3drt_native_strides)");
  // every in-plane offset of a cube is formed in 64 bits; the planes themselves must not overlap
  M3DRT_HIP(hipSetDevice(ctx->device));
  if (d->nx == 1 && d->ny == 1) return sweep_lc2(ctx, d, b);
  return sweep_sc2(ctx, d, b);
}

}  // namespace m3drt

extern "C" {

int m3drt_sc_sweep(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  return m3drt::sc_sweep_checked(ctx, d, b, "m3drt_sc_sweep");
}

int m3drt_sc_sweep_hermite(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b,
                           const m3drt_hermite_fields *h) {
  if (!ctx || !d || !b || !h) return set_error(M3DRT_EINVAL, "m3drt_sc_sweep_hermite: NULL argument");
  if (d->nf < 1 || d->nx < 1 || d->ny < 1 || d->nz < 1)
    return set_error(M3DRT_EINVAL, "m3drt_sc_sweep_hermite: empty grid");
  if (d->dir != 1 && d->dir != -1) return set_error(M3DRT_EINVAL, "m3drt_sc_sweep_hermite: dir must be +1 or -1");
  if (!(d->ds[0] > 0 && d->ds[1] > 0 && d->ds[2] > 0) || d->vec[2] == 0.0 || (d->vec[2] > 0) != (d->dir > 0))
    return set_error(M3DRT_EINVAL, "m3drt_sc_sweep_hermite: bad geometry (ds > 0, dir following vec(3))");
  if (!b->chi || !b->S || !b->ghost_chi || !b->ghost_S || !b->ghost_I)
    return set_error(M3DRT_EINVAL, "m3drt_sc_sweep_hermite: missing input array");
  if (!h->beta || !h->pe4 || !h->temp || !h->dpe_dx || !h->dpe_dy || !h->dpe_dz || !h->dlnt_dx || !h->dlnt_dy ||
      !h->dlnt_dz || !h->hknu)
    return set_error(M3DRT_EINVAL, "m3drt_sc_sweep_hermite: missing composition field");
  if ((d->flags & M3DRT_SPLIT_RATES) && (!b->Irem_out || !b->escape_out))
    return set_error(M3DRT_EINVAL, "m3drt_sc_sweep_hermite: M3DRT_SPLIT_RATES needs Irem_out and escape_out");
  if (d->cs_f < 1 || d->cs_x < 1 || d->cs_y < 1 || d->cs_z < 1 || d->gs_f < 1 || d->gs_x < 1 || d->gs_y < 1)
    return set_error(M3DRT_EINVAL, "m3drt_sc_sweep_hermite: strides must be >= 1");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::sweep_sch(ctx, d, b, h);
}

static int lc_hermite_entry(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b,
                            const m3drt_hermite_fields *h, const m3drt_lc_ghost *g, bool hybrid) {
  const std::string w = hybrid ? "m3drt_lc_sweep_hermite2: " : "m3drt_lc_sweep_hermite: ";
  if (!ctx || !d || !b || !h || !g) return set_error(M3DRT_EINVAL, w + "NULL argument");
  if (d->nf < 1 || d->nx < 1 || d->ny < 1) return set_error(M3DRT_EINVAL, w + "empty grid");
  // SRC/rt_solvers_mod.f90:1076-1077, 1552-1553
  if (d->nz < 2) return set_error(M3DRT_EINVAL, w + "needs nz >= 2");
  if (d->dir != 1 && d->dir != -1) return set_error(M3DRT_EINVAL, w + "dir must be +1 or -1");
  if (!(d->ds[2] > 0) || d->vec[2] == 0.0 || (d->vec[2] > 0) != (d->dir > 0))
    return set_error(M3DRT_EINVAL, w + "bad geometry (ds(3) > 0, dir following vec(3))");
  if (d->prec != M3DRT_F32) return set_error(M3DRT_ELIMIT, w + "FP32 only (as the reference)");
  if ((d->flags & M3DRT_SPLIT_RATES) || b->eta_row)
    return set_error(M3DRT_EINVAL, w + "takes neither split rates nor eta flags");
  if (!b->chi || !b->S || !b->ghost_chi || !b->ghost_S || !b->ghost_I)
    return set_error(M3DRT_EINVAL, w + "missing input array");
  if (!h->pe4 || !h->dpe_dz || (!hybrid && (!h->beta || !h->temp || !h->dlnt_dz || !h->hknu)))
    return set_error(M3DRT_EINVAL, w + "missing composition field");
  if (g->kind != M3DRT_LCH_TOP && g->kind != M3DRT_LCH_BOTTOM && g->kind != M3DRT_LCH_UPSTREAM)
    return set_error(M3DRT_EINVAL, w + "unknown ghost kind");
  if (g->kind != M3DRT_LCH_TOP && (!g->pe || !g->dpe_dz || (!hybrid && (!g->temp || !g->dlnt_dz))))
    return set_error(M3DRT_EINVAL, w + "the bottom and upstream ghosts need their Pe, dPe/dz (and T, dlnT/dz)");
  if (d->cs_f < 1 || d->cs_x < 1 || d->cs_y < 1 || d->cs_z < 1 || d->gs_f < 1 || d->gs_x < 1 || d->gs_y < 1)
    return set_error(M3DRT_EINVAL, w + "strides must be >= 1");
  if ((int64_t)d->nf * d->nx * d->ny > (int64_t)0xffffffffLL * 256 - 255)
    return set_error(M3DRT_ELIMIT, w + "more columns times frequencies than one launch covers");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::sweep_lch(ctx, d, b, h, g, hybrid);
}

int m3drt_lc_sweep_hermite(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b,
                           const m3drt_hermite_fields *h, const m3drt_lc_ghost *g) {
  return lc_hermite_entry(ctx, d, b, h, g, false);
}

int m3drt_lc_sweep_hermite2(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b,
                            const m3drt_hermite_fields *h, const m3drt_lc_ghost *g) {
  return lc_hermite_entry(ctx, d, b, h, g, true);
}

int m3drt_sc_adjoint(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_adjoint_bufs *b) {
  if (!ctx || !d || !b) return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: NULL argument");
  if (d->nf < 1 || d->nx < 1 || d->ny < 1 || d->nz < 1) return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: empty grid");
  if (d->nx == 1 && d->ny == 1) return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: 3D planes only (nx*ny > 1)");
  if (d->dir != 1 && d->dir != -1) return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: dir must be +1 or -1");
  if (!(d->ds[0] > 0 && d->ds[1] > 0 && d->ds[2] > 0) || d->vec[2] == 0.0 || (d->vec[2] > 0) != (d->dir > 0))
    return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: bad geometry (ds > 0, dir following vec(3))");
  if (d->prec != M3DRT_F32 || d->solver != M3DRT_LINEAR2)
    return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: FP32 short_char2 only");
  if (d->flags & M3DRT_SPLIT_RATES) return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: no split rates");
  if (!b->chi || !b->S || !b->ghost_chi || !b->ghost_S || !b->ghost_I || !b->chi500 || !b->chi500_up || !b->g500 ||
      !b->I)
    return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: missing input array");
  if (!b->chibar || !b->Sbar) return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: missing chibar or Sbar");
  if (d->cs_f < 1 || d->cs_x < 1 || d->cs_y < 1 || d->cs_z < 1 || d->gs_f < 1 || d->gs_x < 1 || d->gs_y < 1)
    return set_error(M3DRT_EINVAL, "m3drt_sc_adjoint: strides must be >= 1");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::sweep_adjoint(ctx, d, b);
}

static int line_opacity_entry(m3drt_ctx *ctx, const m3drt_linelist *ll, const m3drt_opacity_args *o,
                              bool general) {
  const std::string w = general ? "m3drt_line_opacity: " : "m3drt_line_opacity_equi: ";
  if (!ctx || !ll || !o) return set_error(M3DRT_EINVAL, w + "NULL argument");
  if (ll->nline < 0 || o->ncell < 0 || o->nf < 0) return set_error(M3DRT_EINVAL, w + "negative size");
  // as the reference: nothing to add (the equidistant form needs two nodes for its step)
  if (ll->nline == 0 || o->ncell == 0 || o->nf < (general ? 1 : 2)) return M3DRT_OK;
  if (!ll->llam0 || !ll->lower_ev || !ll->gpref || !ll->grad || !ll->gw1 || !ll->gw2 || !ll->gstk ||
      !ll->unsold || !ll->lev_i || !ll->lev_j)
    return set_error(M3DRT_EINVAL, w + "missing line-list array");
  if ((ll->reach_in || ll->active) && !ll->sel_work)
    return set_error(M3DRT_EINVAL, w + "a list with reaches or active flags needs sel_work (nline + 1 int32)");
  if (!o->llam || !o->vproj || !o->inv_df || !o->vbar2 || !o->e_kT || !o->ne || !o->tothi ||
      !o->stark_pe || !o->lnbase || !o->chi || !o->eta || !o->cont || !o->logref)
    return set_error(M3DRT_EINVAL, w + "missing cell array");
  if (o->npart < 1) return set_error(M3DRT_EINVAL, w + "npart < 1");
  if ((o->flags & M3DRT_OPAC_NLTE) && (!o->dep || o->nlev < 1))
    return set_error(M3DRT_EINVAL, w + "NLTE needs dep and nlev");
  if ((o->flags & M3DRT_OPAC_SCAT) && (!o->one_m_eps || !o->chi_scat))
    return set_error(M3DRT_EINVAL, w + "scattering needs one_m_eps and chi_scat");
  if (o->sf < 1 || o->sc < 1) return set_error(M3DRT_EINVAL, w + "strides must be >= 1");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::line_opacity(ctx, ll, o, general);
}

int m3drt_line_opacity_equi(m3drt_ctx *ctx, const m3drt_linelist *ll, const m3drt_opacity_args *o) {
  return line_opacity_entry(ctx, ll, o, false);
}

int m3drt_line_opacity(m3drt_ctx *ctx, const m3drt_linelist *ll, const m3drt_opacity_args *o) {
  return line_opacity_entry(ctx, ll, o, true);
}

static int check_atom(const m3drt_rate_solve_args *r, const char *who) {
  char msg[160];
  if (r->nk < 1 || r->nrad < 0 || r->nline < 0 || r->nline > r->nrad || r->npair < 0 || r->ncell < 0) {
    snprintf(msg, sizeof msg, "%s: inconsistent sizes", who);
    return set_error(M3DRT_EINVAL, msg);
  }
  if ((r->nrad && (!r->trans_i || !r->trans_j || !r->trans_active || !r->gamma_ij || !r->gamma_ji)) ||
      !r->nstar) {
    snprintf(msg, sizeof msg, "%s: missing transition or population array", who);
    return set_error(M3DRT_EINVAL, msg);
  }
  return M3DRT_OK;
}

int m3drt_rate_solve_f64(m3drt_ctx *ctx, const m3drt_rate_solve_args *r, m3drt_rate_solve_result *res) {
  if (!ctx || !r) return set_error(M3DRT_EINVAL, "m3drt_rate_solve_f64: NULL argument");
  int rc = check_atom(r, "m3drt_rate_solve_f64");
  if (rc) return rc;
  if (!r->g || !r->ev || !r->e_kT || !r->temp || !r->totn || !r->b || (r->nline && !r->Aji) ||
      (r->npair && r->cup && (!r->pair_i || !r->pair_j)))
    return set_error(M3DRT_EINVAL, "m3drt_rate_solve_f64: missing atom or cell array");
  if (res) {
    res->maxerr = 0;
    res->epos = -1;
    res->elvl = -1;
    res->nclamp = 0;
  }
  if (r->ncell == 0) return M3DRT_OK;
  M3DRT_HIP(hipSetDevice(ctx->device));
  rc = m3drt::rate_solve(ctx, r, res);
  if (rc == M3DRT_OK && res && res->maxerr == 0.0) {
    res->epos = -1;
    res->elvl = -1;
  }
  return rc;
}

int m3drt_normalize_rates(m3drt_ctx *ctx, const m3drt_rate_solve_args *r, const float *wphi) {
  if (!ctx || !r) return set_error(M3DRT_EINVAL, "m3drt_normalize_rates: NULL argument");
  int rc = check_atom(r, "m3drt_normalize_rates");
  if (rc) return rc;
  if (r->ncell == 0 || r->nrad == 0) return M3DRT_OK;
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::normalize_rates(ctx, r, wphi);
}

int m3drt_bb_rates(m3drt_ctx *ctx, const m3drt_bb_rates_args *r) {
  if (!ctx || !r) return set_error(M3DRT_EINVAL, "m3drt_bb_rates: NULL argument");
  if (r->nline < 0 || r->nf < 0 || r->ncell < 0 || r->nk < 1)
    return set_error(M3DRT_EINVAL, "m3drt_bb_rates: inconsistent sizes");
  if (r->nline == 0 || r->ncell == 0) return M3DRT_OK;
  if (!r->iblue || !r->ired || !r->lev_i || !r->lev_j || !r->Aij || !r->lower_off || !r->q || !r->wnu ||
      !r->g || !r->b || !r->nstar || !r->inv_df || !r->vproj || !r->e_kT || !r->adamp || !r->Inu ||
      !r->Bnu || !r->gamma_ij || !r->gamma_ji)
    return set_error(M3DRT_EINVAL, "m3drt_bb_rates: missing array");
  const bool precond = r->flags & M3DRT_RATES_PRECOND, largest = r->flags & M3DRT_RATES_LARGEST;
  if (precond && (!r->beta || !r->Bij || !r->pop_ratio))
    return set_error(M3DRT_EINVAL, "m3drt_bb_rates: preconditioning needs beta, Bij, pop_ratio");
  if (precond && !largest && !r->Psi)
    return set_error(M3DRT_EINVAL, "m3drt_bb_rates: preconditioning needs Psi");
  if (precond && largest &&
      (!r->Psi || !r->Irem || !r->escape || !r->chi || !r->split_opac1 || !r->split_opac2 ||
       !r->split_chi1 || !r->split_chi2 || !r->split_id || !r->split_chi_id))
    return set_error(M3DRT_EINVAL,
                     "m3drt_bb_rates: 'largest' with preconditioning needs Psi, Irem, escape, chi and the "
                     "tracked contributions");
  if ((r->flags & M3DRT_RATES_SCAT) && (!r->chi_scat || !r->dJnu))
    return set_error(M3DRT_EINVAL, "m3drt_bb_rates: line scattering needs chi_scat and dJnu");
  if ((r->flags & M3DRT_RATES_FLAGS) && (!r->ntrans || !r->psi_flags_in || !r->psi_flags_out || r->npw < 1))
    return set_error(M3DRT_EINVAL, "m3drt_bb_rates: psi flags need ntrans, psi_flags_in/out and npw");
  if ((r->flags & M3DRT_RATES_OWN_NORM) && !r->wq)
    return set_error(M3DRT_EINVAL, "m3drt_bb_rates: own normalisation needs wq");
  if (r->sf < 1 || r->sc < 1) return set_error(M3DRT_EINVAL, "m3drt_bb_rates: strides must be >= 1");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::bb_rates(ctx, r);
}

int m3drt_bb_opacity(m3drt_ctx *ctx, const m3drt_bb_opacity_args *r) {
  if (!ctx || !r) return set_error(M3DRT_EINVAL, "m3drt_bb_opacity: NULL argument");
  if (r->nline < 0 || r->nf < 0 || r->ncell < 0 || r->nk < 1)
    return set_error(M3DRT_EINVAL, "m3drt_bb_opacity: inconsistent sizes");
  if (r->nline == 0 || r->ncell == 0) return M3DRT_OK;
  if (!r->iblue || !r->ired || !r->lev_i || !r->lev_j || !r->Bij || !r->lower_off || !r->q || !r->nstar ||
      !r->pop_ratio || !r->inv_df || !r->vproj || !r->e_kT || !r->adamp || !r->xhat || !r->ehat)
    return set_error(M3DRT_EINVAL, "m3drt_bb_opacity: missing array");
  if (!(r->flags & M3DRT_BBOP_LTE) && !r->b)
    return set_error(M3DRT_EINVAL, "m3drt_bb_opacity: departures b needed outside LTE");
  if ((r->flags & M3DRT_BBOP_SPLIT) && (!r->split_opac1 || !r->split_opac2 || !r->split_chi1 ||
                                        !r->split_chi2 || !r->split_id || !r->split_chi_id))
    return set_error(M3DRT_EINVAL, "m3drt_bb_opacity: split tracking needs its six arrays");
  if (r->wphi_line && (!r->wphi || !r->wq))
    return set_error(M3DRT_EINVAL, "m3drt_bb_opacity: wphi needs wphi and wq");
  if (r->sf < 1 || r->sc < 1) return set_error(M3DRT_EINVAL, "m3drt_bb_opacity: strides must be >= 1");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::bb_opacity(ctx, r);
}

int m3drt_window_opacity(m3drt_ctx *ctx, const m3drt_bb_opacity_args *r, const m3drt_window_args *w) {
  if (!ctx || !r || !w) return set_error(M3DRT_EINVAL, "m3drt_window_opacity: NULL argument");
  if (r->nline < 0 || r->nf < 0 || r->ncell < 0 || r->nk < 1)
    return set_error(M3DRT_EINVAL, "m3drt_window_opacity: inconsistent sizes");
  if (r->nf == 0 || r->ncell == 0) return M3DRT_OK;
  if (!r->iblue || !r->ired || !r->lev_i || !r->lev_j || !r->Bij || !r->lower_off || !r->q || !r->nstar ||
      !r->pop_ratio || !r->inv_df || !r->vproj || !r->e_kT || !r->adamp || !r->xhat || !r->ehat || !w->chi_cont ||
      !w->Bnu || !w->beta)
    return set_error(M3DRT_EINVAL, "m3drt_window_opacity: missing array");
  if (!(r->flags & M3DRT_BBOP_LTE) && !r->b)
    return set_error(M3DRT_EINVAL, "m3drt_window_opacity: departures b needed outside LTE");
  if ((r->flags & M3DRT_BBOP_SPLIT) && (!r->split_opac1 || !r->split_opac2 || !r->split_chi1 ||
                                        !r->split_chi2 || !r->split_id || !r->split_chi_id))
    return set_error(M3DRT_EINVAL, "m3drt_window_opacity: split tracking needs its six arrays");
  if (r->wphi_line) return set_error(M3DRT_EINVAL, "m3drt_window_opacity: no wphi sums (use m3drt_bb_opacity)");
  if (w->flags & ~M3DRT_UNITS_SUPPRESS_MASERS)
    return set_error(M3DRT_EINVAL, "m3drt_window_opacity: flags other than M3DRT_UNITS_SUPPRESS_MASERS");
  if (r->sf < 1 || r->sc < 1) return set_error(M3DRT_EINVAL, "m3drt_window_opacity: strides must be >= 1");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::window_opacity(ctx, r, w);
}

int m3drt_rotate_array(m3drt_ctx *ctx, const m3drt_rotate_args *r) {
  if (!ctx || !r) return set_error(M3DRT_EINVAL, "m3drt_rotate_array: NULL argument");
  if (r->nf < 1 || r->nx < 1 || r->ny < 1 || r->nz < 1 || r->nk < 1 || r->nz > 65535 || r->nk > 65535)
    return set_error(M3DRT_EINVAL, "m3drt_rotate_array: bad sizes");
  if (!r->in || !r->out) return set_error(M3DRT_EINVAL, "m3drt_rotate_array: missing array");
  if (r->in == r->out) return set_error(M3DRT_EINVAL, "m3drt_rotate_array: out of place only");
  if (!(r->vec[2] != 0.0) || !(r->ds[0] > 0.0) || !(r->ds[1] > 0.0))
    return set_error(M3DRT_EINVAL, "m3drt_rotate_array: vec[2] must be nonzero and ds positive");
  if (r->cs_f < 1 || r->cs_x < 1 || r->cs_y < 1 || r->cs_z < 1 || r->cs_k < 1)
    return set_error(M3DRT_EINVAL, "m3drt_rotate_array: strides must be >= 1");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::rotate_array(ctx, r);
}

int m3drt_bf_rates(m3drt_ctx *ctx, const m3drt_bf_rates_args *r) {
  if (!ctx || !r) return set_error(M3DRT_EINVAL, "m3drt_bf_rates: NULL argument");
  if (r->ncont < 0 || r->nline < 0 || r->nf < 0 || r->ncell < 0 || r->nk < 1)
    return set_error(M3DRT_EINVAL, "m3drt_bf_rates: inconsistent sizes");
  if (r->ncont == 0 || r->ncell == 0) return M3DRT_OK;
  if (!r->iblue || !r->ired || !r->lev_i || !r->lev_j || !r->ev || !r->wnu || !r->alpha || !r->twohnu3_c2 ||
      !r->e_kT || !r->Jnu || !r->beta || !r->gamma_ij || !r->gamma_ji)
    return set_error(M3DRT_EINVAL, "m3drt_bf_rates: missing array");
  if (r->precond && (!r->PsiJ || !r->nstar || !r->b))
    return set_error(M3DRT_EINVAL, "m3drt_bf_rates: preconditioning needs PsiJ, nstar and b");
  if (!r->precond && !r->PsiJ)
    return set_error(M3DRT_EINVAL, "m3drt_bf_rates: PsiJ is read by the sums (pass zeros without preconditioning)");
  if (r->sf < 1 || r->sc < 1) return set_error(M3DRT_EINVAL, "m3drt_bf_rates: strides must be >= 1");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::bf_rates(ctx, r);
}

int m3drt_work_to_transfer(m3drt_ctx *ctx, const m3drt_units_args *r) {
  if (!ctx || !r) return set_error(M3DRT_EINVAL, "m3drt_work_to_transfer: NULL argument");
  if (r->n < 0) return set_error(M3DRT_EINVAL, "m3drt_work_to_transfer: n < 0");
  if (r->n == 0) return M3DRT_OK;
  if (!r->work1 || !r->work2 || !r->beta) return set_error(M3DRT_EINVAL, "m3drt_work_to_transfer: missing array");
  if ((r->flags & M3DRT_UNITS_SPLIT) && (!r->split_opac1 || !r->split_opac2))
    return set_error(M3DRT_EINVAL, "m3drt_work_to_transfer: the split form needs split_opac1/2");
  if ((r->flags & M3DRT_UNITS_SCAT) && (!r->chi_scat || !r->dJnu))
    return set_error(M3DRT_EINVAL, "m3drt_work_to_transfer: line scattering needs chi_scat and dJnu");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::work_to_transfer(ctx, r);
}

int m3drt_background_emission(m3drt_ctx *ctx, const m3drt_background_args *r) {
  if (!ctx || !r) return set_error(M3DRT_EINVAL, "m3drt_background_emission: NULL argument");
  if (r->n < 0) return set_error(M3DRT_EINVAL, "m3drt_background_emission: n < 0");
  if (r->n == 0) return M3DRT_OK;
  if (!r->J || !r->B || !r->chi_cont || !r->sigma || !r->eta)
    return set_error(M3DRT_EINVAL, "m3drt_background_emission: missing array");
  if ((r->flags & M3DRT_BG_SCAT_ALI) && (!r->PsiJ || !r->Jali_old))
    return set_error(M3DRT_EINVAL, "m3drt_background_emission: the ALI step needs PsiJ and Jali_old");
  if ((r->flags & M3DRT_BG_SCAT_ALI) && !(r->scat_op_max < 1.0))
    return set_error(M3DRT_EINVAL, "m3drt_background_emission: scat_op_max must be < 1");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::background_emission(ctx, r);
}

int m3drt_convert_to_work_units(m3drt_ctx *ctx, int64_t n, const float *chi_cont, float *xhat, float *ehat,
                                const float *beta) {
  if (!ctx) return set_error(M3DRT_EINVAL, "m3drt_convert_to_work_units: NULL context");
  if (n < 0) return set_error(M3DRT_EINVAL, "m3drt_convert_to_work_units: n < 0");
  if (n == 0) return M3DRT_OK;
  if (!chi_cont || !xhat || !ehat || !beta)
    return set_error(M3DRT_EINVAL, "m3drt_convert_to_work_units: missing array");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::convert_to_work_units(ctx, n, chi_cont, xhat, ehat, beta);
}

int m3drt_update_eta_mask(m3drt_ctx *ctx, const m3drt_eta_args *r) {
  if (!ctx || !r) return set_error(M3DRT_EINVAL, "m3drt_update_eta_mask: NULL argument");
  if (r->nf < 1 || r->nx < 1 || r->ny < 1 || r->nz < 1 || r->rx < 0 || r->ry < 0)
    return set_error(M3DRT_EINVAL, "m3drt_update_eta_mask: bad sizes");
  if (r->dir != 1 && r->dir != -1) return set_error(M3DRT_EINVAL, "m3drt_update_eta_mask: dir must be +1 or -1");
  if (!r->chi || !r->S || !r->beta || !r->src || !r->mask)
    return set_error(M3DRT_EINVAL, "m3drt_update_eta_mask: missing array");
  if ((r->flags & M3DRT_ETA_GHOST) && (!r->ghost_chi || !r->ghost_S || !r->ghost_src))
    return set_error(M3DRT_EINVAL, "m3drt_update_eta_mask: the ghost plane needs ghost_chi, ghost_S, ghost_src");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::update_eta_mask(ctx, r);
}

static int check_b1d(const m3drt_b1d_batch *b, const char *who) {
  char msg[160];
  if (!b || b->natm < 1) {
    snprintf(msg, sizeof msg, "%s: empty batch", who);
    return set_error(M3DRT_EINVAL, msg);
  }
  if (!b->nd || !b->nf || !b->nmu || !b->nline || !b->d_off || !b->l_off || !b->ld_off || !b->mu_off ||
      !b->f_off || !b->i_off || !b->lam0 || !b->dlam || !b->dtau || !b->hckT || !b->ivth || !b->cont_a ||
      !b->cont_b || !b->mu || !b->wflux || !b->I_top || !b->flux) {
    snprintf(msg, sizeof msg, "%s: missing array", who);
    return set_error(M3DRT_EINVAL, msg);
  }
  if (!(b->lam_ref > 0.0)) {
    snprintf(msg, sizeof msg, "%s: lam_ref must be positive", who);
    return set_error(M3DRT_EINVAL, msg);
  }
  return M3DRT_OK;
}

int m3drt_b1d_synth_plain(m3drt_ctx *ctx, const m3drt_b1d_batch *b, const m3drt_b1d_tile *tile) {
  if (!ctx) return set_error(M3DRT_EINVAL, "m3drt_b1d_synth_plain: NULL context");
  int rc = check_b1d(b, "m3drt_b1d_synth_plain");
  if (rc) return rc;
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::b1d_synth_plain(ctx, b, tile);
}

int m3drt_b1d_plan_create(m3drt_ctx *ctx, const m3drt_b1d_batch *b, const m3drt_b1d_tile *tile,
                          m3drt_b1d_plan **plan) {
  if (!ctx || !plan) return set_error(M3DRT_EINVAL, "m3drt_b1d_plan_create: NULL argument");
  *plan = nullptr;
  int rc = check_b1d(b, "m3drt_b1d_plan_create");
  if (rc) return rc;
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::b1d_plan_create(ctx, b, tile, plan);
}

int m3drt_b1d_plan_run(m3drt_b1d_plan *plan) {
  if (!plan) return set_error(M3DRT_EINVAL, "m3drt_b1d_plan_run: NULL plan");
  return m3drt::b1d_plan_run(plan);
}

int m3drt_b1d_plan_info(const m3drt_b1d_plan *plan, m3drt_b1d_tile *tile, double *times, int32_t *ntimes) {
  if (!plan) return set_error(M3DRT_EINVAL, "m3drt_b1d_plan_info: NULL plan");
  return m3drt::b1d_plan_info(plan, tile, times, ntimes);
}

int m3drt_b1d_plan_destroy(m3drt_b1d_plan *plan) {
  if (!plan) return M3DRT_OK;
  return m3drt::b1d_plan_destroy(plan);
}

int m3drt_linear_weights(m3drt_ctx *ctx, int64_t n, const float *t, float *E, float *W, float *Psi, float *Esc) {
  if (!ctx) return set_error(M3DRT_EINVAL, "m3drt_linear_weights: ctx is NULL");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::step_weights(ctx, 0, n, t, E, W, Psi, Esc);
}

int m3drt_parab_moments(m3drt_ctx *ctx, int64_t n, const float *t, float *E, float *U0, float *U1, float *q) {
  if (!ctx) return set_error(M3DRT_EINVAL, "m3drt_parab_moments: ctx is NULL");
  M3DRT_HIP(hipSetDevice(ctx->device));
  return m3drt::step_weights(ctx, 1, n, t, E, U0, U1, q);
}

}  // extern "C"

