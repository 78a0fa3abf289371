// The line-list opacity of a composition (composition_type%add_linelist_opacities,
// HOP/composition_mod.f90:1663-1954, without its hydrogen Stark branch): the
// tolerance reference logref = ln(tol/voigt_peak) + ln chi from the working
// opacity before any line (:1766-1772), formed by the summation kernel at the
// nodes it reads (the working opacity is only added to at the end); per species
// and isotope the population
// base and Doppler width of add_particle_opacity (HOP/linelist_mod.f90:970-1070)
// and the Voigt summation (line_opacity, the kernels of m3drt_line_opacity[_equi])
// into the line increments; the increments added to the working xhat/ehat
// (:1862-1873); the LTE scattering opacity taken to physical units (:1879-1890).
//
// Host orchestration over device kernels: the species loop is the reference's,
// the per-cell and per-(frequency, cell) passes are elementwise kernels, lanes over
// cells.
#include "m3drt_internal.hpp"
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <string>

namespace m3drt {
namespace {

// HOP/const_mod.f90:40-45
constexpr double kCC = 2.99792458e10, kUU = 1.6605655e-24, kBK = 1.380662e-16, kMH = 1.6726e-24;
constexpr double kPI = 3.14159265359;

// inverse_doppler_factor and compute_vbar2 (HOP/ion_mod.f90:198-221) and the
// population base of add_particle_opacity (HOP/linelist_mod.f90:1022-1030), in the
// reference's mixed precision
__global__ void __launch_bounds__(256) particle_base_kernel(int64_t ncell, const float *__restrict__ temp,
                                                            const float *__restrict__ vmic,
                                                            const double *__restrict__ n, double scale,
                                                            const double *__restrict__ part, double awgt,
                                                            double frac, float *__restrict__ idf,
                                                            float *__restrict__ vb2, float *__restrict__ lnbase) {
#pragma clang fp contract(off)
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const float cdop = (float)(2.0 * kBK / kUU);
  const float cvb = (float)(8.0 * kBK / (kPI * kUU)), inv_mh = (float)(kUU / kMH);
  const float T = temp[c];
  float dop = (float)((double)(cdop * T) / awgt);
  const float vm = vmic[c];
  if (vm > 0) {
    const float v = vm * 1e5f;
    dop = dop + v * v;
  }
  const float id = (float)(kCC / (double)sqrtf(dop));
  idf[c] = id;
  vb2[c] = (float)((double)(cvb * T) * ((double)inv_mh + 1.0 / awgt));
  float lb = (float)log(fmax(n[c], DBL_MIN));
  if (scale != 1.0) lb = lb + (float)log(scale);
  lnbase[c] = (float)((((double)lb - part[c]) + (double)(float)log((double)id)) + (double)(float)log(frac));
}

__global__ void __launch_bounds__(256) zero_kernel(int64_t n, float *__restrict__ p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

__global__ void __launch_bounds__(256) add_lines_kernel(int nf, int64_t ncell, int64_t sf, int64_t sc,
                                                        const float *__restrict__ chi_l,
                                                        const float *__restrict__ eta_l, float *__restrict__ chi,
                                                        float *__restrict__ eta, float *__restrict__ chi_out,
                                                        float *__restrict__ eta_out) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (c >= ncell) return;
  const int64_t o = f * sf + c * sc, l = (int64_t)f * ncell + c;
  const float a = chi_l[l], b = eta_l[l];
  eta[o] = eta[o] + b;
  chi[o] = chi[o] + a;
  if (chi_out) chi_out[o] = a;
  if (eta_out) eta_out[o] = b;
}

__global__ void __launch_bounds__(256) scat_physical_kernel(int nf, int64_t ncell, int64_t sf, int64_t sc,
                                                            const float *__restrict__ beta,
                                                            float *__restrict__ chi_scat) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (c >= ncell) return;
  const int64_t o = f * sf + c * sc;
  chi_scat[o] = chi_scat[o] * (1.0f - beta[o]);
}

}  // namespace
}  // namespace m3drt

using m3drt::set_error;

extern "C" int m3drt_add_linelist_opacities(m3drt_ctx *ctx, const m3drt_linelist_opacity_args *a) {
  const std::string w = "m3drt_add_linelist_opacities: ";
  if (!ctx || !a) return set_error(M3DRT_EINVAL, w + "NULL argument");
  if (a->nf < 0 || a->ncell < 0 || a->nsp < 0 || a->nf > 65535) return set_error(M3DRT_EINVAL, w + "bad sizes");
  if (a->nf == 0 || a->ncell == 0) return M3DRT_OK;
  if (!a->llam || !a->temp || !a->vmic || !a->vproj || !a->e_kT || !a->ne || !a->tothi || !a->stark_pe ||
      !a->all_chi || !a->all_eta || (a->nsp && !a->sp))
    return set_error(M3DRT_EINVAL, w + "missing array");
  if ((a->chi_scat != nullptr) != (a->one_m_eps != nullptr))
    return set_error(M3DRT_EINVAL, w + "one_m_eps and chi_scat come together");
  for (int s = 0; s < a->nsp; ++s) {
    const m3drt_species &sp = a->sp[s];
    if (sp.niso < 0 || (sp.niso && !sp.iso) || !sp.n || !sp.part)
      return set_error(M3DRT_EINVAL, w + "species " + std::to_string(s) + ": missing isotope list, n or part");
    if (sp.dep && sp.nlev < 1) return set_error(M3DRT_EINVAL, w + "species with departures needs nlev >= 1");
    if (sp.scat_group < 0 || sp.scat_group > 2) return set_error(M3DRT_EINVAL, w + "scat_group is 0, 1 or 2");
    for (int k = 0; k < sp.niso; ++k)
      if ((sp.iso[k].lines.reach_in || sp.iso[k].lines.active) && !sp.iso[k].lines.sel_work)
        return set_error(M3DRT_EINVAL, w + "a list with reaches or active flags needs sel_work (nline + 1 int32)");
  }
  M3DRT_HIP(hipSetDevice(ctx->device));
  const int nf = a->nf;
  const int64_t nc = a->ncell, n2 = (int64_t)nf * nc;
  const int64_t sf = nc, sc = 1;
  // One list of at most 4 lines and nothing else (no other list, no scattering split,
  // no copy of the increments asked for): the summation kernel deposits each cell's
  // lines once, so it adds their sums straight to the working opacity, with the
  // increments' rounding, and the increments with their zeroing and adding passes
  // fall away.
  int nlists = 0, nl_direct = 0;
  bool scat_direct = false;
  for (int s = 0; s < a->nsp; ++s)
    for (int k = 0; k < a->sp[s].niso; ++k)
      if (a->sp[s].iso[k].lines.nline > 0) {
        ++nlists;
        nl_direct = a->sp[s].iso[k].lines.nline;
        scat_direct = !a->sp[s].dep && a->sp[s].scat_group && a->chi_scat;
      }
  const bool direct = nlists == 1 && nl_direct <= 4 && !scat_direct && !a->chi_lines && !a->eta_lines &&
                      !getenv("M3DRT_NO_DIRECT_LINES");
  // scratch: chi_l, eta_l (f, cell) unless direct, and idf, vb2, lnbase (cell).
  // logref = ln(tol/voigt_peak) + ln all_chi is formed by the opacity kernel at the
  // nodes it needs: all_chi is not otherwise touched until every list has been summed.
  const int64_t nl2 = direct ? 0 : n2;
  const size_t bytes = sizeof(float) * (size_t)(2 * nl2 + 3 * nc);
  int rc = m3drt::ensure_scratch(ctx, bytes);
  if (rc) return rc;
  float *chi_l = static_cast<float *>(ctx->scratch);
  float *eta_l = chi_l + nl2, *idf = eta_l + nl2, *vb2 = idf + nc, *lnb = vb2 + nc;
  hipStream_t st = ctx->stream;
  const dim3 g2((unsigned)((nc + 255) / 256), (unsigned)nf);
  const unsigned g1 = (unsigned)((nc + 255) / 256);
  if (direct) {
    chi_l = a->all_chi;
    eta_l = a->all_eta;
    if (getenv("M3DRT_VERBOSE"))
      fprintf(stderr, "[m3drt] add_linelist_opacities: one list of %d lines, deposited directly\n", nl_direct);
  } else {
    hipLaunchKernelGGL(m3drt::zero_kernel, dim3((unsigned)std::min<int64_t>((2 * n2 + 255) / 256, 1 << 20)),
                       dim3(256), 0, st, 2 * n2, chi_l);
    M3DRT_HIP(hipGetLastError());
  }
  for (int s = 0; s < a->nsp; ++s) {
    const m3drt_species &sp = a->sp[s];
    for (int k = 0; k < sp.niso; ++k) {
      const m3drt_isotope &iso = sp.iso[k];
      if (iso.lines.nline <= 0) continue;
      hipLaunchKernelGGL(m3drt::particle_base_kernel, dim3(g1), dim3(256), 0, st, nc, a->temp, a->vmic, sp.n,
                         sp.scale, sp.part, iso.awgt, iso.frac, idf, vb2, lnb);
      M3DRT_HIP(hipGetLastError());
      m3drt_opacity_args o = {};
      o.nf = nf;
      o.ncell = (int32_t)nc;
      o.npart = 1;
      o.nlev = sp.dep ? sp.nlev : 0;
      o.flags = (sp.dep ? M3DRT_OPAC_NLTE : 0) | (!sp.dep && sp.scat_group && a->chi_scat ? M3DRT_OPAC_SCAT : 0);
      o.tolerance = a->tolerance;
      o.line_reject_frac = a->line_reject_frac;
      o.sf = nc;  // every (f, cell) array of the call: cells fastest
      o.sc = 1;
      o.llam = a->llam;
      o.vproj = a->vproj;
      o.inv_df = idf;
      o.vbar2 = vb2;
      o.e_kT = a->e_kT;
      o.ne = a->ne;
      o.tothi = a->tothi;
      o.stark_pe = a->stark_pe;
      o.lnbase = lnb;
      o.dep = sp.dep;
      o.chi = chi_l;
      o.eta = eta_l;
      o.cont = a->all_chi;
      o.logref = nullptr;
      o.one_m_eps = (o.flags & M3DRT_OPAC_SCAT) ? a->one_m_eps + (int64_t)(sp.scat_group - 1) * n2 : nullptr;
      o.chi_scat = (o.flags & M3DRT_OPAC_SCAT) ? a->chi_scat : nullptr;
      rc = m3drt::line_opacity(ctx, &iso.lines, &o, a->equidistant == 0, true, direct);
      if (rc) return rc;
    }
  }
  if (!direct) {
    hipLaunchKernelGGL(m3drt::add_lines_kernel, g2, dim3(256), 0, st, nf, nc, sf, sc, chi_l, eta_l, a->all_chi,
                       a->all_eta, a->chi_lines, a->eta_lines);
    M3DRT_HIP(hipGetLastError());
  }
  if (a->chi_scat && a->beta) {
    hipLaunchKernelGGL(m3drt::scat_physical_kernel, g2, dim3(256), 0, st, nf, nc, sf, sc, a->beta, a->chi_scat);
    M3DRT_HIP(hipGetLastError());
  }
  return M3DRT_OK;
}
