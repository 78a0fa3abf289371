// Per-cell statistical equilibrium in FP64 for gfx950: the rate-matrix assembly of
// setup_transition_matrix (SRC/atom_mod.f90:740-851, line_rate_coefficients
// 862-884), the no-pivot LU of lineq_no_pivot (HOP/op_math_mod.f90:420-449) and
// the update of update_populations (SRC/atom_mod.f90:439-732): limiter,
// under-relaxation, particle conservation, population floor and the convergence
// measure.  Also the rate normalisation of normalize_profiles (377-431).
//
// Mapping.  One thread per cell.  A workgroup of NT cells keeps every cell's
// nk x nk matrix, right-hand side and b*n* vector in LDS, interleaved cell-fastest
// (element e of cell t at e*NT + t), so the lanes of a wavefront touch
// consecutive doubles and the in-place elimination runs without bank conflicts.
// The atom (levels, transitions, pairs) is uniform and read through the scalar
// cache.  The largest change and the per-level changes are reduced first per
// workgroup and then by one workgroup, both keeping the first cell in the
// reference's loop order on ties, so the result is deterministic.
#include "m3drt_internal.hpp"
#include <cmath>
#include <algorithm>
#include <vector>

namespace m3drt {

namespace {

struct SeArgs {
  int nk, nrad, nline, npair, ncell, nt;
  const int32_t *ti, *tj, *active, *pi, *pj;
  const double *g, *ev, *Aji;
  const double *e_kT, *temp, *totn, *nstar, *cup;
  double *b;
  const float *gij, *gji;
  double max_update, min_pop_frac, conv_pop_floor, sor_w, cswitch, cut_x, bf_cut_rate, tmax_update;
  double *blk_err;       // (nblk) signed weighted change of the block's winner
  int32_t *blk_cell;     // (nblk) its cell, -1 for none
  int32_t *blk_lvl;      // (nblk) its level
  double *blk_lerr;      // (nblk, nk) per-level signed change
  int32_t *blk_lcell;    // (nblk, nk) its cell
  unsigned long long *nclamp;
  int32_t *status;       // smallest 1-based cell without a finite solution
};

// keep the larger |value|; on ties the smaller cell (earlier in the reference's loop)
__device__ __forceinline__ bool better(double va, int ca, double vb, int cb) {
  if (cb < 0) return ca >= 0;
  if (ca < 0) return false;
  const double a = fabs(va), b = fabs(vb);
  return a > b || (a == b && ca < cb);
}

__global__ void __launch_bounds__(64) se_update_kernel(const SeArgs a) {
  extern __shared__ double sm[];
  const int nk = a.nk, NT = a.nt;
  const int t = threadIdx.x;
  const int c = blockIdx.x * NT + t;
  double *W = sm;                        // nk*nk*NT
  double *E = W + (size_t)nk * nk * NT;  // nk*NT
  double *BN = E + (size_t)nk * NT;      // nk*NT (b*n*, then b_old)
  double *RV = BN + (size_t)nk * NT;     // 2*NT reduction values
  int *RC = reinterpret_cast<int *>(RV + 2 * NT);  // 2*NT reduction cells
#define Wm(i, j) W[((size_t)(i) * nk + (j)) * NT + t]
#define Ev(i) E[(size_t)(i) * NT + t]
#define Bv(i) BN[(size_t)(i) * NT + t]
  const bool live = c < a.ncell && a.totn[c] > 0 && a.temp[c] < a.tmax_update;
  double my_err = 0.0;
  int my_lvl = -1;
  unsigned long long clamps = 0;
  if (live) {
    const double ekt = a.e_kT[c];
    auto NS = [&](int i) { return a.nstar[(size_t)i * a.ncell + c]; };
    for (int i = 0; i < nk; ++i) {
      Ev(i) = 0.0;
      for (int j = 0; j < nk; ++j) Wm(i, j) = 0.0;
    }
    for (int kr = 0; kr < a.nrad; ++kr) {
      if (!a.active[kr]) continue;
      const int i = a.ti[kr] - 1, j = a.tj[kr] - 1;
      if ((a.ev[j] - a.ev[i]) * ekt > a.cut_x) {
        // a detailed-balance pair for a transition beyond cut_x kT
        const double cut = kr < a.nline ? a.Aji[kr] : a.bf_cut_rate;
        Wm(i, j) = Wm(i, j) + cut;
        Wm(j, i) = Wm(j, i) + cut * NS(j) / NS(i);
        continue;
      }
      double r_up, r_down;
      if (kr < a.nline) {
        const double th = a.gij[(size_t)kr * a.ncell + c], dev = a.gji[(size_t)kr * a.ncell + c];
        const double g_ij = a.g[i] / a.g[j];
        if (NS(i) > 0 && NS(j) > 0) {
          const double beta = NS(j) * g_ij / NS(i);
          r_up = th * (beta / (1.0 - beta)) + dev;
          r_down = g_ij * (th / (1.0 - beta) + dev);
        } else {
          r_up = 0.0;
          r_down = 0.0;
        }
      } else {
        r_up = a.gij[(size_t)kr * a.ncell + c];
        r_down = a.gji[(size_t)kr * a.ncell + c];
      }
      Wm(j, i) = Wm(j, i) + r_up;
      Wm(i, j) = Wm(i, j) + r_down;
    }
    if (a.cswitch > 0 && a.cup) {
      for (int q = 0; q < a.npair; ++q) {
        const int i = a.pi[q] - 1, j = a.pj[q] - 1;
        const double cu = a.cup[(size_t)q * a.ncell + c];
        if (cu == 0) continue;
        Wm(j, i) = Wm(j, i) + cu * a.cswitch;
        Wm(i, j) = Wm(i, j) + cu * (NS(i) / NS(j)) * a.cswitch;
      }
    }
    for (int i = 0; i < nk; ++i) {
      double s = 0.0;
      for (int k = 0; k < nk; ++k) s += Wm(k, i);
      Wm(i, i) = -s;
    }
    int pos = 0;
    for (int i = 0; i < nk; ++i) {
      Bv(i) = a.b[(size_t)i * a.ncell + c] * NS(i);
      if (Bv(i) > Bv(pos)) pos = i;
    }
    Ev(pos) = a.totn[c];
    for (int i = 0; i < nk; ++i)
      for (int j = 0; j < nk; ++j) Wm(i, j) = Wm(i, j) * Bv(j);
    for (int j = 0; j < nk; ++j) Wm(pos, j) = Bv(j);

    // lineq_no_pivot: Doolittle without pivoting, zero upper entries skipped
    for (int k = 0; k < nk - 1; ++k) {
      const double akk = Wm(k, k);
      for (int i = k + 1; i < nk; ++i) Wm(i, k) = Wm(i, k) / akk;
      for (int j = k + 1; j < nk; ++j) {
        const double akj = Wm(k, j);
        if (akj != 0)
          for (int i = k + 1; i < nk; ++i) Wm(i, j) = Wm(i, j) - Wm(i, k) * akj;
      }
    }
    for (int i = 0; i < nk; ++i)
      for (int j = 0; j < i; ++j) Ev(i) = Ev(i) - Wm(i, j) * Ev(j);
    for (int i = nk - 1; i >= 0; --i) {
      for (int j = i + 1; j < nk; ++j) Ev(i) = Ev(i) - Wm(i, j) * Ev(j);
      Ev(i) = Ev(i) / Wm(i, i);
    }
    bool finite = true;
    for (int i = 0; i < nk; ++i) finite = finite && (fabs(Ev(i)) <= 1.79769313486231570e308);
    if (!finite) {
      atomicMin(a.status, c + 1);
    } else {
      const double eup = a.max_update > 1.0 ? a.max_update : 1.0, elo = 1.0 / eup;
      double sum = 0.0;
      for (int i = 0; i < nk; ++i) {
        double e = Ev(i);
        if (e < elo || e > eup) clamps++;
        if (e < elo)
          e = elo * exp(fmax(e, 0.0) / elo - 1.0);
        else if (e > eup)
          e = eup * (2.0 - eup / e);
        if (a.sor_w != 1.0) e = pow(e, a.sor_w);
        const size_t o = (size_t)i * a.ncell + c;
        const double bo = a.b[o];
        Bv(i) = bo;                       // b_old
        const double bn = bo * e;
        Ev(i) = bn;                       // b, before normalisation
        sum += bn * a.nstar[o];
      }
      const double norm = a.totn[c] / sum;
      double wmax = -1.0;
      for (int i = 0; i < nk; ++i) {
        const size_t o = (size_t)i * a.ncell + c;
        double bn = Ev(i) * norm;
        if (a.nstar[o] > 0) bn = fmax(bn, fmin(1.0, a.min_pop_frac * a.totn[c] / a.nstar[o]));
        a.b[o] = bn;
        const double e = Bv(i) > 0 ? bn / Bv(i) - 1.0 : 0.0;
        Ev(i) = e;                        // the change, for the per-level reduction
        double ew = fabs(e);
        if (a.conv_pop_floor > 0) ew = fabs(e) * fmin(1.0, bn * a.nstar[o] / (a.conv_pop_floor * a.totn[c]));
        if (ew > wmax) {
          wmax = ew;
          my_lvl = i;
        }
      }
      my_err = copysign(wmax, Ev(my_lvl));
    }
  }
  if (clamps) atomicAdd(a.nclamp, clamps);

  // workgroup reduction: the largest weighted change, then every level's change
  const int cell = (live && my_lvl >= 0) ? c : -1;
  const int nblk = blockIdx.x;
  __syncthreads();
  RV[t] = my_err;
  RC[t] = cell;
  RC[NT + t] = my_lvl;
  __syncthreads();
  if (t == 0) {
    double bv = 0.0;
    int bc = -1, bl = -1;
    for (int s = 0; s < NT; ++s)
      if (better(RV[s], RC[s], bv, bc)) {
        bv = RV[s];
        bc = RC[s];
        bl = RC[NT + s];
      }
    a.blk_err[nblk] = bv;
    a.blk_cell[nblk] = bc;
    a.blk_lvl[nblk] = bl;
  }
  for (int i = 0; i < nk; ++i) {
    __syncthreads();
    RV[t] = cell >= 0 ? Ev(i) : 0.0;
    RC[t] = cell;
    __syncthreads();
    if (t == 0) {
      double bv = 0.0;
      int bc = -1;
      for (int s = 0; s < NT; ++s)
        if (better(RV[s], RC[s], bv, bc)) {
          bv = RV[s];
          bc = RC[s];
        }
      a.blk_lerr[(size_t)nblk * nk + i] = bv;
      a.blk_lcell[(size_t)nblk * nk + i] = bc;
    }
  }
#undef Wm
#undef Ev
#undef Bv
}

// the blocks' winners, reduced by one workgroup; out[0] = maxerr, out[1..nk] = lvlerr,
// iout[0] = cell, iout[1] = level
__global__ void __launch_bounds__(256) se_final_kernel(int nblk, int nk, const double *blk_err,
                                                       const int32_t *blk_cell, const int32_t *blk_lvl,
                                                       const double *blk_lerr, const int32_t *blk_lcell,
                                                       double *out, int32_t *iout) {
  __shared__ double rv[256];
  __shared__ int rc[256], rl[256];
  const int t = threadIdx.x;
  for (int q = -1; q < nk; ++q) {
    double bv = 0.0;
    int bc = -1, bl = -1;
    for (int s = t; s < nblk; s += blockDim.x) {
      const double v = q < 0 ? blk_err[s] : blk_lerr[(size_t)s * nk + q];
      const int cc = q < 0 ? blk_cell[s] : blk_lcell[(size_t)s * nk + q];
      if (better(v, cc, bv, bc)) {
        bv = v;
        bc = cc;
        bl = q < 0 ? blk_lvl[s] : q;
      }
    }
    rv[t] = bv;
    rc[t] = bc;
    rl[t] = bl;
    __syncthreads();
    if (t == 0) {
      for (int s = 1; s < (int)blockDim.x; ++s)
        if (better(rv[s], rc[s], bv, bc)) {
          bv = rv[s];
          bc = rc[s];
          bl = rl[s];
        }
      if (q < 0) {
        out[0] = bc >= 0 ? bv : 0.0;
        iout[0] = bc;
        iout[1] = bl;
      } else {
        out[1 + q] = bc >= 0 ? bv : 0.0;
      }
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) normalize_kernel(int nrad, int nline, int ncell, const int32_t *active,
                                                        const int32_t *ti, const int32_t *tj,
                                                        const double *nstar, const float *wphi,
                                                        float *gij, float *gji) {
  const int64_t n = (int64_t)nrad * ncell;
  for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < n;
       id += (int64_t)gridDim.x * blockDim.x) {
    const int kr = (int)(id / ncell), c = (int)(id % ncell);
    if (!active[kr]) continue;
    if (kr < nline) {
      if (wphi) {
        gij[id] = gij[id] * wphi[id];
        gji[id] = gji[id] * wphi[id];
      }
    } else {
      const int i = ti[kr] - 1, j = tj[kr] - 1;
      gji[id] = (float)((double)gji[id] * (nstar[(int64_t)i * ncell + c] / nstar[(int64_t)j * ncell + c]));
    }
  }
}

}  // namespace

int rate_solve(m3drt_ctx *ctx, const m3drt_rate_solve_args *r, m3drt_rate_solve_result *res) {
  const int nk = r->nk;
  const size_t per_cell = ((size_t)nk * nk + 2 * (size_t)nk + 2) * sizeof(double) + 2 * sizeof(int);
  int nt = (int)std::min<size_t>(64, (size_t)ctx->lds_max / per_cell);
  if (nt >= 8) nt -= nt % 8;
  if (nt < 1) return set_error(M3DRT_ELIMIT, "m3drt_rate_solve_f64: nk too large for the LDS");
  const int nblk = (r->ncell + nt - 1) / nt;
  const size_t lds = per_cell * nt;
  // scratch: per-block winners, then the final values
  const size_t bytes = (size_t)nblk * (sizeof(double) + 2 * sizeof(int32_t)) +
                       (size_t)nblk * nk * (sizeof(double) + sizeof(int32_t)) +
                       (nk + 1) * sizeof(double) + 4 * sizeof(int32_t) + sizeof(unsigned long long) + 64;
  int rc = ensure_scratch(ctx, bytes);
  if (rc) return rc;
  char *s = static_cast<char *>(ctx->scratch);
  double *blk_err = reinterpret_cast<double *>(s);
  double *blk_lerr = blk_err + nblk;
  double *out = blk_lerr + (size_t)nblk * nk;
  unsigned long long *nclamp = reinterpret_cast<unsigned long long *>(out + nk + 1);
  int32_t *blk_cell = reinterpret_cast<int32_t *>(nclamp + 1);
  int32_t *blk_lvl = blk_cell + nblk;
  int32_t *blk_lcell = blk_lvl + nblk;
  int32_t *iout = blk_lcell + (size_t)nblk * nk;
  int32_t *status = iout + 2;
  const int32_t big = 0x7fffffff;
  M3DRT_HIP(hipMemsetAsync(nclamp, 0, sizeof(unsigned long long), ctx->stream));
  M3DRT_HIP(hipMemcpyAsync(status, &big, sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));

  SeArgs a;
  a.nk = nk; a.nrad = r->nrad; a.nline = r->nline; a.npair = r->npair; a.ncell = r->ncell; a.nt = nt;
  a.ti = r->trans_i; a.tj = r->trans_j; a.active = r->trans_active; a.pi = r->pair_i; a.pj = r->pair_j;
  a.g = r->g; a.ev = r->ev; a.Aji = r->Aji;
  a.e_kT = r->e_kT; a.temp = r->temp; a.totn = r->totn; a.nstar = r->nstar; a.cup = r->cup;
  a.b = r->b; a.gij = r->gamma_ij; a.gji = r->gamma_ji;
  a.max_update = r->max_update; a.min_pop_frac = r->min_pop_frac; a.conv_pop_floor = r->conv_pop_floor;
  a.sor_w = r->sor_w; a.cswitch = r->cswitch; a.cut_x = r->cut_x; a.bf_cut_rate = r->bf_cut_rate;
  a.tmax_update = r->tmax_update;
  a.blk_err = blk_err; a.blk_cell = blk_cell; a.blk_lvl = blk_lvl; a.blk_lerr = blk_lerr;
  a.blk_lcell = blk_lcell; a.nclamp = nclamp; a.status = status;
  M3DRT_HIP(hipFuncSetAttribute((const void *)se_update_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  hipLaunchKernelGGL(se_update_kernel, dim3(nblk), dim3(nt), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  hipLaunchKernelGGL(se_final_kernel, dim3(1), dim3(256), 0, ctx->stream, nblk, nk, blk_err, blk_cell, blk_lvl,
                     blk_lerr, blk_lcell, out, iout);
  M3DRT_HIP(hipGetLastError());
  if (!res) return M3DRT_OK;
  std::vector<double> hout(nk + 1);
  int32_t hi[3];
  unsigned long long hcl = 0;
  M3DRT_HIP(hipMemcpyAsync(hout.data(), out, (nk + 1) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  M3DRT_HIP(hipMemcpyAsync(hi, iout, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  M3DRT_HIP(hipMemcpyAsync(&hcl, nclamp, sizeof(hcl), hipMemcpyDeviceToHost, ctx->stream));
  M3DRT_HIP(hipStreamSynchronize(ctx->stream));
  res->maxerr = hout[0];
  res->epos = hi[0];
  res->elvl = hi[1];
  res->nclamp = (int64_t)hcl;
  if (res->lvlerr)
    for (int i = 0; i < nk; ++i) res->lvlerr[i] = hout[1 + i];
  if (hi[2] != big) {
    char msg[160];
    snprintf(msg, sizeof msg,
             "m3drt_rate_solve_f64: statistical equilibrium has no finite solution at cell %d "
             "(0-based, x fastest); check the rate coefficients for overflow", hi[2] - 1);
    res->epos = hi[2] - 1;
    return set_error(M3DRT_EINVAL, msg);
  }
  return M3DRT_OK;
}

int normalize_rates(m3drt_ctx *ctx, const m3drt_rate_solve_args *r, const float *wphi) {
  const int64_t n = (int64_t)r->nrad * r->ncell;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(normalize_kernel, dim3(blocks), dim3(256), 0, ctx->stream, r->nrad, r->nline, r->ncell,
                     r->trans_active, r->trans_i, r->trans_j, r->nstar, wphi, r->gamma_ij, r->gamma_ji);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

}  // namespace m3drt
