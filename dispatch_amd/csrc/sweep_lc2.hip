// Column sweep with the step of long_char2 (SRC/rt_solvers_mod.f90:527-663), fused
// with integrate_intensity (SRC/rt_patch_mod.f90:3399-3431): 1D atmospheres and
// column-by-column runs, where the upwind point of a cell is the previous cell of
// its own column.  One thread per (frequency, column) walks the column; lanes run
// along the unit-stride axis of the layout (frequency in the reference layout) so
// every access of a wavefront is contiguous.  The step lengths in tau500 depend on
// the cell only, so a first kernel forms them once per direction in FP64.
#include "m3drt_internal.hpp"
#include <cmath>
#include <algorithm>

namespace m3drt {

template <typename T>
struct Lc2Args {
  const T *chi, *S, *gchi, *gS, *gI;
  const T *prep;  // per cell (dt500, rc, rcu) for planes 0..nz-1
  const int32_t *eta_row;
  const uint8_t *eta_mask;
  T *J, *PsiJ, *Iout, *Psiout, *Irem, *esc, *exitI;
  int64_t cs_f, cs_x, cs_y, cs_z, gs_f, gs_x, gs_y, eta_ld, eta_nrow;
  int nf, nx, ny, nz, dir, split, f_fast;
  T mdtau, imax, wj;
};

// dt500 of each step of a column, from chi500 of the cell and of the previous cell
// of the column (the ghost plane's g500 on the first), SRC/rt_solvers_mod.f90:579-593.
template <typename T>
__global__ void __launch_bounds__(256) lc2_prep(const float *__restrict__ chi500,
                                                const float *__restrict__ g500, T *__restrict__ prep,
                                                int nxy, int nz, int dir, double dsup) {
  const int64_t n = (int64_t)nxy * nz;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i / nxy), c = (int)(i % nxy);
    const int zup = z - dir;
    const double cc = (double)chi500[i];
    const double cu = (zup < 0 || zup >= nz) ? (double)g500[c] : (double)chi500[(int64_t)zup * nxy + c];
    const double r = cu / cc;
    double d;
    if (fabs(r - 1.0) < 1e-3)
      d = dsup * cc * (1.0 + (r - 1.0) * (0.5 - (r - 1.0) / 12.0));
    else
      d = dsup * cc * (r - 1.0) / log(r);
    prep[3 * i] = (T)d;
    prep[3 * i + 1] = (T)(1.0 / cc);
    prep[3 * i + 2] = (T)(1.0 / cu);
  }
}

template <typename T, bool ETA>
__global__ void __launch_bounds__(256) lc2_sweep(const Lc2Args<T> a) {
  using namespace dev;
  const int64_t nxy = (int64_t)a.nx * a.ny;
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= nxy * a.nf) return;
  int f, col;
  if (a.f_fast) {
    f = (int)(id % a.nf);
    col = (int)(id / a.nf);
  } else {
    col = (int)(id % nxy);
    f = (int)(id / nxy);
  }
  const int x = col % a.nx, y = col / a.nx;
  const int64_t o = f * a.cs_f + x * a.cs_x + y * a.cs_y;
  const int64_t go = f * a.gs_f + x * a.gs_x + y * a.gs_y;
  T chi_up = a.gchi[go], S_up = a.gS[go], I_up = a.gI[go];
  for (int step = 0; step < a.nz; ++step) {
    const int z = a.dir > 0 ? step : a.nz - 1 - step;
    const int64_t oz = o + (int64_t)z * a.cs_z;
    const T *pp = a.prep + 3 * ((int64_t)z * nxy + col);
    const T dt500 = pp[0], rc = pp[1], rcu = pp[2];
    const T chiv = a.chi[oz], ss = a.S[oz];
    const T xxup = chi_up * rcu;
    const T xx = chiv * rc;
    const T dt = fmax((T)0.5 * (xxup + xx) * dt500, a.mdtau);
    T E, W, Ps, Es;
    lin_weights(dt, E, W, Ps, Es);
    const T sgnchi = copysign(fmax(fabs(chiv), (T)1.17549435e-38f), chiv);
    bool special = false;
    if (ETA) {
      const int row = a.eta_row[col + nxy * z];
      special = row > 0 && row <= a.eta_nrow && a.eta_mask[f + a.eta_ld * (int64_t)(row - 1)] != 0;
    }
    const T eup = xxup * S_up, ee = xx * ss;
    const T rxb = dt500 * rcp(dt != (T)0 ? dt : (T)1);
    const T Ir = special ? I_up * E + W * eup * rxb : I_up * E + W * S_up;
    T Inu = special ? Ir + Ps * ee * rxb : Ir + Ps * ss;
    bool off = fmin(fmin(dt, chiv), special ? xxup : (T)1) <= (T)0;
    if (a.split) {
      off = off || (a.imax - fabs(Inu)) <= (T)0;
      Es = special ? Es + (T)0.5 * Ps * (xxup - xx) * rxb : Es;
    }
    const T psiout = off ? (T)0 : (special ? Ps * rxb * rc : Ps * rcp(sgnchi));
    Inu = fmax(-a.imax, fmin(a.imax, Inu));
    if (a.J) a.J[oz] = a.J[oz] + a.wj * Inu;
    if (a.PsiJ) a.PsiJ[oz] = a.PsiJ[oz] + a.wj * psiout;
    if (a.Iout) a.Iout[oz] = Inu;
    if (a.Psiout) a.Psiout[oz] = psiout;
    if (a.split) {
      if (a.Irem) a.Irem[oz] = off ? Inu : Ir;
      if (a.esc) a.esc[oz] = off ? (T)1 : Es;
    }
    chi_up = chiv;
    S_up = ss;
    I_up = Inu;
  }
  if (a.exitI) a.exitI[go] = I_up;
}

template <typename T>
static int run_lc2(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  const int64_t nxy = (int64_t)d->nx * d->ny;
  const size_t need = sizeof(T) * 3 * (size_t)nxy * d->nz;
  int rc = ensure_scratch(ctx, need);
  if (rc) return rc;
  T *prep = static_cast<T *>(ctx->scratch);
  const int dir = d->dir > 0 ? 1 : -1;
  {
    const int64_t n = nxy * d->nz;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(lc2_prep<T>, dim3(blocks), dim3(256), 0, ctx->stream, b->chi500, b->g500, prep,
                       (int)nxy, d->nz, dir, d->ds[2] / fabs(d->vec[2]));
    M3DRT_HIP(hipGetLastError());
  }
  Lc2Args<T> a;
  a.chi = static_cast<const T *>(b->chi);
  a.S = static_cast<const T *>(b->S);
  a.gchi = static_cast<const T *>(b->ghost_chi);
  a.gS = static_cast<const T *>(b->ghost_S);
  a.gI = static_cast<const T *>(b->ghost_I);
  a.prep = prep;
  a.eta_row = b->eta_row;
  a.eta_mask = b->eta_mask;
  a.J = static_cast<T *>(b->J);
  a.PsiJ = (d->flags & M3DRT_GET_RATES) ? static_cast<T *>(b->PsiJ) : nullptr;
  a.Iout = static_cast<T *>(b->I_out);
  a.Psiout = (d->flags & M3DRT_GET_RATES) ? static_cast<T *>(b->Psi_out) : nullptr;
  a.Irem = static_cast<T *>(b->Irem_out);
  a.esc = static_cast<T *>(b->escape_out);
  a.exitI = static_cast<T *>(b->exit_I);
  a.cs_f = d->cs_f; a.cs_x = d->cs_x; a.cs_y = d->cs_y; a.cs_z = d->cs_z;
  a.gs_f = d->gs_f; a.gs_x = d->gs_x; a.gs_y = d->gs_y;
  a.eta_ld = d->eta_ld;
  a.eta_nrow = d->eta_nrow;
  a.nf = d->nf; a.nx = d->nx; a.ny = d->ny; a.nz = d->nz; a.dir = dir;
  a.split = (d->flags & M3DRT_SPLIT_RATES) ? 1 : 0;
  a.f_fast = d->cs_f <= d->cs_x ? 1 : 0;
  a.mdtau = -(T)d->maser_dtau_max;
  a.imax = (T)d->maser_Inu_max;
  a.wj = (T)d->wj;
  if (sizeof(T) == 4) {
    a.mdtau = (T)(-(float)d->maser_dtau_max);
    a.imax = (T)(float)d->maser_Inu_max;
  }
  const int64_t n = nxy * d->nf;
  const int blocks = (int)((n + 255) / 256);
  if (a.eta_row)
    hipLaunchKernelGGL((lc2_sweep<T, true>), dim3(blocks), dim3(256), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL((lc2_sweep<T, false>), dim3(blocks), dim3(256), 0, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

int sweep_lc2(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  if (d->prec == M3DRT_F64) return run_lc2<double>(ctx, d, b);
  return run_lc2<float>(ctx, d, b);
}

}  // namespace m3drt
