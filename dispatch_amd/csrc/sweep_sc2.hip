// 3D short-characteristics sweep (the semantics of short_char2,
// SRC/rt_solvers_mod.f90:224-507) fused with the angle integration of
// integrate_intensity (SRC/rt_patch_mod.f90:3399-3431), for gfx950.
//
// Mapping.  A ray's sweep is sequential in z and, for one frequency, couples a
// cell only to the previous plane through a 4x4 periodic stencil.  Frequencies are
// independent.  So one workgroup owns ONE frequency and the WHOLE periodic plane:
// the upwind plane's intensity I and x = chi/chi500 sit in LDS, and the workgroup
// advances plane by plane with two workgroup barriers per plane and no traffic
// between workgroups.  The grid is the nf frequencies of one direction.
//
// Threads.  nx*nstrips threads: thread t takes column x = t % nx and a strip of
// L consecutive rows y.  Down a strip the 4 stencil rows of consecutive cells
// overlap in 3, so each cell reads ONE new upwind row (4 I, 2 x, 2 S values) and
// reduces it in x at once; the y-reduction runs over a 4-deep register window.
// Lanes of a wavefront are consecutive x, so with the native layout every HBM
// access of a wavefront is one contiguous run.
//
// Traffic per ray-frequency-cell update: chi and S read, J (and PsiJ) read and
// written.  The upwind plane stays on chip; S of the upwind plane, when it does not
// fit the LDS next to I and x, is re-read through L1/L2 where this workgroup read
// it one plane earlier.
#include "m3drt_internal.hpp"
#include "sweep_sc2.hpp"
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <algorithm>

namespace m3drt {

template <typename T>
struct Sc2Args {
  const T *chi, *S;
  const T *gchi, *gS, *gI;
  // per cell, planes 0..nz-1, then the ghost plane: dt500 of the cell's step and
  // rc = 1/chi500 (the ghost plane's entry is 1/g500)
  const T *dt500, *rc;
  const int32_t *eta_row;
  const uint8_t *eta_mask;
  T *J, *PsiJ, *Iout, *Psiout, *Irem, *esc, *exitI;
  int64_t cs_f, cs_z, gs_f, eta_ld, eta_nrow;
  uint32_t cs_x, cs_y, gs_x, gs_y;
  int nf, nx, ny, nz, dir, L, xo, yo, split;
  T xk0, xk1, xk2, xk3, yk0, yk1, yk2, yk3, hx, hy, mdtau, imax, wj;
};

template <typename T>
struct Row {
  T rI, rmin, rX, rS, rE;
};

// prep: dt500 of every cell's step, in FP64 from the FP32 chi500 as the reference
// forms it (SRC/rt_solvers_mod.f90:355-370), and rc = 1/chi500; then 1/g500 of the
// ghost plane.
template <typename T>
__global__ void __launch_bounds__(256) sc2_prep(const float *__restrict__ chi500,
                                                const float *__restrict__ chi500_up,
                                                const float *__restrict__ g500, T *__restrict__ dt500,
                                                T *__restrict__ rc, int nxy, int nz, double dsup) {
  const int64_t n = (int64_t)nxy * nz;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n + nxy;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < n) {
      const double c = (double)chi500[i];
      const double cu = (double)chi500_up[i];
      const double r = cu / c;
      double d;
      if (fabs(r - 1.0) < 1e-3)
        d = dsup * c * (1.0 + (r - 1.0) * (0.5 - (r - 1.0) / 12.0));
      else
        d = dsup * c * (r - 1.0) / log(r);
      dt500[i] = (T)d;
      rc[i] = (T)(1.0 / c);
    } else {
      dt500[i] = (T)0;
      rc[i] = (T)(1.0 / (double)g500[i - n]);
    }
  }
}

int sc2_prep_f32(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const float *chi500, const float *chi500_up,
                 const float *g500, float **dt500, float **rc) {
  const int64_t nxy = (int64_t)d->nx * d->ny;
  const size_t nprep = (size_t)nxy * (d->nz + 1);
  const int st = ensure_scratch(ctx, sizeof(float) * 2 * nprep);
  if (st) return st;
  *dt500 = static_cast<float *>(ctx->scratch);
  *rc = *dt500 + nprep;
  const double dsup = d->ds[2] / fabs(d->vec[2]);
  const int blocks = (int)std::min<int64_t>(((int64_t)nprep + 255) / 256, 65536);
  hipLaunchKernelGGL(sc2_prep<float>, dim3(blocks), dim3(256), 0, ctx->stream, chi500, chi500_up, g500, *dt500,
                     *rc, (int)nxy, d->nz, dsup);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

// Global inputs of one cell: opacity, source function, the cell's dt500 and
// 1/chi500, and the J / PsiJ it accumulates into.
template <typename T>
struct CellIn {
  T chi, S, dt, rc, J, PJ;
};

// Uniform base pointers of one plane of every cube (per frequency).
template <typename T>
struct Plane {
  const T *chi, *S, *dt, *rc;
  T *J, *PJ, *Iout, *Psiout, *Irem, *esc;
};

// loads and stores at a wave-uniform base plus a 32-bit per-lane byte offset, the
// form the global_load/store saddr encoding takes
template <typename T>
__device__ __forceinline__ T ldb(const T *base, uint32_t off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T>
__device__ __forceinline__ void stb(T *base, uint32_t off, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}

// PD: cells of global loads in flight ahead of the one computed; the loads of a
// plane's first PD cells are issued before the barriers that end the previous
// plane, so no wavefront waits on HBM when a plane opens.
// HOLD: keep x and S of the new plane in registers until they go to LDS (FP32);
// otherwise (FP64) re-read them through the cache after the barrier.
// EXACT: every strip holds exactly LMAX rows (ny = LMAX * strips), so the cell loop
// is straight-line code: no lane-dependent branch joins the registers of the load
// pipeline, and the loads stay in flight instead of being waited on at every cell.
// OUT: the outputs, fixed at compile time (OUT_J, OUT_J | OUT_PJ, OUT_I: the
// intensity cube alone, as the adjoint's recomputation asks), or OUT_ANY for
// the general form that tests every output pointer at run time.
constexpr int OUT_ANY = 0, OUT_J = 1, OUT_PJ = 2, OUT_I = 4;

template <typename T, int LMAX, bool EXACT, int OUT, bool X_LDS, bool S_LDS, bool ETA, bool HOLD>
__global__ void __launch_bounds__(sizeof(T) == 8 ? 512 : 1024) sc2_sweep(const Sc2Args<T> a) {
  using namespace dev;
  constexpr int PD = sizeof(T) == 8 ? 1 : 2;
  constexpr uint32_t TB = sizeof(T);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nx = a.nx, ny = a.ny, nxy = nx * ny;
  T *sI = reinterpret_cast<T *>(smem_raw);
  T *sX = sI + nxy;
  T *sS = sI + nxy * (X_LDS ? 2 : 1);

  const int f = blockIdx.x;
  const int tid = threadIdx.x;
  const int x = tid % nx;
  const int strip = tid / nx;
  const int ybeg = strip * a.L;
  const int ylen = EXACT ? LMAX : max(0, min(a.L, ny - ybeg));

  // stencil columns: the reference's xup_t(1:4, x) (SRC/rt_solvers_mod.f90:316-319)
  const int xc0 = wrap(x + a.xo, nx), xc1 = wrap(x + a.xo + 1, nx);
  const int xc2 = wrap(x + a.xo + 2, nx), xc3 = wrap(x + a.xo + 3, nx);
  const uint32_t csy = a.cs_y * TB, gsy = a.gs_y * TB;      // row strides in bytes
  const uint32_t ox = x * a.cs_x * TB;                        // this lane's column
  const uint32_t xo1 = xc1 * a.cs_x * TB, xo2 = xc2 * a.cs_x * TB;
  const uint32_t gxo1 = xc1 * a.gs_x * TB, gxo2 = xc2 * a.gs_x * TB;
  const T ohx = (T)1 - a.hx, ohy = (T)1 - a.hy;
  const bool want_J = OUT == OUT_ANY ? a.J != nullptr : (OUT & OUT_J) != 0;
  const bool want_PJ = OUT == OUT_ANY ? a.PsiJ != nullptr : (OUT & OUT_PJ) != 0;
  const bool split = OUT == OUT_ANY && a.split;

  const T *__restrict__ gchi = a.gchi + f * a.gs_f;
  const T *__restrict__ gS = a.gS + f * a.gs_f;
  const T *__restrict__ gI = a.gI + f * a.gs_f;
  const T *__restrict__ ghost_rc = a.rc + (int64_t)a.nz * nxy;

  auto plane_of = [&](int step) { return a.dir > 0 ? step : a.nz - 1 - step; };
  auto ptrs = [&](int step) -> Plane<T> {
    Plane<T> P;
    const int z = plane_of(step);
    const int64_t zoff = (int64_t)z * a.cs_z + (int64_t)f * a.cs_f;
    P.chi = a.chi + zoff;
    P.S = a.S + zoff;
    P.dt = a.dt500 + (int64_t)z * nxy;
    P.rc = a.rc + (int64_t)z * nxy;
    P.J = a.J ? a.J + zoff : nullptr;
    P.PJ = a.PsiJ ? a.PsiJ + zoff : nullptr;
    P.Iout = (OUT == OUT_ANY || (OUT & OUT_I)) && a.Iout ? a.Iout + zoff : nullptr;
    P.Psiout = OUT == OUT_ANY && a.Psiout ? a.Psiout + zoff : nullptr;
    P.Irem = a.Irem ? a.Irem + zoff : nullptr;
    P.esc = a.esc ? a.esc + zoff : nullptr;
    return P;
  };
  auto fetch = [&](const Plane<T> &P, int y) -> CellIn<T> {
    CellIn<T> v;
    const uint32_t o = ox + y * csy;
    const uint32_t cb = (uint32_t)(y * nx + x) * TB;
    v.chi = ldb(P.chi, o);
    v.S = ldb(P.S, o);
    v.dt = ldb(P.dt, cb);
    v.rc = ldb(P.rc, cb);
    v.J = want_J ? ldb((const T *)P.J, o) : (T)0;
    v.PJ = want_PJ ? ldb((const T *)P.PJ, o) : (T)0;
    return v;
  };

  // loads of the first plane's first cells, then the upstream ghost plane
  CellIn<T> pre[PD], ring[PD];
  {
    const Plane<T> P0 = ptrs(0);
#pragma unroll
    for (int k = 0; k < PD; ++k)
      if (EXACT || k < ylen) pre[k] = fetch(P0, ybeg + k);
  }
  for (int c = tid; c < nxy; c += blockDim.x) {
    const int cx = c % nx, cy = c / nx;
    const uint32_t go = (cx * a.gs_x + cy * a.gs_y) * TB;
    sI[c] = ldb(gI, go);
    if (X_LDS) sX[c] = ldb(gchi, go) * ghost_rc[c];
    if (S_LDS) sS[c] = ldb(gS, go);
  }
  lds_barrier();

  for (int step = 0; step < a.nz; ++step) {
    const int z = plane_of(step);
    const Plane<T> P = ptrs(step);
    // ybeg made opaque per plane: the row addresses derived from it are a few
    // integer ops each, cheaper than the registers that hoisting them would hold
    int yb = ybeg;
    asm volatile("" : "+v"(yb));
    // the upwind plane in global memory, for the values not held in LDS
    const bool up_ghost = step == 0;
    const int64_t uoff = (int64_t)(z - a.dir) * a.cs_z + (int64_t)f * a.cs_f;
    const T *__restrict__ up_chi = up_ghost ? gchi : a.chi + uoff;
    const T *__restrict__ up_S = up_ghost ? gS : a.S + uoff;
    const uint32_t uy = up_ghost ? gsy : csy;
    const uint32_t ux1 = up_ghost ? gxo1 : xo1, ux2 = up_ghost ? gxo2 : xo2;
    const T *__restrict__ up_rc = up_ghost ? ghost_rc : a.rc + (int64_t)(z - a.dir) * nxy;

    // raw upwind row values (4 I, 2 x, 2 S) and their reduction in x
    struct RawRow { T i0, i1, i2, i3, x1, x2, s1, s2; };
    auto read_row = [&](int yr) -> RawRow {
      RawRow r;
      const int base = yr * nx;
      r.i0 = sI[base + xc0]; r.i1 = sI[base + xc1]; r.i2 = sI[base + xc2]; r.i3 = sI[base + xc3];
      if (X_LDS) {
        r.x1 = sX[base + xc1];
        r.x2 = sX[base + xc2];
      } else {
        r.x1 = ldb(up_chi, ux1 + yr * uy) * up_rc[base + xc1];
        r.x2 = ldb(up_chi, ux2 + yr * uy) * up_rc[base + xc2];
      }
      if (S_LDS) {
        r.s1 = sS[base + xc1];
        r.s2 = sS[base + xc2];
      } else {
        r.s1 = ldb(up_S, ux1 + yr * uy);
        r.s2 = ldb(up_S, ux2 + yr * uy);
      }
      return r;
    };
    auto reduce_row = [&](const RawRow &q) -> Row<T> {
      Row<T> r;
      r.rI = a.xk0 * q.i0 + a.xk1 * q.i1 + a.xk2 * q.i2 + a.xk3 * q.i3;
      r.rmin = fmin(q.i1, q.i2);
      r.rX = ohx * q.x1 + a.hx * q.x2;
      r.rS = ohx * q.s1 + a.hx * q.s2;
      r.rE = ETA ? ohx * q.x1 * q.s1 + a.hx * q.x2 * q.s2 : (T)0;
      return r;
    };

    T newI[LMAX], newX[HOLD ? LMAX : 1], newS[HOLD ? LMAX : 1];
    Row<T> w0 = reduce_row(read_row(wrap_near(yb + a.yo, ny)));
    Row<T> w1 = reduce_row(read_row(wrap_near(yb + a.yo + 1, ny)));
    Row<T> w2 = reduce_row(read_row(wrap_near(yb + a.yo + 2, ny)));
    RawRow next = read_row(wrap_near(yb + a.yo + 3, ny));
#pragma unroll
    for (int k = 0; k < LMAX; ++k) {
      if (EXACT || k < ylen) {
        __builtin_amdgcn_sched_barrier(0);
        const int y = yb + k;
        const CellIn<T> in = k < PD ? pre[k] : ring[k % PD];
        // keep PD cells of global loads in flight, and one upwind row of LDS reads
        if (EXACT ? k + PD < LMAX : k + PD < ylen) ring[k % PD] = fetch(P, y + PD);
        const Row<T> w3 = reduce_row(next);
        if (EXACT ? k + 1 < LMAX : k + 1 < ylen) next = read_row(wrap_near(y + a.yo + 4, ny));
        // upwind values, SRC/rt_solvers_mod.f90:379-394
        const T xxup = ohy * w1.rX + a.hy * w2.rX;
        const T ssup = ohy * w1.rS + a.hy * w2.rS;
        T Inuup = a.yk0 * w0.rI + a.yk1 * w1.rI + a.yk2 * w2.rI + a.yk3 * w3.rI;
        Inuup = fmax(Inuup, fmin(w1.rmin, w2.rmin));
        const uint32_t o = ox + y * csy;
        const T chiv = in.chi, ss = in.S, dt500 = in.dt, rc = in.rc;
        const T xx = chiv * rc;
        T dt = (T)0.5 * (xxup + xx) * dt500;
        dt = fmax(dt, a.mdtau);
        T E, W, Ps, Es;
        lin_weights(dt, E, W, Ps, Es);
        const T sgnchi = copysign(fmax(fabs(chiv), (T)1.17549435e-38f), chiv);
        T Inu, Ir = (T)0, psiout;
        bool off;
        if (!ETA) {
          // a cell without eta flags: the S step (SRC/rt_solvers_mod.f90:419-440, 465-475)
          Ir = Inuup * E + W * ssup;
          Inu = Ir + Ps * ss;
          off = fmin(dt, chiv) <= (T)0;
          if (split) off = off || (a.imax - fabs(Inu)) <= (T)0;
          psiout = off ? (T)0 : Ps * rcp(sgnchi);
        } else {
          // flagged pairs take the eta step (SRC/rt_solvers_mod.f90:441-463, 476-490)
          const int row = a.eta_row[(int64_t)z * nxy + y * nx + x];
          const bool special = row > 0 && row <= a.eta_nrow && a.eta_mask[f + a.eta_ld * (int64_t)(row - 1)] != 0;
          const T eup = ohy * w1.rE + a.hy * w2.rE;
          const T ee = xx * ss;
          const T rxb = dt500 * rcp(dt != (T)0 ? dt : (T)1);
          Ir = special ? Inuup * E + W * eup * rxb : Inuup * E + W * ssup;
          Inu = special ? Ir + Ps * ee * rxb : Ir + Ps * ss;
          off = fmin(fmin(dt, chiv), special ? xxup : (T)1) <= (T)0;
          if (split) {
            off = off || (a.imax - fabs(Inu)) <= (T)0;
            Es = special ? Es + (T)0.5 * Ps * (xxup - xx) * rxb : Es;
          }
          psiout = off ? (T)0 : (special ? Ps * rxb * rc : Ps * rcp(sgnchi));
        }
        Inu = fmax(-a.imax, fmin(a.imax, Inu));
        newI[k] = Inu;
        if (HOLD) {
          newX[k] = xx;
          newS[k] = ss;
        }
        if (want_J) stb(P.J, o, in.J + a.wj * Inu);
        if (want_PJ) stb(P.PJ, o, in.PJ + a.wj * psiout);
        if (P.Iout) stb(P.Iout, o, Inu);
        if (P.Psiout) stb(P.Psiout, o, psiout);
        if (split) {
          stb(P.Irem, o, off ? Inu : Ir);
          stb(P.esc, o, off ? (T)1 : Es);
        }
        w0 = w1;
        w1 = w2;
        w2 = w3;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    // the next plane's first cells, in flight across the barriers
    if (step + 1 < a.nz) {
      const Plane<T> Pn = ptrs(step + 1);
#pragma unroll
      for (int k = 0; k < PD; ++k)
        if (EXACT || k < ylen) pre[k] = fetch(Pn, yb + k);
    }
    lds_barrier();  // every read of the upwind plane is done
    const bool exit_plane = a.exitI && step == a.nz - 1;
#pragma unroll
    for (int k = 0; k < LMAX; ++k) {
      if (EXACT || k < ylen) {
        const int y = yb + k;
        const int c = y * nx + x;
        sI[c] = newI[k];
        if (exit_plane) a.exitI[f * a.gs_f + x * a.gs_x + y * a.gs_y] = newI[k];
        if (HOLD) {
          if (X_LDS) sX[c] = newX[k];
          if (S_LDS) sS[c] = newS[k];
        } else {
          const uint32_t o = ox + y * csy;
          if (X_LDS) sX[c] = ldb(P.chi, o) * P.rc[c];
          if (S_LDS) sS[c] = ldb(P.S, o);
        }
      }
    }
    lds_barrier();
  }
}

template <typename T, int LMAX, bool EXACT, int OUT, bool X_LDS, bool S_LDS, bool ETA>
static int launch_sc2(m3drt_ctx *ctx, const Sc2Args<T> &a, int nthreads, size_t lds) {
  auto kern = sc2_sweep<T, LMAX, EXACT, OUT, X_LDS, S_LDS, ETA, (sizeof(T) == 4)>;
  M3DRT_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(a.nf), dim3(nthreads), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

template <typename T, int LMAX, bool X_LDS, bool S_LDS>
static int dispatch_eta(m3drt_ctx *ctx, const Sc2Args<T> &a, int nt, size_t lds) {
  if (a.eta_row) return launch_sc2<T, LMAX, false, OUT_ANY, X_LDS, S_LDS, true>(ctx, a, nt, lds);
  return launch_sc2<T, LMAX, false, OUT_ANY, X_LDS, S_LDS, false>(ctx, a, nt, lds);
}

template <typename T, int LMAX>
static int dispatch_lds(m3drt_ctx *ctx, const Sc2Args<T> &a, int nt, int nplanes) {
  const size_t plane = (size_t)a.nx * a.ny * sizeof(T);
  if (nplanes >= 3) return dispatch_eta<T, LMAX, true, true>(ctx, a, nt, 3 * plane);
  if (nplanes == 2) return dispatch_eta<T, LMAX, true, false>(ctx, a, nt, 2 * plane);
  return dispatch_eta<T, LMAX, false, false>(ctx, a, nt, plane);
}

// The straight-line form: exact strips of L rows, J (and PsiJ) as the only outputs,
// no eta flags.  L comes from the set instantiated for the plane sizes in use
// (60, 96, 120 cells square: FP32 4, 12, 15; FP64 12, 24, 30).
template <typename T, int L, int OUT>
static int exact_lds(m3drt_ctx *ctx, const Sc2Args<T> &a, int nt, int nplanes) {
  const size_t plane = (size_t)a.nx * a.ny * sizeof(T);
  if (nplanes >= 3) return launch_sc2<T, L, true, OUT, true, true, false>(ctx, a, nt, 3 * plane);
  if (nplanes == 2) return launch_sc2<T, L, true, OUT, true, false, false>(ctx, a, nt, 2 * plane);
  return launch_sc2<T, L, true, OUT, false, false, false>(ctx, a, nt, plane);
}

template <typename T, int L>
static int exact_out(m3drt_ctx *ctx, const Sc2Args<T> &a, int nt, int nplanes, bool any) {
  // any other set of outputs (FP32, strips of at most 12 rows: at 15 and more the
  // run-time outputs need more registers than a lane has): every output checked at
  // run time
  if constexpr (sizeof(T) == 4 && L <= 12)
    if (any) return exact_lds<T, L, OUT_ANY>(ctx, a, nt, nplanes);
  if (!a.J) return exact_lds<T, L, OUT_I>(ctx, a, nt, nplanes);
  if (a.PsiJ) return exact_lds<T, L, OUT_J | OUT_PJ>(ctx, a, nt, nplanes);
  return exact_lds<T, L, OUT_J>(ctx, a, nt, nplanes);
}

static const int kExactL32[] = {4, 12, 15};
static const int kExactL64[] = {12, 24, 30};

template <typename T>
static int exact_dispatch(m3drt_ctx *ctx, const Sc2Args<T> &a, int L, int nt, int nplanes, bool any) {
  if (sizeof(T) == 4) {
    if (L == 4) return exact_out<T, 4>(ctx, a, nt, nplanes, any);
    if (L == 12) return exact_out<T, 12>(ctx, a, nt, nplanes, any);
    return exact_out<T, 15>(ctx, a, nt, nplanes, any);
  }
  if (L == 12) return exact_out<T, 12>(ctx, a, nt, nplanes, false);
  if (L == 24) return exact_out<T, 24>(ctx, a, nt, nplanes, false);
  return exact_out<T, 30>(ctx, a, nt, nplanes, false);
}

template <typename T>
static int run_sc2(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  const int64_t nxy = (int64_t)nx * ny;
  if (nx > 1024) return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep: nx > 1024");
  const size_t plane = (size_t)nxy * sizeof(T);
  const int nplanes = (int)std::min<size_t>(3, (size_t)ctx->lds_max / plane);
  if (nplanes < 1) return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep: the plane does not fit the LDS");
  // strips: as many rows per thread as the workgroup needs (1024 threads in FP32;
  // 512 in FP64, whose registers hold twice the bytes)
  const int tmax = sizeof(T) == 8 ? 512 : 1024;
  if (nx > tmax) return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep: nx beyond the workgroup size");
  const int smax = std::max(1, tmax / nx);
  int L = (ny + smax - 1) / smax;
  if (L > 32) return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep: ny too large for the strip mapping");
  // the straight-line form when only J (and PsiJ), or only the intensity cube, are
  // asked for and an instantiated strip length divides ny
  int Lexact = 0;
  const bool psi_out = (d->flags & M3DRT_GET_RATES) && (b->Psi_out || b->PsiJ);
  const bool plain_out = !b->eta_row && !(d->flags & M3DRT_SPLIT_RATES) &&
                         ((!b->I_out && b->J && !((d->flags & M3DRT_GET_RATES) && b->Psi_out)) ||
                          (b->I_out && !b->J && !psi_out));
  // FP32 with any other outputs but the eta step (the ALI's I, Psi, Irem, escape): the
  // exact strips with run-time output checks
  const bool any_out = !plain_out && !b->eta_row && sizeof(T) == 4;
  if ((plain_out || any_out) && !getenv("M3DRT_NO_EXACT")) {
    for (int c : (sizeof(T) == 4 ? kExactL32 : kExactL64))
      if (c >= L && ny % c == 0 && !Lexact && !(any_out && c > 12)) Lexact = c;
  }
  if (Lexact) L = Lexact;
  const int nstrips = (ny + L - 1) / L;
  const int nt = nx * nstrips;

  // geometry of the stencil, SRC/ray_mod.f90:493-508 (set_stencil) and 581-596 (cc_kernel)
  const double dsup = d->ds[2] / fabs(d->vec[2]);
  const double wt1 = -d->vec[0] / d->ds[0] * dsup, wt2 = -d->vec[1] / d->ds[1] * dsup;
  const double h1 = wt1 > 0 ? wt1 : wt1 + 1, h2 = wt2 > 0 ? wt2 : wt2 + 1;
  auto cc = [](double s, double k[4]) {
    const double s2 = s * s, s3 = s2 * s;
    k[0] = 0.5 * (-s3 + 2 * s2 - s);
    k[1] = 0.5 * (3 * s3 - 5 * s2 + 2);
    k[2] = 0.5 * (-3 * s3 + 4 * s2 + s);
    k[3] = 0.5 * (s3 - s2);
  };
  double kx[4], ky[4];
  cc(h1, kx);
  cc(h2, ky);

  // in-plane offsets are 32-bit: every (x, y) offset of a cube and plane must fit
  const uint64_t span = (uint64_t)(nx - 1) * d->cs_x + (uint64_t)(ny - 1) * d->cs_y;
  const uint64_t gspan = (uint64_t)(nx - 1) * d->gs_x + (uint64_t)(ny - 1) * d->gs_y;
  if ((span + 1) * sizeof(T) >= (1ull << 32) || (gspan + 1) * sizeof(T) >= (1ull << 32))
    return set_error(M3DRT_ELIMIT, "m3drt_sc_sweep: in-plane offsets exceed 32 bits");

  const size_t nprep = (size_t)nxy * (nz + 1);
  int rc = ensure_scratch(ctx, sizeof(T) * 2 * nprep);
  if (rc) return rc;
  T *dt500 = static_cast<T *>(ctx->scratch);
  T *rcp = dt500 + nprep;
  {
    const int blocks = (int)std::min<int64_t>(((int64_t)nprep + 255) / 256, 65536);
    hipLaunchKernelGGL(sc2_prep<T>, dim3(blocks), dim3(256), 0, ctx->stream, b->chi500, b->chi500_up,
                       b->g500, dt500, rcp, (int)nxy, nz, dsup);
    M3DRT_HIP(hipGetLastError());
  }

  Sc2Args<T> a;
  a.chi = static_cast<const T *>(b->chi);
  a.S = static_cast<const T *>(b->S);
  a.gchi = static_cast<const T *>(b->ghost_chi);
  a.gS = static_cast<const T *>(b->ghost_S);
  a.gI = static_cast<const T *>(b->ghost_I);
  a.dt500 = dt500;
  a.rc = rcp;
  a.eta_row = b->eta_row;
  a.eta_mask = b->eta_mask;
  a.J = static_cast<T *>(b->J);
  a.PsiJ = (d->flags & M3DRT_GET_RATES) ? static_cast<T *>(b->PsiJ) : nullptr;
  a.Iout = static_cast<T *>(b->I_out);
  a.Psiout = (d->flags & M3DRT_GET_RATES) ? static_cast<T *>(b->Psi_out) : nullptr;
  a.Irem = static_cast<T *>(b->Irem_out);
  a.esc = static_cast<T *>(b->escape_out);
  a.exitI = static_cast<T *>(b->exit_I);
  a.cs_f = d->cs_f; a.cs_z = d->cs_z; a.gs_f = d->gs_f;
  a.cs_x = (uint32_t)d->cs_x; a.cs_y = (uint32_t)d->cs_y;
  a.gs_x = (uint32_t)d->gs_x; a.gs_y = (uint32_t)d->gs_y;
  a.eta_ld = d->eta_ld;
  a.eta_nrow = d->eta_nrow;
  a.nf = d->nf; a.nx = nx; a.ny = ny; a.nz = nz; a.dir = d->dir > 0 ? 1 : -1;
  a.L = L;
  a.xo = wt1 > 0 ? -1 : -2;
  a.yo = wt2 > 0 ? -1 : -2;
  a.split = (d->flags & M3DRT_SPLIT_RATES) ? 1 : 0;
  a.xk0 = (T)(float)kx[0]; a.xk1 = (T)(float)kx[1]; a.xk2 = (T)(float)kx[2]; a.xk3 = (T)(float)kx[3];
  a.yk0 = (T)(float)ky[0]; a.yk1 = (T)(float)ky[1]; a.yk2 = (T)(float)ky[2]; a.yk3 = (T)(float)ky[3];
  if (sizeof(T) == 8) {  // the plain FP64 sweep keeps the stencil in FP64
    a.xk0 = (T)kx[0]; a.xk1 = (T)kx[1]; a.xk2 = (T)kx[2]; a.xk3 = (T)kx[3];
    a.yk0 = (T)ky[0]; a.yk1 = (T)ky[1]; a.yk2 = (T)ky[2]; a.yk3 = (T)ky[3];
  }
  // inner_fraction (SRC/rt_solvers_mod.f90:500-506): the first moment of the kernel
  a.hx = -a.xk0 + a.xk2 + (T)2 * a.xk3;
  a.hy = -a.yk0 + a.yk2 + (T)2 * a.yk3;
  a.mdtau = -(T)d->maser_dtau_max;
  a.imax = (T)d->maser_Inu_max;
  a.wj = (T)d->wj;
  if (sizeof(T) == 4) {  // the reference rounds the caps to real(4)
    a.mdtau = (T)(-(float)d->maser_dtau_max);
    a.imax = (T)(float)d->maser_Inu_max;
  }
  if (getenv("M3DRT_VERBOSE"))
    fprintf(stderr, "[m3drt] sc2 %s nf=%d plane=%dx%d nz=%d: %d threads, strip %d%s, %d LDS planes of %zu B (lds_max %d)%s\n",
            sizeof(T) == 8 ? "f64" : "f32", d->nf, nx, ny, nz, nt, L, Lexact ? " exact" : "", nplanes, plane,
            ctx->lds_max, b->eta_row ? ", eta" : "");
  if (Lexact) return exact_dispatch<T>(ctx, a, L, nt, nplanes, any_out);
  if (L <= 4) return dispatch_lds<T, 4>(ctx, a, nt, nplanes);
  if (L <= 8) return dispatch_lds<T, 8>(ctx, a, nt, nplanes);
  if (L <= 16) return dispatch_lds<T, 16>(ctx, a, nt, nplanes);
  return dispatch_lds<T, 32>(ctx, a, nt, nplanes);
}

int sweep_sc2(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  const int fast = sweep_sc2_fast(ctx, d, b);
  if (fast >= 0) return fast;
  if (d->prec == M3DRT_F64) return run_sc2<double>(ctx, d, b);
  return run_sc2<float>(ctx, d, b);
}

}  // namespace m3drt
