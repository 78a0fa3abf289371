// Reverse (adjoint) sweep of the 3D short-characteristics formal solution with the
// semantics of short_char2's S step (SRC/rt_solvers_mod.f90:224-507, 409-440).  The
// reference has no adjoint; this is the exact derivative of the forward kernel's
// arithmetic, taken plane by plane against the sweep direction:
//
//   I    = clamp(E Iup + W Sup + Psi S, +-Imax)
//   Iup  = max(cubic(I'), min2x2(I'))        Sup = bilinear(S')
//   xup  = bilinear(chi' / chi500')          dt  = max(dt500 (xup + chi/chi500)/2, -dtmax)
//
// (primes on the upwind plane).  The cotangent g of a cell's I sends g Psi to its S,
// g E to the upwind intensities (the Keys weights, or the smallest inner value where
// the clamp holds), g W to the upwind S (bilinear) and g (E' Iup + W' Sup + Psi' S)
// through dt to chi of the cell and of the upwind plane.  chi500 is held fixed.
//
// Mapping.  One workgroup per frequency owns the periodic plane, as the forward.
// Each thread owns the cells c = t + k * blockDim (k < MAXC) of every plane.  Per
// plane, against the sweep:
//   (a) the upwind plane's I, chi/chi500 and S go to LDS;
//   (b) every target recomputes its step from LDS and adds its own chi and S
//       cotangents into the global sums;
//   (c) the per-target upwind cotangents (intensity, chi/chi500, S, and the clamp's
//       choice) replace the upwind plane in LDS;
//   (d) every upwind cell gathers what the targets sent it: the transposed stencils,
//       a gather with mirrored offsets, so no atomics and a fixed summation order;
//       the chi and S shares go into the global sums, the intensity's is held for
//       the next plane.  The cube values a plane reads again one step later (its
//       chi, S and cotangent sums) are L2 hits, so each is one HBM pass per ray.
// The forward intensity of every plane is an input (the forward sweep's I_out).
#include "m3drt_internal.hpp"
#include "sweep_sc2.hpp"
#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <type_traits>

namespace m3drt {

namespace {

struct AdjArgs {
  const float *chi, *S, *I, *gchi, *gS, *gI;
  const float *dt500, *rc;     // per cell, planes 0..nz-1 then the ghost plane's rc
  const float *Jbar, *Ibar, *exitbar;
  float *chibar, *Sbar, *gchibar, *gSbar, *gIbar;
  int64_t cs_f, cs_x, cs_y, cs_z, gs_f, gs_x, gs_y;
  int nf, nx, ny, nz, dir, xo, yo;
  float kx[4], ky[4], wx[2], wy[2];
  float mdtau, imax, wj;
};

// The weights of the linear step (dev::lin_weights, the forward's) and their
// derivatives: E' = -E, Psi' = (u0 - t E) / t^2, W' = E - Psi'.  For |t| <= 1,
// Psi' = sum_{n>=2} (-1)^n (n-1)/n! t^(n-2), the series of the same quantity,
// to t^9 (the first omitted term is below 5e-8 of Psi'(0)).
__device__ __forceinline__ void lin_weights_deriv(float t, float &E, float &W, float &Ps, float &dE, float &dW,
                                                  float &dPs) {
  float Es;
  dev::lin_weights(t, E, W, Ps, Es);
  const float ser = 0.5f + t * (-3.333333433e-01f + t * (1.250000000e-01f + t * (-3.333333507e-02f +
                    t * (6.944444496e-03f + t * (-1.190476236e-03f + t * (1.736111112e-04f +
                    t * (-2.204585518e-05f + t * (2.480158764e-06f + t * (-2.505210591e-07f)))))))));
  const float u0 = 1.0f - E;
  const float far = (u0 - t * E) / (t * t);
  dPs = fabsf(t) <= 1.0f ? ser : far;
  dE = -E;
  dW = E - dPs;
}

// A float at a byte offset from a wave-uniform base: the offset is a 32-bit lane
// value, so the access takes the scalar-base + 32-bit-offset form and no 64-bit
// address is formed per lane and row.  The host takes the kernels that use it only
// where a plane's byte span fits 32 bits.
__device__ __forceinline__ float ldb(const float *base, uint32_t off) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + off);
}
__device__ __forceinline__ void stb(float *base, uint32_t off, float v) {
  *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + off) = v;
}

// MAXC: cells per thread.  FULL: every thread owns exactly MAXC cells (nx*ny =
// MAXC * blockDim), so the per-cell loops carry no guards and the loads of all of a
// thread's cells can be in flight together.  SEED: a J or intensity-cube cotangent
// is given (the exit plane's alone otherwise).  The upstream ghost plane's step is
// peeled off the plane loop, so no plane of the loop branches on it.
template <int MAXC, bool FULL, bool SEED>
__global__ void __launch_bounds__(1024) sc2_adjoint(const AdjArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nx = a.nx, ny = a.ny, nxy = nx * ny;
  float *sI = reinterpret_cast<float *>(smem_raw);  // upwind I, then the intensity cotangents
  float *sX = sI + nxy;                               // upwind chi/chi500, then their cotangents
  float *sS = sX + nxy;                               // upwind S, then its cotangents
  uint8_t *sC = reinterpret_cast<uint8_t *>(sS + nxy);  // the clamp's choice per target
  const int f = blockIdx.x, nt = blockDim.x, tid = threadIdx.x;
  const float *gchi = a.gchi + f * a.gs_f, *gS = a.gS + f * a.gs_f, *gI = a.gI + f * a.gs_f;
  const float *ghost_rc = a.rc + (int64_t)a.nz * nxy;
  auto plane_of = [&](int step) { return a.dir > 0 ? step : a.nz - 1 - step; };
  auto in_plane = [&](uint32_t xy) { return (int64_t)(xy & 0xffff) * a.cs_x + (int64_t)(xy >> 16) * a.cs_y; };
  auto in_ghost = [&](uint32_t xy) { return (int64_t)(xy & 0xffff) * a.gs_x + (int64_t)(xy >> 16) * a.gs_y; };
  // a cell's coordinates made opaque at each use: everything derived from them (the
  // stencil's LDS addresses, the cube offsets) is then recomputed per plane, a few
  // integer ops, instead of being hoisted out of the plane loop into registers held
  // for every owned cell
  auto opaque = [](uint32_t v) {
    asm volatile("" : "+v"(v));
    return v;
  };

  // per owned cell: (y << 16 | x), and the intensity cotangent gathered from the
  // plane downstream; the chi and S cotangents sent upwind go straight into the
  // global sums (the same thread adds the cell's own share one plane later, so
  // both read-modify-writes of a cell are in one thread's program order)
  uint32_t xy[MAXC];
  float acc[MAXC];
#pragma unroll
  for (int k = 0; k < MAXC; ++k) {
    const int c = tid + k * nt;
    const bool ok = FULL || c < nxy;
    xy[k] = ok ? ((uint32_t)(c / nx) << 16) | (uint32_t)(c % nx) : 0xffffffffu;
    acc[k] = (ok && a.exitbar) ? a.exitbar[f * a.gs_f + in_ghost(xy[k])] : 0.0f;
  }
  auto valid = [&](int k) { return FULL || xy[k] != 0xffffffffu; };

  auto plane = [&](int step, auto ghost_c) {
    constexpr bool GHOST = decltype(ghost_c)::value;
    const int z = plane_of(step);
    const int zu = GHOST ? 0 : plane_of(step - 1);
    const float *rcu = GHOST ? ghost_rc : a.rc + (int64_t)zu * nxy;
    const int64_t bz = (int64_t)z * a.cs_z + (int64_t)f * a.cs_f;
    const int64_t bu = (int64_t)zu * a.cs_z + (int64_t)f * a.cs_f;
    // (a) the upwind plane
#pragma unroll
    for (int k = 0; k < MAXC; ++k) {
      if (valid(k)) {
        const uint32_t q = opaque(xy[k]);
        const int c = (int)(q >> 16) * nx + (int)(q & 0xffff);
        float iv, cv, sv;
        if (GHOST) {
          const int64_t g = in_ghost(q);
          cv = gchi[g];
          sv = gS[g];
          iv = gI[g];
        } else {
          const int64_t o = bu + in_plane(q);
          cv = a.chi[o];
          sv = a.S[o];
          iv = a.I[o];
        }
        sI[c] = iv;
        sX[c] = cv * rcu[c];
        sS[c] = sv;
      }
    }
    dev::lds_barrier();
    // (b) the targets
    float gi[MAXC], gx[MAXC], gs[MAXC];
    uint64_t codes = 0;
    const float *dtp = a.dt500 + (int64_t)z * nxy, *rcp = a.rc + (int64_t)z * nxy;
#pragma unroll
    for (int k = 0; k < MAXC; ++k) {
      gi[k] = gx[k] = gs[k] = 0.0f;
      if (valid(k)) {
        const uint32_t q = opaque(xy[k]);
        const int x = (int)(q & 0xffff), y = (int)(q >> 16), c = y * nx + x;
        int xs[4], ys[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          xs[i] = dev::wrap_near(x + a.xo + i, nx);
          ys[i] = dev::wrap_near(y + a.yo + i, ny);
        }
        // the forward's upwind values (row reductions in x, then y, as the forward)
        float rI[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float *row = sI + ys[j] * nx;
          rI[j] = a.kx[0] * row[xs[0]] + a.kx[1] * row[xs[1]] + a.kx[2] * row[xs[2]] + a.kx[3] * row[xs[3]];
        }
        const float iraw = a.ky[0] * rI[0] + a.ky[1] * rI[1] + a.ky[2] * rI[2] + a.ky[3] * rI[3];
        const int r1 = ys[1] * nx, r2 = ys[2] * nx;
        const float xr1 = a.wx[0] * sX[r1 + xs[1]] + a.wx[1] * sX[r1 + xs[2]];
        const float xr2 = a.wx[0] * sX[r2 + xs[1]] + a.wx[1] * sX[r2 + xs[2]];
        const float sr1 = a.wx[0] * sS[r1 + xs[1]] + a.wx[1] * sS[r1 + xs[2]];
        const float sr2 = a.wx[0] * sS[r2 + xs[1]] + a.wx[1] * sS[r2 + xs[2]];
        const float xxup = a.wy[0] * xr1 + a.wy[1] * xr2;
        const float ssup = a.wy[0] * sr1 + a.wy[1] * sr2;
        // the last smallest of (x2,y2), (x3,y2), (x2,y3), (x3,y3)
        const float v0 = sI[r1 + xs[1]], v1 = sI[r1 + xs[2]], v2 = sI[r2 + xs[1]], v3 = sI[r2 + xs[2]];
        float mn = v0;
        int arg = 0;
        if (!(mn < v1)) { mn = v1; arg = 1; }
        if (!(mn < v2)) { mn = v2; arg = 2; }
        if (!(mn < v3)) { mn = v3; arg = 3; }
        const bool clamp = !(iraw > mn);
        const float Iup = clamp ? mn : iraw;
        const int64_t o = bz + in_plane(q);
        const float chv = a.chi[o], ss = a.S[o];
        const float dt500 = dtp[c], rc = rcp[c];
        const float xx = chv * rc;
        const float dtr = 0.5f * (xxup + xx) * dt500;
        const float dt = fmaxf(dtr, a.mdtau);
        float E, W, Ps, dE, dW, dPs;
        lin_weights_deriv(dt, E, W, Ps, dE, dW, dPs);
        const float Inu = Iup * E + W * ssup + Ps * ss;
        float g = acc[k];
        if (SEED) {
          if (a.Jbar) g += a.wj * a.Jbar[o];
          if (a.Ibar) g += a.Ibar[o];
        }
        g = (Inu >= -a.imax && Inu <= a.imax) ? g : 0.0f;
        const float gdt = dtr > a.mdtau ? g * (Iup * dE + ssup * dW + ss * dPs) : 0.0f;
        const float gxx = 0.5f * gdt * dt500;
        a.chibar[o] = a.chibar[o] + gxx * rc;
        a.Sbar[o] = a.Sbar[o] + g * Ps;
        gi[k] = g * E;
        gx[k] = gxx;
        gs[k] = g * W;
        codes |= (uint64_t)(clamp ? arg + 1 : 0) << (3 * k);
      }
    }
    dev::lds_barrier();
    // (c) the targets' upwind cotangents replace the upwind plane
#pragma unroll
    for (int k = 0; k < MAXC; ++k) {
      if (valid(k)) {
        const uint32_t q = opaque(xy[k]);
        const int c = (int)(q >> 16) * nx + (int)(q & 0xffff);
        sI[c] = gi[k];
        sX[c] = gx[k];
        sS[c] = gs[k];
        sC[c] = (uint8_t)((codes >> (3 * k)) & 7);
      }
    }
    dev::lds_barrier();
    // (d) every upwind cell gathers what the targets whose stencil holds it sent:
    // target column x - xo - i sees this column in stencil slot i
#pragma unroll
    for (int k = 0; k < MAXC; ++k) {
      if (valid(k)) {
        const uint32_t q = opaque(xy[k]);
        const int x = (int)(q & 0xffff), y = (int)(q >> 16), c = y * nx + x;
        int tx[4], ty[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          tx[i] = dev::wrap_near(x - a.xo - i, nx);
          ty[i] = dev::wrap_near(y - a.yo - i, ny);
        }
        float ib = 0.0f, xb = 0.0f, sb = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int row = ty[j] * nx;
          float ri = 0.0f;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int t = row + tx[i];
            const float v = sI[t];
            const int cd = sC[t];
            ri += cd == 0 ? a.kx[i] * v : 0.0f;
            if (i >= 1 && i <= 2 && j >= 1 && j <= 2) ib += cd == (i - 1) + 2 * (j - 1) + 1 ? v : 0.0f;
          }
          ib += a.ky[j] * ri;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int row = ty[j + 1] * nx;
          xb += a.wy[j] * (a.wx[0] * sX[row + tx[1]] + a.wx[1] * sX[row + tx[2]]);
          sb += a.wy[j] * (a.wx[0] * sS[row + tx[1]] + a.wx[1] * sS[row + tx[2]]);
        }
        if (GHOST) {
          const int64_t g = f * a.gs_f + in_ghost(q);
          if (a.gIbar) a.gIbar[g] = a.gIbar[g] + ib;
          if (a.gchibar) a.gchibar[g] = a.gchibar[g] + xb * rcu[c];
          if (a.gSbar) a.gSbar[g] = a.gSbar[g] + sb;
        } else {
          const int64_t o = bu + in_plane(q);
          acc[k] = ib;
          a.chibar[o] = a.chibar[o] + xb * rcu[c];
          a.Sbar[o] = a.Sbar[o] + sb;
        }
      }
    }
    dev::lds_barrier();
  };

  for (int step = a.nz - 1; step >= 1; --step) plane(step, std::false_type{});
  plane(0, std::true_type{});
}

// The same reverse sweep with the forward's strip mapping (every strip exactly L
// rows: ny = L * strips, nx * strips threads).  Thread (x, strip) owns the cells of
// column x in its L rows; its stencil columns are fixed, and down the strip the
// 4-row stencils of consecutive cells share three rows, so in (b) each target reads
// ONE new upwind row (4 I, 2 chi/chi500, 2 S) and in (d) each upwind cell ONE new
// target row (4 intensity cotangents and clamp codes, 2 + 2 chi/S cotangents), the
// row reductions sliding over a 4-deep register window.  Global accesses of a
// wavefront are rows of consecutive x.
//
// The upwind cells of one step are the targets of the next, and a thread owns the
// same (x, y) in both: the chi and S shares (d) gathers for a cell stay in registers
// with its intensity cotangent and are added to the cell's own in the next step's
// (b), so each cell's chi and S cotangents take one read-modify-write per ray, and
// (d) touches no global memory between ghost steps.  (b) keeps the global reads of
// the next two cells in flight while a cell is computed (each cell's read-modify-
// write is ordered after the previous cell's, which the compiler cannot reorder:
// the cotangent arrays may alias the inputs as far as it knows).  Arithmetic as
// sc2_adjoint's, except that a cell's two chi shares are summed before the one
// factor 1/chi500 they both carry.
template <int L, bool SEED>
__global__ void __launch_bounds__(L >= 8 ? 768 : 1024) sc2_adjoint_strip(const AdjArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nx = a.nx, ny = a.ny, nxy = nx * ny;
  float *sI = reinterpret_cast<float *>(smem_raw);
  float *sX = sI + nxy;
  float *sS = sX + nxy;
  uint8_t *sC = reinterpret_cast<uint8_t *>(sS + nxy);
  const int f = blockIdx.x, tid = threadIdx.x;
  const int x = tid % nx, ybeg = (tid / nx) * L;
  const float *gchi = a.gchi + f * a.gs_f, *gS = a.gS + f * a.gs_f, *gI = a.gI + f * a.gs_f;
  const float *ghost_rc = a.rc + (int64_t)a.nz * nxy;
  auto plane_of = [&](int step) { return a.dir > 0 ? step : a.nz - 1 - step; };
  // byte steps of one row in a cube plane, in the ghost plane, in a per-cell plane
  const uint32_t rowb = (uint32_t)a.cs_y * 4u, growb = (uint32_t)a.gs_y * 4u, crowb = (uint32_t)nx * 4u;
  int xs[4], tx[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    xs[i] = dev::wrap_near(x + a.xo + i, nx);
    tx[i] = dev::wrap_near(x - a.xo - i, nx);
  }
  const int64_t ox = (int64_t)x * a.cs_x, gox = (int64_t)x * a.gs_x;
  // per owned cell, carried from one step's (d) to the next step's (b): the
  // intensity cotangent and the chi (before its 1/chi500) and S shares sent upwind
  float acc[L], ub[L], vb[L];
  {
    const uint32_t go0 = (uint32_t)(gox + (int64_t)ybeg * a.gs_y) * 4u;
    const float *eb = a.exitbar ? a.exitbar + f * a.gs_f : nullptr;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      acc[k] = eb ? ldb(eb, go0 + (uint32_t)k * growb) : 0.0f;
      ub[k] = 0.0f;
      vb[k] = 0.0f;
    }
  }

  struct URow { float rI, i1, i2, rX, rS; };
  struct TRow { float rG, g1, g2, rX, rS; int c1, c2; };
  struct TIn { float chv, ss, dt500, rc, cb, sb, jb, ib; };

  auto plane = [&](int step, auto ghost_c) {
    constexpr bool GHOST = decltype(ghost_c)::value;
    int yb = ybeg;
    asm volatile("" : "+v"(yb));
    const int z = plane_of(step);
    const int zu = GHOST ? 0 : plane_of(step - 1);
    const float *rcu = GHOST ? ghost_rc : a.rc + (int64_t)zu * nxy;
    // wave-uniform bases of the target plane z and the upwind plane zu, and this
    // thread's byte offsets of its first row (in a cube plane, the ghost plane, a
    // per-cell plane)
    const int64_t pz = (int64_t)z * a.cs_z + (int64_t)f * a.cs_f, pu = (int64_t)zu * a.cs_z + (int64_t)f * a.cs_f;
    const uint32_t vo = (uint32_t)(ox + (int64_t)yb * a.cs_y) * 4u;
    const uint32_t go = (uint32_t)(gox + (int64_t)yb * a.gs_y) * 4u;
    const uint32_t co = (uint32_t)(yb * nx + x) * 4u;
    // (a) the upwind plane
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int y = yb + k, c = y * nx + x;
      float iv, cv, sv;
      if (GHOST) {
        const uint32_t g = go + (uint32_t)k * growb;
        cv = ldb(gchi, g);
        sv = ldb(gS, g);
        iv = ldb(gI, g);
      } else {
        const uint32_t o = vo + (uint32_t)k * rowb;
        cv = ldb(a.chi + pu, o);
        sv = ldb(a.S + pu, o);
        iv = ldb(a.I + pu, o);
      }
      sI[c] = iv;
      sX[c] = cv * ldb(rcu, co + (uint32_t)k * crowb);
      sS[c] = sv;
    }
    dev::lds_barrier();
    // (b) the targets, the upwind rows sliding down the strip
    asm volatile("" : "+v"(yb));
    auto urow = [&](int yr) -> URow {
      URow r;
      const int base = dev::wrap_near(yr, ny) * nx;
      const float *row = sI + base;
      const float i1 = row[xs[1]], i2 = row[xs[2]];
      r.rI = a.kx[0] * row[xs[0]] + a.kx[1] * i1 + a.kx[2] * i2 + a.kx[3] * row[xs[3]];
      r.i1 = i1;
      r.i2 = i2;
      r.rX = a.wx[0] * sX[base + xs[1]] + a.wx[1] * sX[base + xs[2]];
      r.rS = a.wx[0] * sS[base + xs[1]] + a.wx[1] * sS[base + xs[2]];
      return r;
    };
    const float *dtp = a.dt500 + (int64_t)z * nxy, *rcp = a.rc + (int64_t)z * nxy;
    auto tload = [&](int k) -> TIn {
      TIn t;
      const uint32_t o = vo + (uint32_t)k * rowb, oc = co + (uint32_t)k * crowb;
      t.chv = ldb(a.chi + pz, o);
      t.ss = ldb(a.S + pz, o);
      t.dt500 = ldb(dtp, oc);
      t.rc = ldb(rcp, oc);
      t.cb = ldb(a.chibar + pz, o);
      t.sb = ldb(a.Sbar + pz, o);
      t.jb = (SEED && a.Jbar) ? ldb(a.Jbar + pz, o) : 0.0f;
      t.ib = (SEED && a.Ibar) ? ldb(a.Ibar + pz, o) : 0.0f;
      return t;
    };
    float gi[L], gx[L], gs[L];
    uint64_t codes = 0;
    TIn q0 = tload(0), q1 = tload(L > 1 ? 1 : 0);
    URow w0 = urow(yb + a.yo), w1 = urow(yb + a.yo + 1), w2 = urow(yb + a.yo + 2);
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int y = yb + k;
      const TIn t = q0;
      q0 = q1;
      if (k + 2 < L) q1 = tload(k + 2);
      const URow w3 = urow(y + a.yo + 3);
      const float iraw = a.ky[0] * w0.rI + a.ky[1] * w1.rI + a.ky[2] * w2.rI + a.ky[3] * w3.rI;
      const float xxup = a.wy[0] * w1.rX + a.wy[1] * w2.rX;
      const float ssup = a.wy[0] * w1.rS + a.wy[1] * w2.rS;
      float mn = w1.i1;
      int arg = 0;
      if (!(mn < w1.i2)) { mn = w1.i2; arg = 1; }
      if (!(mn < w2.i1)) { mn = w2.i1; arg = 2; }
      if (!(mn < w2.i2)) { mn = w2.i2; arg = 3; }
      const bool clamp = !(iraw > mn);
      const float Iup = clamp ? mn : iraw;
      const float xx = t.chv * t.rc;
      const float dtr = 0.5f * (xxup + xx) * t.dt500;
      const float dt = fmaxf(dtr, a.mdtau);
      float E, W, Ps, dE, dW, dPs;
      lin_weights_deriv(dt, E, W, Ps, dE, dW, dPs);
      const float Inu = Iup * E + W * ssup + Ps * t.ss;
      float g = acc[k];
      if (SEED) g += a.wj * t.jb + t.ib;
      g = (Inu >= -a.imax && Inu <= a.imax) ? g : 0.0f;
      const float gdt = dtr > a.mdtau ? g * (Iup * dE + ssup * dW + t.ss * dPs) : 0.0f;
      const float gxx = 0.5f * gdt * t.dt500;
      const uint32_t o = vo + (uint32_t)k * rowb;
      stb(a.chibar + pz, o, t.cb + (ub[k] + gxx) * t.rc);
      stb(a.Sbar + pz, o, t.sb + (vb[k] + g * Ps));
      gi[k] = g * E;
      gx[k] = gxx;
      gs[k] = g * W;
      codes |= (uint64_t)(clamp ? arg + 1 : 0) << (3 * k);
      w0 = w1;
      w1 = w2;
      w2 = w3;
    }
    dev::lds_barrier();
    // (c) the targets' upwind cotangents replace the upwind plane
    asm volatile("" : "+v"(yb));
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int c = (yb + k) * nx + x;
      sI[c] = gi[k];
      sX[c] = gx[k];
      sS[c] = gs[k];
      sC[c] = (uint8_t)((codes >> (3 * k)) & 7);
    }
    dev::lds_barrier();
    // (d) the gather, the target rows sliding down the strip: target row
    // y - yo - j sees this row in stencil row j
    asm volatile("" : "+v"(yb));
    auto trow = [&](int yr) -> TRow {
      TRow r;
      const int base = dev::wrap_near(yr, ny) * nx;
      float ri = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float v = sI[base + tx[i]];
        const int cd = sC[base + tx[i]];
        ri += cd == 0 ? a.kx[i] * v : 0.0f;
        if (i == 1) { r.g1 = v; r.c1 = cd; }
        if (i == 2) { r.g2 = v; r.c2 = cd; }
      }
      r.rG = ri;
      r.rX = a.wx[0] * sX[base + tx[1]] + a.wx[1] * sX[base + tx[2]];
      r.rS = a.wx[0] * sS[base + tx[1]] + a.wx[1] * sS[base + tx[2]];
      return r;
    };
    TRow t1 = trow(yb - a.yo - 1), t2 = trow(yb - a.yo - 2), t3 = trow(yb - a.yo - 3);
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int y = yb + k;
      const TRow t0 = trow(y - a.yo);
      float ib = 0.0f;
      ib += a.ky[0] * t0.rG;
      ib += t1.c1 == 1 ? t1.g1 : 0.0f;
      ib += t1.c2 == 2 ? t1.g2 : 0.0f;
      ib += a.ky[1] * t1.rG;
      ib += t2.c1 == 3 ? t2.g1 : 0.0f;
      ib += t2.c2 == 4 ? t2.g2 : 0.0f;
      ib += a.ky[2] * t2.rG;
      ib += a.ky[3] * t3.rG;
      float xb = 0.0f, sb = 0.0f;
      xb += a.wy[0] * t1.rX;
      sb += a.wy[0] * t1.rS;
      xb += a.wy[1] * t2.rX;
      sb += a.wy[1] * t2.rS;
      if (GHOST) {
        const uint32_t g = go + (uint32_t)k * growb;
        const int64_t gf = f * a.gs_f;
        if (a.gIbar) stb(a.gIbar + gf, g, ldb(a.gIbar + gf, g) + ib);
        if (a.gchibar) stb(a.gchibar + gf, g, ldb(a.gchibar + gf, g) + xb * ldb(rcu, co + (uint32_t)k * crowb));
        if (a.gSbar) stb(a.gSbar + gf, g, ldb(a.gSbar + gf, g) + sb);
      } else {
        acc[k] = ib;
        ub[k] = xb;
        vb[k] = sb;
      }
      t3 = t2;
      t2 = t1;
      t1 = t0;
    }
    dev::lds_barrier();
  };

#pragma nounroll
  for (int step = a.nz - 1; step >= 1; --step) plane(step, std::false_type{});
  plane(0, std::true_type{});
}

template <int L, bool SEED>
int launch_strip(m3drt_ctx *ctx, const AdjArgs &a, int nt, size_t lds) {
  auto kern = sc2_adjoint_strip<L, SEED>;
  M3DRT_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(a.nf), dim3(nt), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

template <int L>
int strip_seed(m3drt_ctx *ctx, const AdjArgs &a, int nt, size_t lds) {
  return (a.Jbar || a.Ibar) ? launch_strip<L, true>(ctx, a, nt, lds) : launch_strip<L, false>(ctx, a, nt, lds);
}

// strip lengths instantiated; 8 and 12 are built for 768 threads (163 VGPRs at 12)
static const int kStripL[] = {4, 6, 8, 12};
static int strip_max_threads(int L) { return L >= 8 ? 768 : 1024; }

template <int MAXC, bool FULL, bool SEED>
int launch_adj(m3drt_ctx *ctx, const AdjArgs &a, int nt, size_t lds) {
  auto kern = sc2_adjoint<MAXC, FULL, SEED>;
  M3DRT_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(a.nf), dim3(nt), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

template <int MAXC>
int launch_full(m3drt_ctx *ctx, const AdjArgs &a, int nt, size_t lds, bool full) {
  const bool seed = a.Jbar || a.Ibar;
  if (full) return seed ? launch_adj<MAXC, true, true>(ctx, a, nt, lds) : launch_adj<MAXC, true, false>(ctx, a, nt, lds);
  return seed ? launch_adj<MAXC, false, true>(ctx, a, nt, lds) : launch_adj<MAXC, false, false>(ctx, a, nt, lds);
}

}  // namespace

int sweep_adjoint(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_adjoint_bufs *b) {
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  const int64_t nxy = (int64_t)nx * ny;
  const size_t lds = (size_t)nxy * (3 * sizeof(float) + 1);
  if (lds > (size_t)ctx->lds_max)
    return set_error(M3DRT_ELIMIT, "m3drt_sc_adjoint: three FP32 planes and the clamp bytes exceed the LDS");
  // a workgroup size that divides the plane exactly when one does (every thread then
  // owns the same number of cells), else the fewest cells per thread
  int nt = (int)std::min<int64_t>(1024, (nxy + 63) / 64 * 64);
  for (int t = 1024; t >= 256; t -= 64)
    if (nxy % t == 0 && nxy / t <= 13) {
      nt = t;
      break;
    }
  const int64_t maxc = (nxy + nt - 1) / nt;
  float *dt500 = nullptr, *rc = nullptr;
  int st = sc2_prep_f32(ctx, d, b->chi500, b->chi500_up, b->g500, &dt500, &rc);
  if (st) return st;
  const Sc2Geom g = sc2_geometry(d);
  AdjArgs a;
  a.chi = static_cast<const float *>(b->chi);
  a.S = static_cast<const float *>(b->S);
  a.I = static_cast<const float *>(b->I);
  a.gchi = static_cast<const float *>(b->ghost_chi);
  a.gS = static_cast<const float *>(b->ghost_S);
  a.gI = static_cast<const float *>(b->ghost_I);
  a.dt500 = dt500;
  a.rc = rc;
  a.Jbar = static_cast<const float *>(b->Jbar);
  a.Ibar = static_cast<const float *>(b->Ibar);
  a.exitbar = static_cast<const float *>(b->exit_bar);
  a.chibar = static_cast<float *>(b->chibar);
  a.Sbar = static_cast<float *>(b->Sbar);
  a.gchibar = static_cast<float *>(b->ghost_chibar);
  a.gSbar = static_cast<float *>(b->ghost_Sbar);
  a.gIbar = static_cast<float *>(b->ghost_Ibar);
  a.cs_f = d->cs_f; a.cs_x = d->cs_x; a.cs_y = d->cs_y; a.cs_z = d->cs_z;
  a.gs_f = d->gs_f; a.gs_x = d->gs_x; a.gs_y = d->gs_y;
  a.nf = d->nf; a.nx = nx; a.ny = ny; a.nz = nz; a.dir = d->dir > 0 ? 1 : -1;
  a.xo = g.xo; a.yo = g.yo;
  for (int i = 0; i < 4; ++i) {
    a.kx[i] = (float)g.kx[i];
    a.ky[i] = (float)g.ky[i];
  }
  // inner_fraction (SRC/rt_solvers_mod.f90:500-506), in FP32 as the forward forms it
  const float hx = -a.kx[0] + a.kx[2] + 2.0f * a.kx[3], hy = -a.ky[0] + a.ky[2] + 2.0f * a.ky[3];
  a.wx[0] = 1.0f - hx; a.wx[1] = hx;
  a.wy[0] = 1.0f - hy; a.wy[1] = hy;
  a.mdtau = -(float)d->maser_dtau_max;
  a.imax = (float)d->maser_Inu_max;
  a.wj = (float)d->wj;
  // the strip form where an instantiated strip length divides ny and the strips fit
  // one workgroup, and where a plane of the cubes and of the ghost planes spans less
  // than 4 GiB (the strip kernel addresses within a plane by 32-bit byte offsets);
  // the cell-owner form otherwise
  const auto span32 = [&](int64_t sx, int64_t sy) {
    return ((int64_t)(nx - 1) * sx + (int64_t)(ny - 1) * sy + 1) * 4 <= (int64_t)UINT32_MAX;
  };
  if (!getenv("M3DRT_ADJ_CELLS") && span32(a.cs_x, a.cs_y) && span32(a.gs_x, a.gs_y)) {
    for (int Ls : kStripL) {
      if (ny % Ls || (int64_t)nx * (ny / Ls) > strip_max_threads(Ls)) continue;
      const int nts = nx * (ny / Ls);
      if (Ls == 4) return strip_seed<4>(ctx, a, nts, lds);
      if (Ls == 6) return strip_seed<6>(ctx, a, nts, lds);
      if (Ls == 8) return strip_seed<8>(ctx, a, nts, lds);
      return strip_seed<12>(ctx, a, nts, lds);
    }
  }
  // the exact-divisor specialisation only where it stays within the register file
  const auto full = [&](int m) { return m <= 4 && nxy == (int64_t)m * nt; };
  if (maxc <= 1) return launch_full<1>(ctx, a, nt, lds, full(1));
  if (maxc <= 2) return launch_full<2>(ctx, a, nt, lds, full(2));
  if (maxc <= 4) return launch_full<4>(ctx, a, nt, lds, full(4));
  if (maxc <= 6) return launch_full<6>(ctx, a, nt, lds, full(6));
  if (maxc <= 9) return launch_full<9>(ctx, a, nt, lds, full(9));
  if (maxc <= 13) return launch_full<13>(ctx, a, nt, lds, full(13));
  return set_error(M3DRT_ELIMIT, "m3drt_sc_adjoint: plane too large");
}

}  // namespace m3drt
