// Straight-line form of the 3D short-characteristics sweep (short_char2,
// SRC/rt_solvers_mod.f90:224-507, with integrate_intensity,
// SRC/rt_patch_mod.f90:3399-3431, fused) for the case the formal solution of an
// iteration runs: J, or J and PsiJ, or only the exit plane as the outputs, no eta flags, no split
// rates.  Same mapping and arithmetic as sweep_sc2.hip's general kernel (one
// workgroup per frequency owning the whole periodic plane, strips of L rows per
// thread, a sliding 4-row window, PD cells of global loads in flight), with the
// address work taken out of the cell loop:
//
//  * the upwind plane sits in LDS with a periodic halo of 2 cells on every side
//    ((nx+4) x (ny+4)), so a thread's whole 4-column x (L+3)-row stencil reach is
//    ONE base address plus row steps, with no wrap per cell;
//  * I and x = chi/chi500 are interleaved as pairs, so a stencil row is two
//    64-bit pair reads (ds_read2_b64) instead of six 32-bit ones;
//  * the cubes are read at a per-thread 32-bit offset that never changes from
//    plane to plane against a wave-uniform base per (plane, row);
//  * dt500 and 1/chi500 of a cell come as one pair.
// The strips are exact (ny = L * strips, L from the instantiated set), so every
// per-cell branch is resolved at compile time.
#include "sweep_sc2.hpp"
#include <cstdio>
#include <cstdlib>
#include <algorithm>

namespace m3drt {
namespace {

template <typename T>
struct alignas(2 * sizeof(T)) Pair {
  T a, b;
};

template <typename T>
struct FastArgs {
  const T *chi, *S;
  const T *gchi, *gS, *gI;
  const Pair<T> *dr;  // per cell (dt500, 1/chi500), planes 0..nz-1, then the ghost plane (0, 1/g500)
  T *J, *PsiJ, *exitI;
  T *Iout, *Psiout, *Irem, *esc;  // OM 1: the ALI's split outputs instead of J; OM 2: Iout alone
  int64_t cs_f, cs_z, gs_f;
  uint32_t cs_x, cs_y, gs_x, gs_y;
  int nx, ny, nz, dir, xo, yo, PW, nstrips;
  uint32_t plane_bytes, gplane_bytes;  // byte reach of one (z, f) plane of a cube / the ghost plane
  T xk0, xk1, xk2, xk3, yk0, yk1, yk2, yk3, hx, hy, mdtau, imax, wj;
};

template <typename T>
__device__ __forceinline__ T ldo(const T *base, uint32_t off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

// Buffer access: a wave-uniform descriptor (base, size in bytes), a per-lane 32-bit
// offset and a wave-uniform 32-bit offset, so a cell's address costs no vector
// instruction.  Accesses beyond the size return zero / are dropped.
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bld(Rsrc r, uint32_t v, uint32_t s) {
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
  else if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
}
template <typename T>
__device__ __forceinline__ void bst(Rsrc r, uint32_t v, uint32_t s, T x) {
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), r, v, s, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b64(
        __builtin_bit_cast(unsigned int __attribute__((ext_vector_type(2))), x), r, v, s, 0);
}

// (dt500, 1/chi500) per cell as sc2_prep in sweep_sc2.hip forms them
// (SRC/rt_solvers_mod.f90:355-370), then (0, 1/g500) for the ghost plane.
template <typename T>
__global__ void __launch_bounds__(256) sc2_prep_pairs(const float *__restrict__ chi500,
                                                      const float *__restrict__ chi500_up,
                                                      const float *__restrict__ g500, Pair<T> *__restrict__ dr,
                                                      int nxy, int nz, double dsup) {
  const int64_t n = (int64_t)nxy * nz;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n + nxy;
       i += (int64_t)gridDim.x * blockDim.x) {
    Pair<T> p;
    if (i < n) {
      const double c = (double)chi500[i];
      const double cu = (double)chi500_up[i];
      const double r = cu / c;
      double d;
      if (fabs(r - 1.0) < 1e-3)
        d = dsup * c * (1.0 + (r - 1.0) * (0.5 - (r - 1.0) / 12.0));
      else
        d = dsup * c * (r - 1.0) / log(r);
      p.a = (T)d;
      p.b = (T)(1.0 / c);
    } else {
      p.a = (T)0;
      p.b = (T)(1.0 / (double)g500[i - n]);
    }
    dr[i] = p;
  }
}

template <typename T, int N>
struct alignas(N * sizeof(T) >= 16 ? 16 : N * sizeof(T)) Arr {
  T v[N];
};

// MODE 2: (I, x) pairs and S in LDS; 1: (I, x) pairs in LDS, upwind S from the
// cache; 4: (I, x) pairs in LDS and, for planes where a whole S plane no longer
// fits beside them, the S of the first SR of every strip's L+1 tap rows in LDS
// (SR rows of nx per strip, then a spare row), the other tap rows from the cache.
// Every strip takes the same share from the cache: the workgroup waits at each
// plane for its slowest wave, so holding whole strips in LDS and leaving others
// on the cache gains nothing.  CPL: adjacent columns per lane (1, or 2 with the native layout: the
// cube rows are then read and written 8 bytes per lane and the (dt500, 1/chi500)
// pairs 16, so a pair of cells costs the vector-memory instructions of one).
// threads of a workgroup with strip length L (fewer threads per workgroup leave
// more registers per lane: 168 at 768 threads, 256 at 512)
template <int L>
constexpr int fast_max_threads() { return L >= 30 ? 512 : L >= 20 ? 768 : 1024; }

// OM (outputs): 0, J (and PsiJ) and the exit plane; 1, no J: every cell's I, Psi
// (= Lambda*/chi), Irem and escape written instead, as the general kernel's split-rate
// form writes them for a cell without eta flags (SRC/rt_solvers_mod.f90:419-440,
// 465-475; the maser limit on |I| before the clamp); 2, no J: every cell's I (the
// intensity cube an adjoint sweep recomputes).
template <int L, bool PJ, int MODE, int CPL, int SR = 0, int OM = 0>
__global__ void __launch_bounds__(fast_max_threads<L>()) sc2_fast(const FastArgs<float> a) {
  using namespace dev;
  using T = float;
  using V = Arr<T, CPL>;
  using DV = Arr<Pair<T>, CPL>;
  // rows of loads in flight ahead of the row being computed (two cells' loads per
  // step with CPL = 2).  4 measured best for strips of 15 and 20 rows on the
  // 120 x 120 plane (13.3 ms against 13.9 with 3 and 14.1 with 6 for strips of 20).
  constexpr int PD = CPL == 1 ? (L >= 15 ? 4 : 3) : 1;
  constexpr uint32_t TB = sizeof(T);
  static_assert(MODE != 4 || (CPL == 1 && SR >= 1 && SR <= L), "S rows in LDS: one column per lane, SR <= L");
  constexpr bool SPL = OM == 1, IOUT = OM >= 1;
  static_assert(OM == 0 || (CPL == 1 && !PJ && MODE != 4), "cell outputs: one column per lane, no PsiJ, S rows whole");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nx = a.nx, ny = a.ny, PW = a.PW, nxy = nx * ny;
  const int np = PW * (ny + 4);
  Pair<T> *sP = reinterpret_cast<Pair<T> *>(smem_raw);
  T *sS = reinterpret_cast<T *>(smem_raw + (size_t)np * sizeof(Pair<T>));

  const int f = blockIdx.x;
  const int tid = threadIdx.x;
  const int nxp = nx / CPL;
  const int x0 = (tid % nxp) * CPL;
  const int strip = tid / nxp;
  const int yb = strip * L;
  const int rb = (yb + a.yo + 2) * PW + (x0 + a.xo + 2);  // row 0 of the window, padded index
  int xc[CPL + 1];                                        // upwind columns of S (wrapped)
#pragma unroll
  for (int c = 0; c <= CPL; ++c) xc[c] = wrap(x0 + a.xo + 1 + c, nx);
  const int g0 = wrap(yb + a.yo, ny);                     // row 0 of the window, unpadded
  const uint32_t voff = (x0 * a.cs_x + yb * a.cs_y) * TB; // first cell of the strip in a cube plane
  const uint32_t doff = (yb * nx + x0) * (uint32_t)sizeof(Pair<T>);
  const T ohx = (T)1 - a.hx, ohy = (T)1 - a.hy;

  const T *__restrict__ gchi = a.gchi + f * a.gs_f;
  const T *__restrict__ gS = a.gS + f * a.gs_f;
  const T *__restrict__ gI = a.gI + f * a.gs_f;
  const Pair<T> *__restrict__ gdr = a.dr + (int64_t)a.nz * nxy;

  struct In {
    V chi, S, J, PsJ;
    DV dr;
  };
  // descriptors of one plane of every cube
  struct Planes {
    Rsrc chi, S, J, PsJ, dr, Io, Po, Ro, Eo;
  };
  auto plane_of = [&](int step) { return a.dir > 0 ? step : a.nz - 1 - step; };
  auto planes = [&](int z) -> Planes {
    Planes P;
    const int64_t o = (int64_t)z * a.cs_z + (int64_t)f * a.cs_f;
    P.chi = rsrc(a.chi + o, a.plane_bytes);
    P.S = rsrc(a.S + o, a.plane_bytes);
    P.J = a.J ? rsrc(a.J + o, a.plane_bytes) : rsrc(a.chi + o, 0u);
    P.PsJ = PJ ? rsrc(a.PsiJ + o, a.plane_bytes) : P.J;
    P.dr = rsrc(a.dr + (int64_t)z * nxy, (uint32_t)(nxy * sizeof(Pair<T>)));
    if (IOUT) P.Io = rsrc(a.Iout + o, a.plane_bytes);
    if (SPL) {
      P.Po = rsrc(a.Psiout + o, a.plane_bytes);
      P.Ro = rsrc(a.Irem + o, a.plane_bytes);
      P.Eo = rsrc(a.esc + o, a.plane_bytes);
    }
    return P;
  };
  const uint32_t rowb = a.cs_y * TB, drowb = nx * (uint32_t)sizeof(Pair<T>);
  auto fetch = [&](const Planes &P, int k) -> In {
    In v;
    v.chi = bld<V>(P.chi, voff, k * rowb);
    v.S = bld<V>(P.S, voff, k * rowb);
    if (!IOUT) v.J = bld<V>(P.J, voff, k * rowb);
    if (PJ) v.PsJ = bld<V>(P.PsJ, voff, k * rowb);
    v.dr = bld<DV>(P.dr, doff, k * drowb);
    return v;
  };

  In pre[PD], ring[PD];
  {
    const Planes P0 = planes(plane_of(0));
#pragma unroll
    for (int k = 0; k < PD; ++k) pre[k] = fetch(P0, k);
  }
  // the upstream plane, halo included
  for (int c = tid; c < np; c += blockDim.x) {
    const int pr = c / PW, pc = c - pr * PW;
    const int cy = wrap(pr - 2, ny), cx = wrap(pc - 2, nx);
    const uint32_t go = (cx * a.gs_x + cy * a.gs_y) * TB;
    Pair<T> p;
    p.a = ldo(gI, go);
    p.b = ldo(gchi, go) * gdr[cy * nx + cx].b;
    sP[c] = p;
    if (MODE == 2) sS[c] = ldo(gS, go);
  }
  if (MODE == 4) {
    // slot s*SR + r-1 holds window row r (row s*L+yo+r) of strip s, r = 1..SR
    for (int c = tid; c < a.nstrips * SR * nx; c += blockDim.x) {
      const int q = c / nx, cx = c - q * nx;
      const int sq = q / SR;
      const int cy = wrap(sq * L + a.yo + 1 + (q - sq * SR), ny);
      sS[c] = ldo(gS, (cx * a.gs_x + cy * a.gs_y) * TB);
    }
  }
  lds_barrier();

  for (int step = 0; step < a.nz; ++step) {
    const int z = plane_of(step);
    const Planes P = planes(z);
    int e0 = rb;  // opaque per plane: cheaper to re-derive the row addresses than to hold them
    asm volatile("" : "+v"(e0));
    // the upwind plane in global memory, for S when it is not held in LDS
    const bool up_ghost = step == 0;
    const int zu = z - a.dir;
    const uint32_t ub = up_ghost ? a.gplane_bytes : a.plane_bytes;
    const Rsrc up_S = rsrc(up_ghost ? gS : a.S + (int64_t)zu * a.cs_z + (int64_t)f * a.cs_f, ub);
    const uint32_t uy = up_ghost ? a.gs_y : a.cs_y;
    const uint32_t ux = up_ghost ? a.gs_x : a.cs_x;
    // MODE 4: the taps' LDS slots less r*nx (window row r in slot strip*SR + r-1),
    // opaque per plane like e0
    int sa[CPL + 1];
#pragma unroll
    for (int j = 0; j <= CPL; ++j) {
      sa[j] = (strip * SR - 1) * nx + xc[j];
      if (MODE == 4) asm volatile("" : "+v"(sa[j]));
    }

    // one upwind row for the CPL cells: I at CPL+3 columns, x and S at CPL+1
    struct RawRow { T i[CPL + 3], x[CPL + 1], s[CPL + 1]; };
    auto read_row = [&](int r) -> RawRow {
      RawRow q;
      const int e = e0 + r * PW;
#pragma unroll
      for (int j = 0; j < CPL + 3; ++j) {
        const Pair<T> p = sP[e + j];
        q.i[j] = p.a;
        if (j >= 1 && j <= CPL + 1) q.x[j - 1] = p.b;
      }
      if (MODE == 2) {
#pragma unroll
        for (int j = 0; j <= CPL; ++j) q.s[j] = sS[e + 1 + j];
      } else if (MODE == 4 && (r < 1 || r > L + 1)) {
#pragma unroll
        for (int j = 0; j <= CPL; ++j) q.s[j] = (T)0;  // rows whose S no step uses
      } else if (MODE == 4 && r <= SR) {
        // (r is a constant of the unrolled loop: LDS or cache is decided at compile time)
#pragma unroll
        for (int j = 0; j <= CPL; ++j) q.s[j] = sS[sa[j] + r * nx];
      } else {
        int g = g0 + r;
        g = g >= ny ? g - ny : g;
        const uint32_t ro = g * uy * TB;
#pragma unroll
        for (int j = 0; j <= CPL; ++j) q.s[j] = bld<T>(up_S, ro + xc[j] * ux * TB, 0);
      }
      return q;
    };
    struct Row { T rI, rmin, rX, rS; };
    struct Rows { Row c[CPL]; };
    auto reduce_row = [&](const RawRow &q) -> Rows {
      Rows w;
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        w.c[c].rI = a.xk0 * q.i[c] + a.xk1 * q.i[c + 1] + a.xk2 * q.i[c + 2] + a.xk3 * q.i[c + 3];
        w.c[c].rmin = fmin(q.i[c + 1], q.i[c + 2]);
        w.c[c].rX = ohx * q.x[c] + a.hx * q.x[c + 1];
        w.c[c].rS = ohx * q.s[c] + a.hx * q.s[c + 1];
      }
      return w;
    };

    T newI[CPL][L], newX[CPL][L], newS[CPL][MODE == 2 || MODE == 4 ? L : 1];
    Rows w0 = reduce_row(read_row(0));
    Rows w1 = reduce_row(read_row(1));
    Rows w2 = reduce_row(read_row(2));
    RawRow next = read_row(3);
#pragma unroll
    for (int k = 0; k < L; ++k) {
      __builtin_amdgcn_sched_barrier(0);
      const In in = k < PD ? pre[k] : ring[k % PD];
      if (k + PD < L) ring[k % PD] = fetch(P, k + PD);
      const Rows w3 = reduce_row(next);
      if (k + 1 < L) next = read_row(k + 4);
      V Jn, PJn, Io, Po, Ro, Eo;
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        // upwind values, SRC/rt_solvers_mod.f90:379-394
        const T xxup = ohy * w1.c[c].rX + a.hy * w2.c[c].rX;
        const T ssup = ohy * w1.c[c].rS + a.hy * w2.c[c].rS;
        T Inuup = a.yk0 * w0.c[c].rI + a.yk1 * w1.c[c].rI + a.yk2 * w2.c[c].rI + a.yk3 * w3.c[c].rI;
        Inuup = fmax(Inuup, fmin(w1.c[c].rmin, w2.c[c].rmin));
        const T chiv = in.chi.v[c], ss = in.S.v[c], dt500 = in.dr.v[c].a, rc = in.dr.v[c].b;
        const T xx = chiv * rc;
        T dt = (T)0.5 * (xxup + xx) * dt500;
        dt = fmax(dt, a.mdtau);
        T E, W, Ps, Es;
        lin_weights(dt, E, W, Ps, Es);
        // the S step of a cell without eta flags (SRC/rt_solvers_mod.f90:419-440, 465-475)
        const T Ir = Inuup * E + W * ssup;
        T Inu = Ir + Ps * ss;
        bool soff = false;
        T psio = (T)0;
        if (SPL) {
          const T sgnchi = copysign(fmax(fabs(chiv), (T)1.17549435e-38f), chiv);
          soff = fmin(dt, chiv) <= (T)0 || (a.imax - fabs(Inu)) <= (T)0;
          psio = soff ? (T)0 : Ps * rcp(sgnchi);
        }
        Inu = fmax(-a.imax, fmin(a.imax, Inu));
        newI[c][k] = Inu;
        if (MODE >= 1) newX[c][k] = xx;
        if (MODE == 2 || (MODE == 4 && (k < SR || k == L - 1))) newS[c][k] = ss;
        if (IOUT) Io.v[c] = Inu;
        if (SPL) {
          Po.v[c] = psio;
          Ro.v[c] = soff ? Inu : Ir;
          Eo.v[c] = soff ? (T)1 : Es;
        }
        if (!IOUT) Jn.v[c] = in.J.v[c] + a.wj * Inu;
        if (PJ) {
          const T sgnchi = copysign(fmax(fabs(chiv), (T)1.17549435e-38f), chiv);
          const bool off = fmin(dt, chiv) <= (T)0;
          const T psiout = off ? (T)0 : Ps * rcp(sgnchi);
          PJn.v[c] = in.PsJ.v[c] + a.wj * psiout;
        }
      }
      if (IOUT) bst<V>(P.Io, voff, k * rowb, Io);
      if (SPL) {
        bst<V>(P.Po, voff, k * rowb, Po);
        bst<V>(P.Ro, voff, k * rowb, Ro);
        bst<V>(P.Eo, voff, k * rowb, Eo);
      }
      if (!IOUT) bst<V>(P.J, voff, k * rowb, Jn);
      if (PJ) bst<V>(P.PsJ, voff, k * rowb, PJn);
      w0 = w1;
      w1 = w2;
      w2 = w3;
    }
    __builtin_amdgcn_sched_barrier(0);
    if (step + 1 < a.nz) {
      const Planes Pn = planes(plane_of(step + 1));
#pragma unroll
      for (int k = 0; k < PD; ++k) pre[k] = fetch(Pn, k);
    }
    lds_barrier();  // every read of the upwind plane is done
    // the new plane, then its halo copies: columns nx-2, nx-1 to the left pad,
    // 0, 1 to the right; rows ny-2, ny-1 to the top pad, 0, 1 to the bottom.  The
    // base index is opaque per plane so that the write addresses are formed here,
    // not held in registers across the whole sweep.
    int wb = (yb + 2) * PW + x0 + 2;
    asm volatile("" : "+v"(wb));
    const bool xedge = x0 < 2 || x0 + CPL - 1 >= nx - 2;
    const int dh = x0 < 2 ? nx : -nx;  // to this group's halo copy
    auto put = [&](int e, int k) {
      Arr<Pair<T>, CPL> pp;
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        pp.v[c].a = newI[c][k];
        pp.v[c].b = newX[c][k];
      }
      *reinterpret_cast<Arr<Pair<T>, CPL> *>(sP + e) = pp;
      if (MODE == 2) {
        V sv;
#pragma unroll
        for (int c = 0; c < CPL; ++c) sv.v[c] = newS[c][k];
        *reinterpret_cast<V *>(sS + e) = sv;
      }
    };
#pragma unroll
    for (int k = 0; k < L; ++k) put(wb + k * PW, k);
    if (xedge) {
#pragma unroll
      for (int k = 0; k < L; ++k) put(wb + k * PW + dh, k);
    }
    if (strip == 0) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        put(wb + (k + ny) * PW, k);
        if (xedge) put(wb + (k + ny) * PW + dh, k);
      }
    }
    if (strip == a.nstrips - 1) {
#pragma unroll
      for (int k = L - 2; k < L; ++k) {
        put(wb + (k - ny) * PW, k);
        if (xedge) put(wb + (k - ny) * PW + dh, k);
      }
    }
    if (MODE == 4) {
      // row yb+k is window row k-yo of this strip (slot strip*SR + k-yo-1 while
      // k-yo <= SR) and, for k = L-1 with yo = -2, window row 1 of the next strip
      // (periodically: strip 0 after the last); the other rows to the spare row
      // (selects, not branches: a branch per row costs registers)
      const int c1 = -a.yo - 1;
      const int spare = a.nstrips * SR;
      const int nxt = strip + 1 == a.nstrips ? 0 : (strip + 1) * SR;
#pragma unroll
      for (int k = 0; k < SR; ++k) {  // rows SR .. L-2 are never held
        const int rr = k + c1;
        sS[(rr < SR ? strip * SR + rr : spare) * nx + x0] = newS[0][k];
      }
      sS[(c1 ? nxt : spare) * nx + x0] = newS[0][L - 1];
    }
    if (a.exitI && step == a.nz - 1) {
      const Rsrc ex = rsrc(a.exitI + f * a.gs_f, a.gplane_bytes);
      uint32_t eo = (x0 * a.gs_x + yb * a.gs_y) * TB;
      asm volatile("" : "+v"(eo));
#pragma unroll
      for (int k = 0; k < L; ++k) {
        if (CPL == 1) {
          bst<T>(ex, eo, k * a.gs_y * TB, newI[0][k]);
        } else {
          V ev;
#pragma unroll
          for (int c = 0; c < CPL; ++c) ev.v[c] = newI[c][k];
          bst<V>(ex, eo, k * a.gs_y * TB, ev);
        }
      }
    }
    lds_barrier();
  }
}

template <int L, bool PJ, int MODE, int CPL, int SR = 0, int OM = 0>
int launch_fast(m3drt_ctx *ctx, const FastArgs<float> &a, int nf, int nt, size_t lds) {
  auto kern = sc2_fast<L, PJ, MODE, CPL, SR, OM>;
  M3DRT_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(nf), dim3(nt), lds, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  return M3DRT_OK;
}

// MODE 4's rows per strip held in LDS for strips of 20 rows (the largest that fits
// and does not exceed the cap in run_fast is taken)
constexpr int kSR20[] = {8, 14};

template <int L, int CPL>
int fast_variant(m3drt_ctx *ctx, const FastArgs<float> &a, int nf, int nt, int mode, size_t np, int sr) {
  const size_t l2 = np * 2 * sizeof(float), l3 = np * 3 * sizeof(float);
  const size_t l4 = l2 + (size_t)(a.nstrips * sr + 1) * a.nx * sizeof(float);
  if (a.Iout) {  // per-cell outputs (the split set, or I alone), for the strip lengths 4 to 20
    if constexpr (CPL == 1 && (L == 4 || L == 12 || L == 15 || L == 20)) {
      if (a.Psiout) {
        if (mode == 2) return launch_fast<L, false, 2, CPL, 0, 1>(ctx, a, nf, nt, l3);
        if (mode == 1) return launch_fast<L, false, 1, CPL, 0, 1>(ctx, a, nf, nt, l2);
      } else {
        if (mode == 2) return launch_fast<L, false, 2, CPL, 0, 2>(ctx, a, nf, nt, l3);
        if (mode == 1) return launch_fast<L, false, 1, CPL, 0, 2>(ctx, a, nf, nt, l2);
      }
    }
    return -1;
  }
  if (a.PsiJ) {
    if constexpr (CPL == 1) {
      if (mode == 2) return launch_fast<L, true, 2, CPL>(ctx, a, nf, nt, l3);
      return launch_fast<L, true, 1, CPL>(ctx, a, nf, nt, l2);
    }
    return -1;
  }
  if (mode == 2) return launch_fast<L, false, 2, CPL>(ctx, a, nf, nt, l3);
  if constexpr (CPL == 1 && L == 20) {
    if (mode == 4 && sr == 14) return launch_fast<L, false, 4, CPL, 14>(ctx, a, nf, nt, l4);
    if (mode == 4 && sr == 8) return launch_fast<L, false, 4, CPL, 8>(ctx, a, nf, nt, l4);
  }
  return launch_fast<L, false, 1, CPL>(ctx, a, nf, nt, l2);
}

// strip lengths instantiated: one lane per column (60, 96, 120-cell planes), two
// adjacent columns per lane (native layout, even nx).  The first that fits is taken,
// except that strips of 20 rows are preferred where they give 512 to 768 threads: on
// the 120 x 120 plane (720 threads, up to 168 registers a lane) they measured 13.3 ms
// per 1024-frequency direction against 13.7 for 15 (960 threads, 128 registers),
// 14.8 for 24 and 14.5 for 30.
constexpr int kL1[] = {4, 12, 15, 20, 24, 30};
constexpr int kL2[] = {4, 6, 8};

int fast_dispatch(m3drt_ctx *ctx, const FastArgs<float> &a, int nf, int nt, int L, int cpl, int mode, size_t np,
                  int sr) {
  if (cpl == 2) {
    if (L == 4) return fast_variant<4, 2>(ctx, a, nf, nt, mode, np, sr);
    if (L == 6) return fast_variant<6, 2>(ctx, a, nf, nt, mode, np, sr);
    if (L == 8) return fast_variant<8, 2>(ctx, a, nf, nt, mode, np, sr);
    return -1;
  }
  if (L == 4) return fast_variant<4, 1>(ctx, a, nf, nt, mode, np, sr);
  if (L == 12) return fast_variant<12, 1>(ctx, a, nf, nt, mode, np, sr);
  if (L == 15) return fast_variant<15, 1>(ctx, a, nf, nt, mode, np, sr);
  if (L == 20) return fast_variant<20, 1>(ctx, a, nf, nt, mode, np, sr);
  if (L == 24) return fast_variant<24, 1>(ctx, a, nf, nt, mode, np, sr);
  if (L == 30) return fast_variant<30, 1>(ctx, a, nf, nt, mode, np, sr);
  return -1;
}

template <typename T>
int run_fast(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  static_assert(sizeof(T) == 4, "the straight-line sweep is instantiated for FP32");
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  const bool get_rates = (d->flags & M3DRT_GET_RATES) != 0;
  // J, or J and PsiJ, or only the exit plane (no J: its descriptor then has size 0,
  // so the kernel's J loads read zero and its J stores are dropped); or the ALI's
  // split set (I, Psi, Irem, escape with rates, no J), with or without the exit plane
  const bool spl = (d->flags & M3DRT_SPLIT_RATES) && get_rates && b->I_out && b->Psi_out && b->Irem_out &&
                   b->escape_out && !b->J && !b->PsiJ;
  // or the intensity cube alone (no J, no rates), with or without the exit plane
  const bool iout = !spl && b->I_out && !b->J && !b->PsiJ && !b->Psi_out && !b->Irem_out && !b->escape_out &&
                    !(d->flags & M3DRT_SPLIT_RATES);
  if (b->eta_row) return -1;
  if (!spl && !iout) {
    if (b->I_out || (d->flags & M3DRT_SPLIT_RATES) || (get_rates && b->Psi_out)) return -1;
    if (!b->J && (!b->exit_I || (get_rates && b->PsiJ))) return -1;
  }
  if (nx < 4 || ny < 4 || nx > 1024 || getenv("M3DRT_NO_FAST")) return -1;
  // one column per lane.  Two (M3DRT_FAST_CPL=2, where the rows are contiguous, nx is
  // even and J is the only output: with PsiJ the pair form needs more registers than a
  // lane has) halve the memory instructions per cell but measured 4 % slower on the
  // 120 x 120 plane (14.9 against 14.3 ms per 1024-frequency direction).
  const char *cenv = getenv("M3DRT_FAST_CPL");
  int cpl = 1;
  if (cenv && atoi(cenv) == 2 && d->cs_x == 1 && d->gs_x == 1 && nx % 2 == 0 && !(get_rates && b->PsiJ) && !spl &&
      !iout)
    cpl = 2;
  // the strip length: the first instantiated one that tiles ny into at least two
  // strips within its workgroup's thread limit (M3DRT_FAST_L picks one of them)
  const char *lenv = getenv("M3DRT_FAST_L");
  const int want = lenv ? atoi(lenv) : 0;
  auto max_threads = [](int l) { return l >= 30 ? 512 : l >= 20 ? 768 : 1024; };
  int L = 0;
  if (!want && cpl == 1 && ny % 20 == 0 && ny / 20 >= 2 && (ny / 20) * nx >= 512 && (ny / 20) * nx <= 768) L = 20;
  // Where the pairs and S of three workgroups fit one CU's LDS, a strip of 15 (or 20)
  // rows that still gives a workgroup more than three waves: four rows of loads in
  // flight instead of three, three workgroups per CU.  On the 60 x 60 plane the
  // split-output sweep took 6.47 ms per 2400-frequency direction with strips of 15,
  // 6.55 with 20, 7.40 with 12 and 7.05 with 4 (900 threads, one workgroup per CU).
  const size_t np_all = (size_t)(nx + 4) * (ny + 4);
  if (!want && !L && cpl == 1 && 3 * np_all * 3 * sizeof(T) <= (size_t)ctx->lds_max)
    for (int c : {15, 20})
      if (!L && ny % c == 0 && ny / c >= 2 && (ny / c) * nx > 192 && (ny / c) * nx <= max_threads(c)) L = c;
  for (int pass = 0; pass < 2 && !L; ++pass) {
    const int nxp = nx / cpl;
    const int *cand = cpl == 2 ? kL2 : kL1;
    const int ncand = cpl == 2 ? (int)(sizeof(kL2) / sizeof(int)) : (int)(sizeof(kL1) / sizeof(int));
    for (int i = 0; i < ncand && !L; ++i) {
      const int c = cand[i];
      if (want && c != want) continue;
      if (ny % c == 0 && ny / c >= 2 && (ny / c) * nxp <= max_threads(c)) L = c;
    }
    if (!L && cpl == 2) cpl = 1;
  }
  if (!L) return -1;
  const int nstrips = ny / L;
  const int nt = (nx / cpl) * nstrips;
  const size_t np = (size_t)(nx + 4) * (ny + 4);
  const uint64_t span = (uint64_t)(nx - 1) * d->cs_x + (uint64_t)(ny - 1) * d->cs_y;
  const uint64_t gspan = (uint64_t)(nx - 1) * d->gs_x + (uint64_t)(ny - 1) * d->gs_y;
  if ((span + 1) * sizeof(T) >= (1ull << 32) || (gspan + 1) * sizeof(T) >= (1ull << 32)) return -1;
  int mode = -1, sr = 0;
  if (np * 3 * sizeof(T) <= (size_t)ctx->lds_max) {
    mode = 2;
  } else if (np * 2 * sizeof(T) <= (size_t)ctx->lds_max) {
    mode = 1;
    // S of the first sr tap rows of every strip (nstrips * sr rows of nx, and a
    // spare row) in the LDS the pairs leave: J only, one column per lane, strips of
    // 20 rows (the form that holds it without spilling registers)
    const size_t rows = ((size_t)ctx->lds_max - np * 2 * sizeof(T)) / ((size_t)nx * sizeof(T));
    const bool j_only = !(get_rates && b->PsiJ) && !spl && !iout;
    if (j_only && cpl == 1 && L == 20 && rows >= 1 && !getenv("M3DRT_FAST_NO_SROWS")) {
      // the most rows per strip to hold: 8 unless M3DRT_FAST_SR says otherwise.  On
      // the 120 x 120 plane, per 1024-frequency direction: 13.35 ms with 8 and 13.45
      // with 14 over all 24 directions, 13.32 and 13.10 over four of them on another
      // box; 13.5-13.8 with none.
      const char *senv = getenv("M3DRT_FAST_SR");
      const int smax = senv ? atoi(senv) : 8;
      for (int c : kSR20)
        if ((size_t)nstrips * c + 1 <= rows && c <= smax) sr = c;
      if (sr) mode = 4;
    }
  }
  if (mode < 0) return -1;

  const Sc2Geom g = sc2_geometry(d);
  const int64_t nxy = (int64_t)nx * ny;
  const size_t npairs = (size_t)nxy * (nz + 1);
  int rc = ensure_scratch(ctx, sizeof(Pair<T>) * npairs);
  if (rc) return rc;
  Pair<T> *dr = static_cast<Pair<T> *>(ctx->scratch);
  {
    const int blocks = (int)std::min<int64_t>(((int64_t)npairs + 255) / 256, 65536);
    hipLaunchKernelGGL(sc2_prep_pairs<T>, dim3(blocks), dim3(256), 0, ctx->stream, b->chi500, b->chi500_up,
                       b->g500, dr, (int)nxy, nz, g.dsup);
    M3DRT_HIP(hipGetLastError());
  }
  FastArgs<T> a;
  a.chi = static_cast<const T *>(b->chi);
  a.S = static_cast<const T *>(b->S);
  a.gchi = static_cast<const T *>(b->ghost_chi);
  a.gS = static_cast<const T *>(b->ghost_S);
  a.gI = static_cast<const T *>(b->ghost_I);
  a.dr = dr;
  a.J = static_cast<T *>(b->J);
  a.PsiJ = get_rates ? static_cast<T *>(b->PsiJ) : nullptr;
  a.exitI = static_cast<T *>(b->exit_I);
  a.Iout = spl || iout ? static_cast<T *>(b->I_out) : nullptr;
  a.Psiout = spl ? static_cast<T *>(b->Psi_out) : nullptr;
  a.Irem = spl ? static_cast<T *>(b->Irem_out) : nullptr;
  a.esc = spl ? static_cast<T *>(b->escape_out) : nullptr;
  a.cs_f = d->cs_f; a.cs_z = d->cs_z; a.gs_f = d->gs_f;
  a.cs_x = (uint32_t)d->cs_x; a.cs_y = (uint32_t)d->cs_y;
  a.gs_x = (uint32_t)d->gs_x; a.gs_y = (uint32_t)d->gs_y;
  a.nx = nx; a.ny = ny; a.nz = nz; a.dir = d->dir > 0 ? 1 : -1;
  a.xo = g.xo; a.yo = g.yo; a.PW = nx + 4; a.nstrips = nstrips;
  a.plane_bytes = (uint32_t)((span + 1) * sizeof(T));
  a.gplane_bytes = (uint32_t)((gspan + 1) * sizeof(T));
  // the stencil in FP32 as the reference casts it (FP64 sweep: in FP64)
  const bool f64 = sizeof(T) == 8;
  auto cast = [&](double v) { return f64 ? (T)v : (T)(float)v; };
  a.xk0 = cast(g.kx[0]); a.xk1 = cast(g.kx[1]); a.xk2 = cast(g.kx[2]); a.xk3 = cast(g.kx[3]);
  a.yk0 = cast(g.ky[0]); a.yk1 = cast(g.ky[1]); a.yk2 = cast(g.ky[2]); a.yk3 = cast(g.ky[3]);
  // inner_fraction (SRC/rt_solvers_mod.f90:500-506): the first moment of the kernel
  a.hx = -a.xk0 + a.xk2 + (T)2 * a.xk3;
  a.hy = -a.yk0 + a.yk2 + (T)2 * a.yk3;
  a.mdtau = f64 ? -(T)d->maser_dtau_max : (T)(-(float)d->maser_dtau_max);
  a.imax = f64 ? (T)d->maser_Inu_max : (T)(float)d->maser_Inu_max;
  a.wj = (T)d->wj;
  if (getenv("M3DRT_VERBOSE"))
    fprintf(stderr, "[m3drt] sc2 fast %s nf=%d plane=%dx%d nz=%d: %d threads, %d column(s) per lane, strip %d, "
            "LDS mode %d (%zu B, %d S rows of %d per strip in LDS)%s\n", f64 ? "f64" : "f32", d->nf, nx, ny, nz,
            nt, cpl, L, mode,
            np * (mode == 2 ? 3 : 2) * sizeof(T) + (mode == 4 ? (size_t)(nstrips * sr + 1) * nx * sizeof(T) : 0),
            mode == 2 ? L + 1 : sr, L + 1, a.PsiJ ? ", PsiJ" : spl ? ", split outputs" : iout ? ", I out" : "");
  return fast_dispatch(ctx, a, d->nf, nt, L, cpl, mode, np, sr);
}

}  // namespace

// FP32 only: the FP64 sweep holds twice the bytes per register and takes the
// general kernel's exact-strip form.
int sweep_sc2_fast(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b) {
  if (d->prec == M3DRT_F64) return -1;
  return run_fast<float>(ctx, d, b);
}

}  // namespace m3drt
