// Shared host-side geometry of the short-characteristics sweeps (sweep_sc2.hip,
// sc2_fast.hip): the upwind stencil of one direction, SRC/ray_mod.f90:493-508
// (set_stencil) and 581-596 (cc_kernel), and the inner fraction of
// SRC/rt_solvers_mod.f90:500-506.
#ifndef M3DRT_SWEEP_SC2_HPP
#define M3DRT_SWEEP_SC2_HPP
#include <cmath>
#include "m3drt_internal.hpp"

namespace m3drt {

struct Sc2Geom {
  double dsup;        // path length to the upwind plane
  int xo, yo;         // first stencil column / row relative to the cell (-1 or -2)
  double kx[4], ky[4];  // Keys (a = -1/2) weights at the fractional upwind position
};

inline Sc2Geom sc2_geometry(const m3drt_sweep_desc *d) {
  Sc2Geom g;
  g.dsup = d->ds[2] / std::fabs(d->vec[2]);
  const double wt1 = -d->vec[0] / d->ds[0] * g.dsup, wt2 = -d->vec[1] / d->ds[1] * g.dsup;
  const double h1 = wt1 > 0 ? wt1 : wt1 + 1, h2 = wt2 > 0 ? wt2 : wt2 + 1;
  auto cc = [](double s, double k[4]) {
    const double s2 = s * s, s3 = s2 * s;
    k[0] = 0.5 * (-s3 + 2 * s2 - s);
    k[1] = 0.5 * (3 * s3 - 5 * s2 + 2);
    k[2] = 0.5 * (-3 * s3 + 4 * s2 + s);
    k[3] = 0.5 * (s3 - s2);
  };
  cc(h1, g.kx);
  cc(h2, g.ky);
  g.xo = wt1 > 0 ? -1 : -2;
  g.yo = wt2 > 0 ? -1 : -2;
  return g;
}

// dt500 of every cell's step and 1/chi500 (FP32, then the ghost plane's 1/g500) in
// the context's scratch: the per-direction preparation of the short_char2 sweeps.
int sc2_prep_f32(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const float *chi500, const float *chi500_up,
                 const float *g500, float **dt500, float **rc);

// The straight-line sweep for J (and PsiJ) only; returns -1 when the case is not
// one it takes (the caller then runs the general kernel), else a status.
int sweep_sc2_fast(m3drt_ctx *ctx, const m3drt_sweep_desc *d, const m3drt_sweep_bufs *b);

}  // namespace m3drt
#endif
