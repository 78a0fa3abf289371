// rotate_array (SRC/rt_patch_mod.f90:3809-3863): every z plane of the cubes shifted
// periodically by the ray's horizontal offset at that height, with bilinear
// interpolation, so that the rays of one direction become columns for the output
// pass's column solvers.  The per-plane integer shift and the four weights are formed
// on the host in the reference's precision and order (real(4), the mixed real(8)
// ray_offset sum, floor); the kernel is a gather of four points per element, in the
// reference's order of the weighted sum -- HBM-bound, one read of each input element
// per output element from the cache and one write.
#include "m3drt_internal.hpp"
#include <climits>
#include <cmath>
#include <cstdlib>
#include <vector>

#pragma clang fp contract(off)

namespace m3drt {
namespace {

struct PlaneShift {
  int up1, up2;
  float a, b, c, d;
};

struct RotArgs {
  const float *in;
  float *out;
  const PlaneShift *sh;
  int64_t cs_f, cs_x, cs_y, cs_z, cs_k;
  int nf, nx, ny, nz, nk;
  int fast_f;  // the unit-stride axis runs along f (else along x)
};

// I: the index type of the element number within a plane (32 bits where the plane
// fits, as integer division is far cheaper there); the shifts are reduced to [0, n)
// on the host, so a shifted index wraps at most once
template <typename I>
__global__ void __launch_bounds__(256) rotate_kernel(const RotArgs a) {
  const I n = (I)a.nf * a.nx * a.ny;
  const int z = blockIdx.y, k = blockIdx.z;
  const PlaneShift s = a.sh[z];
  const int64_t base = (int64_t)z * a.cs_z + (int64_t)k * a.cs_k;
  for (I i = blockIdx.x * (I)blockDim.x + threadIdx.x; i < n; i += (I)gridDim.x * blockDim.x) {
    int f, x, y;
    if (a.fast_f) {
      f = (int)(i % a.nf);
      const I r = i / a.nf;
      x = (int)(r % a.nx);
      y = (int)(r / a.nx);
    } else {
      x = (int)(i % a.nx);
      const I r = i / a.nx;
      y = (int)(r % a.ny);
      f = (int)(r / a.ny);
    }
    // modulo(x + up - 1, n) + 1 in 1-based indices: (x + up) mod n here
    int x1 = x + s.up1, y1 = y + s.up2;
    x1 -= x1 >= a.nx ? a.nx : 0;
    y1 -= y1 >= a.ny ? a.ny : 0;
    const int x2 = x1 + 1 == a.nx ? 0 : x1 + 1, y2 = y1 + 1 == a.ny ? 0 : y1 + 1;
    const float *p = a.in + base + (int64_t)f * a.cs_f;
    float v = s.a * p[x1 * a.cs_x + y1 * a.cs_y];
    v = s.b * p[x2 * a.cs_x + y1 * a.cs_y] + v;
    v = s.c * p[x1 * a.cs_x + y2 * a.cs_y] + v;
    v = s.d * p[x2 * a.cs_x + y2 * a.cs_y] + v;
    a.out[base + (int64_t)f * a.cs_f + x * a.cs_x + y * a.cs_y] = v;
  }
}

// The same gather where x is the unit-stride axis (the native layout): each wave
// takes RW consecutive output rows (k, z, f, y) per pass, its lanes along x.  A row's
// shift, weights and both input rows are wave-uniform, so no index is recovered by
// division per element; a lane's four taps are two adjacent elements of each of two
// input rows.  The taps of all RW rows are loaded before any is stored, so a wave
// has RW rows of loads in flight.  Rows are taken in memory order: the input row y2
// of one output row is the y1 of the next, read from the cache.  The weighted sum is
// formed as in rotate_kernel.
template <int RW>
__global__ void __launch_bounds__(256) rotate_rows_kernel(const RotArgs a, int nrows) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  for (int r0 = (blockIdx.x * (blockDim.x >> 6) + wave) * RW; r0 < nrows; r0 += nwaves * RW) {
    const float *p1[RW], *p2[RW];
    float *o[RW];
    PlaneShift s[RW];
#pragma unroll
    for (int j = 0; j < RW; ++j) {
      // r = ((k * nz + z) * nf + f) * ny + y; rows past the end repeat the last one
      const int r = min(r0 + j, nrows - 1);
      const int y = r % a.ny;
      int q = r / a.ny;
      const int f = q % a.nf;
      q /= a.nf;
      const int z = q % a.nz, k = q / a.nz;
      s[j] = a.sh[z];
      int y1 = y + s[j].up2;
      y1 -= y1 >= a.ny ? a.ny : 0;
      const int y2 = y1 + 1 == a.ny ? 0 : y1 + 1;
      const int64_t base = (int64_t)z * a.cs_z + (int64_t)k * a.cs_k + (int64_t)f * a.cs_f;
      p1[j] = a.in + base + (int64_t)y1 * a.cs_y;
      p2[j] = a.in + base + (int64_t)y2 * a.cs_y;
      o[j] = a.out + base + (int64_t)y * a.cs_y;
    }
    for (int x = lane; x < a.nx; x += 64) {
      float t[RW][4];
#pragma unroll
      for (int j = 0; j < RW; ++j) {
        int x1 = x + s[j].up1;
        x1 -= x1 >= a.nx ? a.nx : 0;
        const int x2 = x1 + 1 == a.nx ? 0 : x1 + 1;
        t[j][0] = p1[j][x1 * a.cs_x];
        t[j][1] = p1[j][x2 * a.cs_x];
        t[j][2] = p2[j][x1 * a.cs_x];
        t[j][3] = p2[j][x2 * a.cs_x];
      }
#pragma unroll
      for (int j = 0; j < RW; ++j) {
        if (r0 + j >= nrows) break;
        float v = s[j].a * t[j][0];
        v = s[j].b * t[j][1] + v;
        v = s[j].c * t[j][2] + v;
        v = s[j].d * t[j][3] + v;
        o[j][x * a.cs_x] = v;
      }
    }
  }
}

}  // namespace

// the per-plane shifts, as rotate_array forms them (real(4) but for the real(8)
// ray_offset in the sum that gives deltak)
void rotate_shifts(const m3drt_rotate_args *r, std::vector<int> &up1, std::vector<int> &up2,
                   std::vector<float> &w) {
  const float delta1 = (float)(r->vec[0] / (r->vec[2] * r->ds[0]));
  const float delta2 = (float)(r->vec[1] / (r->vec[2] * r->ds[1]));
  const float bot2pivot = (float)r->bot2pivot;
  up1.resize(r->nz);
  up2.resize(r->nz);
  w.resize(4 * (size_t)r->nz);
  for (int z = 1; z <= r->nz; ++z) {
    const float height = (float)(z * r->ds[2]);
    const float t = bot2pivot - height;
    const float d1 = (float)((double)(t * delta1) + r->ray_offset[0]);
    const float d2 = (float)((double)(t * delta2) + r->ray_offset[1]);
    const int u1 = (int)std::floor(d1), u2 = (int)std::floor(d2);
    const float f1 = d1 - (float)u1, f2 = d2 - (float)u2;
    up1[z - 1] = u1;
    up2[z - 1] = u2;
    w[4 * (z - 1) + 0] = (1.0f - f1) * (1.0f - f2);
    w[4 * (z - 1) + 1] = f1 * (1.0f - f2);
    w[4 * (z - 1) + 2] = (1.0f - f1) * f2;
    w[4 * (z - 1) + 3] = f1 * f2;
  }
}

int rotate_array(m3drt_ctx *ctx, const m3drt_rotate_args *r) {
  std::vector<int> up1, up2;
  std::vector<float> w;
  rotate_shifts(r, up1, up2, w);
  std::vector<PlaneShift> sh(r->nz);
  for (int z = 0; z < r->nz; ++z) {
    // the integer shift reduced to [0, n) keeps the kernel's index arithmetic small
    auto red = [](int u, int n) { return ((u % n) + n) % n; };
    sh[z] = PlaneShift{red(up1[z], r->nx), red(up2[z], r->ny), w[4 * z], w[4 * z + 1], w[4 * z + 2], w[4 * z + 3]};
  }
  int rc = ensure_scratch(ctx, sizeof(PlaneShift) * sh.size());
  if (rc) return rc;
  M3DRT_HIP(hipMemcpyAsync(ctx->scratch, sh.data(), sizeof(PlaneShift) * sh.size(), hipMemcpyHostToDevice,
                           ctx->stream));
  RotArgs a;
  a.in = r->in;
  a.out = r->out;
  a.sh = static_cast<const PlaneShift *>(ctx->scratch);
  a.cs_f = r->cs_f; a.cs_x = r->cs_x; a.cs_y = r->cs_y; a.cs_z = r->cs_z; a.cs_k = r->cs_k;
  a.nf = r->nf; a.nx = r->nx; a.ny = r->ny; a.nz = r->nz; a.nk = r->nk;
  a.fast_f = r->cs_f <= r->cs_x ? 1 : 0;
  const int64_t n = (int64_t)r->nf * r->nx * r->ny;
  const int bx = (int)std::min<int64_t>((n + 255) / 256, 4096);
  // rows of x when x runs fastest: four rows a wave per pass, four waves a
  // workgroup (the row count and the waves' stride over it fit int32)
  constexpr int RW = 4;
  const int64_t nrows = (int64_t)r->nk * r->nz * r->nf * r->ny;
  const int rb = (int)std::min<int64_t>((nrows + 4 * RW - 1) / (4 * RW), 8192);
  if (!a.fast_f && nrows + (int64_t)rb * 4 * RW <= INT32_MAX && !getenv("M3DRT_ROTATE_FLAT"))
    hipLaunchKernelGGL(rotate_rows_kernel<RW>, dim3(rb), dim3(256), 0, ctx->stream, a, (int)nrows);
  else if (n + (int64_t)bx * 256 <= INT32_MAX)
    hipLaunchKernelGGL(rotate_kernel<int>, dim3(bx, r->nz, r->nk), dim3(256), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL(rotate_kernel<int64_t>, dim3(bx, r->nz, r->nk), dim3(256), 0, ctx->stream, a);
  M3DRT_HIP(hipGetLastError());
  // the shifts stay in the scratch buffer until the kernel has read them
  M3DRT_HIP(hipStreamSynchronize(ctx->stream));
  return M3DRT_OK;
}

}  // namespace m3drt
