// Angular quadratures of the two transfer passes (setup_quadrature,
// SRC/rt_params_mod.f90:170-220; the sets of SRC/ray_mod.f90:36-467 and the
// Gauss-Radau rule of solvers/rt/radau_mod.f90:15-115).  Host code: the ray
// directions and weights are a few hundred doubles, formed once per run.
//
// Directions keep the reference's arithmetic so that the sweeps see the same rays:
// the tabulated Gauss, Lobatto and Carlson nodes are default-real (FP32) literals
// widened to FP64 there, and so are they here; the Gauss-Radau nodes are computed
// (Newton on the Legendre recurrence) and mapped to (0, 1] in the precision of
// radau%init, whose results are FP32.
#include "m3drt_internal.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace m3drt {
namespace {

constexpr double kPi = 3.14159265359;  // HOP/const_mod.f90:14

struct Rays {
  std::vector<double> vec;  // (n, 3)
  std::vector<double> w;
  size_t n() const { return w.size(); }
  void add(double x, double y, double z, double wt) {
    vec.push_back(x);
    vec.push_back(y);
    vec.push_back(z);
    w.push_back(wt);
  }
};

double f32(double v) { return (double)(float)v; }

// Gauss-Legendre on (0, 1), SRC/ray_mod.f90:342-373 (FP32 literals)
bool gauss_table(int n, std::vector<double> &mu, std::vector<double> &wmu) {
  static const float M[6][6] = {{1.0f},
                                {0.78867513f, 0.21132487f},
                                {0.88729833f, 0.50000000f, 0.11270167f},
                                {0.93056816f, 0.66999052f, 0.33000948f, 0.06943184f},
                                {0.95308992f, 0.76923466f, 0.50000000f, 0.23076534f, 0.04691008f},
                                {0.96623476f, 0.83060469f, 0.61930959f, 0.38069041f, 0.16939531f, 0.03376524f}};
  static const float W[6][6] = {{1.0f},
                                {0.50000000f, 0.50000000f},
                                {0.27777778f, 0.44444444f, 0.27777778f},
                                {0.17392742f, 0.32607258f, 0.32607258f, 0.17392742f},
                                {0.11846344f, 0.23931434f, 0.28444444f, 0.23931434f, 0.11846344f},
                                {0.08566225f, 0.18038079f, 0.23395697f, 0.23395697f, 0.18038079f, 0.08566225f}};
  if (n < 1 || n > 6) return false;
  for (int i = 0; i < n; ++i) {
    mu.push_back(M[n - 1][i]);
    wmu.push_back(W[n - 1][i]);
  }
  return true;
}

// Lobatto with mu = 1 included, SRC/ray_mod.f90:296-327 (FP32 literals)
bool lobatto_table(int n, std::vector<double> &mu, std::vector<double> &wmu) {
  static const float M[6][6] = {{1.0f},
                                {1.0f, 0.44721360f},
                                {1.0f, 0.76505532f, 0.28523152f},
                                {1.0f, 0.87174015f, 0.59170018f, 0.20929922f},
                                {1.0f, 0.9195339082f, 0.7387738651f, 0.4779249498f, 0.1652789577f},
                                {1.0f, 0.944899272223f, 0.819279321644f, 0.632876153032f, 0.399530940965f,
                                 0.136552932855f}};
  static const float W[6][6] = {{1.0f},
                                {0.16666667f, 0.83333333f},
                                {0.06666667f, 0.37847496f, 0.55485838f},
                                {0.03571428f, 0.21070422f, 0.34112270f, 0.41245880f},
                                {0.0222222222f, 0.1333059908f, 0.2248893420f, 0.2920426836f, 0.3275397612f},
                                {0.015151515153f, 0.091684517440f, 0.157974705655f, 0.212508417834f,
                                 0.251275603210f, 0.271405240910f}};
  if (n < 1 || n > 6) return false;
  for (int i = 0; i < n; ++i) {
    mu.push_back(M[n - 1][i]);
    wmu.push_back(W[n - 1][i]);
  }
  return true;
}

// Legendre P_{n-1}(x), P_n(x) and P_n'(x) by the three-term recurrence
void legendre(int n, double x, double &pm1, double &p, double &dp) {
  double a = 1.0, b = x;  // P0, P1
  if (n == 0) {
    pm1 = 0.0;
    p = 1.0;
    dp = 0.0;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    const double c = ((2 * k - 1) * x * b - (k - 1) * a) / k;
    a = b;
    b = c;
  }
  pm1 = a;
  p = b;
  dp = n * (x * b - a) / (x * x - 1.0);
}

// Gauss-Radau rule of k points on [-1, 1] with the node -1: the others are the
// roots of (P_{k-1} + P_k)/(1 + x); weights 2/k^2 at -1 and
// (1 - x)/(k P_{k-1}(x))^2 elsewhere.  Ascending nodes.
void gauss_radau(int k, std::vector<double> &x, std::vector<double> &w) {
  x.assign(k, -1.0);
  w.assign(k, 2.0 / ((double)k * k));
  for (int i = 1; i < k; ++i) {
    // Chebyshev-Gauss-Radau start, then Newton on f = P_{k-1} + P_k
    double r = -std::cos(2.0 * M_PI * i / (2.0 * k - 1.0));
    for (int it = 0; it < 100; ++it) {
      double a1, b1, d1, a2, b2, d2;
      legendre(k - 1, r, a1, b1, d1);
      legendre(k, r, a2, b2, d2);
      const double f = b1 + b2, df = d1 + d2;
      // divide out the known root at -1: g = f / (1 + r), g' = (f' - g) / (1 + r)
      const double g = f / (1.0 + r), dg = (df - g) / (1.0 + r);
      const double step = g / dg;
      r -= step;
      if (std::fabs(step) < 1e-16) break;
    }
    double a1, b1, d1;
    legendre(k - 1, r, a1, b1, d1);
    x[i] = r;
    w[i] = (1.0 - r) / ((double)k * k * b1 * b1);
  }
  std::vector<int> ord(k);
  for (int i = 0; i < k; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return x[a] < x[b]; });
  std::vector<double> xs(k), ws(k);
  for (int i = 0; i < k; ++i) {
    xs[i] = x[ord[i]];
    ws[i] = w[ord[i]];
  }
  x = xs;
  w = ws;
}

// radau%init(k) mapped to (0, 1] with mu = 1 the fixed node (FP32, as there), then
// reversed to decreasing mu as set_radau_quadratures takes it
void radau_table(int k, std::vector<double> &mu, std::vector<double> &wmu) {
  std::vector<double> xn, wn;
  gauss_radau(k, xn, wn);
  const float c = 0.5f, d = 0.5f;
  std::vector<float> xi(k), wi(k);
  for (int i = 0; i < k; ++i) {
    // node i (1-based) takes the (k+1-i)-th ascending node, so x = 1 is node k
    xi[i] = (float)(1.0 - ((double)c + (double)d * xn[k - 1 - i]));
    wi[i] = (float)((double)d * wn[k - 1 - i]);
  }
  for (int i = k - 1; i >= 0; --i) {
    mu.push_back(xi[i]);
    wmu.push_back(wi[i]);
  }
}

// nphi_copies, SRC/ray_mod.f90:409-467
void nphi_copies(int nphi_in, bool one_d, const std::vector<double> &mu, const std::vector<double> &wmu,
                 const double *zrot, bool mirror, Rays &out) {
  const int nr = (int)mu.size();
  const bool mu1 = std::find(mu.begin(), mu.end(), 1.0) != mu.end();
  const int nphi = one_d ? 1 : nphi_in;
  int n1 = 0;
  if (mu1) {
    out.add(0.0, 0.0, 1.0, wmu[0]);
    n1 = 1;
  }
  for (int t = n1; t < nr; ++t) {
    const double xoff = zrot ? zrot[t] : 0.0;
    const double th = std::acos(mu[t]);
    for (int p = 0; p < nphi; ++p) {
      const double phi = 2.0 * kPi / (double)nphi * (double)p + xoff;
      out.add(std::sin(th) * std::cos(phi), std::sin(th) * std::sin(phi), mu[t], wmu[t] / (double)nphi);
    }
  }
  for (double &v : out.vec)
    if (std::fabs(v) < 1e-3) v = 0.0;
  const size_t n = out.n();
  if (mirror) {
    for (size_t i = 0; i < n; ++i) out.add(-out.vec[3 * i], -out.vec[3 * i + 1], -out.vec[3 * i + 2], out.w[i]);
    for (double &w : out.w) w /= 2.0;
  } else {
    for (size_t i = 0; i < n; ++i) out.w[i] *= out.vec[3 * i + 2];
  }
}

// Simpson weights on increasing nodes (odd count), SRC/ray_mod.f90:514-540
void simpson(const double *x, int n, double *w) {
  for (int i = 0; i < n; ++i) w[i] = 0.0;
  if (n == 1) {
    w[0] = 1.0;
    return;
  }
  for (int i = 0; i + 2 < n; i += 2) {
    const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
    w[i] += (h0 + h1) / 6 * (2 - h1 / h0);
    w[i + 1] += (h0 + h1) * (h0 + h1) * (h0 + h1) / (6 * h0 * h1);
    w[i + 2] += (h0 + h1) / 6 * (2 - h0 / h1);
  }
}

// set_custom_quadratures, SRC/ray_mod.f90:380-404: the mu weights integrate over
// [0, 1] with a node at 0 that gets no ray; flux weights
int custom(std::vector<double> mu, int nphi, bool one_d, std::vector<double> *zrot, Rays &out) {
  const int n = (int)mu.size();
  std::vector<double> wmu(n, 1.0);
  if (n > 1) {
    // ascending with the permutation (the reference's bubble sort is stable)
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return mu[a] < mu[b]; });
    std::vector<double> s(n + 1), sw(n + 1, 0.0);
    s[0] = 0.0;
    for (int i = 0; i < n; ++i) s[i + 1] = mu[ord[i]];
    for (int i = 1; i <= n; ++i)
      if (!(s[i] > s[i - 1])) return set_error(M3DRT_EINVAL, "m3drt_quadrature: custom mu must be distinct and > 0");
    if ((n + 1) % 2 == 1) {
      simpson(s.data(), n + 1, sw.data());
    } else {
      // the mean of two Simpson rules, each closed by a trapezoid at one end
      std::vector<double> sw2(n);
      simpson(s.data(), n, sw.data());
      simpson(s.data() + 1, n, sw2.data());
      sw[0] += 0.5 * (s[1] - s[0]);
      sw[1] += 0.5 * (s[1] - s[0]);
      sw2[n - 2] += 0.5 * (s[n] - s[n - 1]);
      sw2[n - 1] += 0.5 * (s[n] - s[n - 1]);
      for (int i = 1; i <= n; ++i) sw[i] += sw2[i - 1];
      for (double &v : sw) v /= 2;
    }
    for (int i = 0; i < n; ++i) {
      wmu[i] = sw[n - i];
      mu[i] = s[n - i];
    }
    if (zrot) {
      std::vector<double> z(n);
      for (int i = 0; i < n; ++i) z[i] = (*zrot)[ord[n - 1 - i]];
      *zrot = z;
    }
  }
  nphi_copies(nphi, one_d, mu, wmu, zrot ? zrot->data() : nullptr, false, out);
  return M3DRT_OK;
}

// set_carlson_quadratures, SRC/ray_mod.f90:36-181: one octant tabulated (FP32
// literals), reflected into the four upward octants, mirrored for J
bool carlson(const std::string &s, bool mirror, Rays &out) {
  struct Set {
    const char *name;
    int n;
    float mx[10], my[10], w[10];
  };
  static const Set sets[] = {
      {"set_a2", 1, {0.57735026f}, {0.57735026f}, {1.00000000f}},
      {"set_a4", 3, {0.88191710f, 0.33333333f, 0.33333333f}, {0.33333333f, 0.88191710f, 0.33333333f},
       {0.33333333f, 0.33333333f, 0.33333333f}},
      {"set_a6", 6, {0.93094934f, 0.68313005f, 0.25819889f, 0.68313005f, 0.25819889f, 0.25819889f},
       {0.25819889f, 0.68313005f, 0.93094934f, 0.25819889f, 0.68313005f, 0.25819889f},
       {0.18333333f, 0.15000000f, 0.18333333f, 0.15000000f, 0.15000000f, 0.18333333f}},
      {"set_a8", 10,
       {0.95118973f, 0.78679579f, 0.57735027f, 0.21821789f, 0.78679579f, 0.57735027f, 0.21821789f, 0.57735027f,
        0.21821789f, 0.21821789f},
       {0.21821789f, 0.57735027f, 0.78679579f, 0.95118973f, 0.21821789f, 0.57735027f, 0.78679579f, 0.21821789f,
        0.57735027f, 0.21821789f},
       {0.12698138f, 0.09138353f, 0.09138353f, 0.12698138f, 0.09138353f, 0.07075469f, 0.09138353f, 0.09138353f,
        0.09138353f, 0.12698138f}},
      {"set_b4", 3, {0.70412415f, 0.09175171f, 0.09175171f}, {0.09175171f, 0.70412415f, 0.09175171f},
       {0.33333333f, 0.33333333f, 0.33333333f}},
      {"set_b6", 6, {0.80847426f, 0.57735027f, 0.11417547f, 0.57735027f, 0.11417547f, 0.11417547f},
       {0.11417547f, 0.57735027f, 0.80847426f, 0.11417547f, 0.57735027f, 0.11417547f},
       {0.22222222f, 0.11111111f, 0.22222222f, 0.11111111f, 0.11111111f, 0.22222222f}},
      {"set_b8", 10,
       {0.85708018f, 0.70273364f, 0.50307327f, 0.11104440f, 0.70273364f, 0.50307327f, 0.11104440f, 0.50307327f,
        0.11104440f, 0.11104440f},
       {0.11104440f, 0.50307327f, 0.70273364f, 0.85708018f, 0.11104440f, 0.50307327f, 0.70273364f, 0.11104440f,
        0.50307327f, 0.11104440f},
       {0.17152232f, 0.06900520f, 0.06900520f, 0.17152232f, 0.06900520f, 0.07140185f, 0.06900520f, 0.06900520f,
        0.06900520f, 0.17152232f}},
  };
  if (s == "disk_center") {
    out.add(0.0, 0.0, 1.0, mirror ? 0.5 : 1.0);
    if (mirror) out.add(0.0, 0.0, -1.0, 0.5);
    return true;
  }
  const Set *set = nullptr;
  for (const Set &c : sets)
    if (s == c.name) set = &c;
  if (!set) return false;
  const int n = set->n;
  // octant 1: (-mux, -muy, muz); 2: x reflected; 3, 4: y reflected of 1, 2
  for (int o = 0; o < 4; ++o)
    for (int i = 0; i < n; ++i) {
      const double mx = set->mx[i], my = set->my[i];
      const double sx = (o == 0 || o == 3) ? -1.0 : 1.0, sy = o < 2 ? -1.0 : 1.0;
      out.add(sx * mx, sy * my, std::sqrt(1.0 - mx * mx - my * my), set->w[i]);
    }
  const size_t m = out.n();
  if (mirror) {
    for (size_t i = 0; i < m; ++i) out.add(-out.vec[3 * i], -out.vec[3 * i + 1], -out.vec[3 * i + 2], out.w[i]);
    for (double &w : out.w) w /= 8;
  } else {
    for (size_t i = 0; i < m; ++i) out.w[i] = out.w[i] / 4 * out.vec[3 * i + 2];
  }
  return true;
}

}  // namespace
}  // namespace m3drt

extern "C" int m3drt_quadrature(const m3drt_quadrature_args *q, int32_t cap, double *vec, double *wts,
                                int32_t *nray) {
  using namespace m3drt;
  if (!q || !q->scheme || !nray) return set_error(M3DRT_EINVAL, "m3drt_quadrature: NULL argument");
  const std::string s(q->scheme);
  const bool mirror = q->jnu_weights != 0, one_d = q->one_dimensional != 0;
  Rays r;
  std::vector<double> mu, wmu;
  if (s == "radau") {
    if (q->ntheta < 1 || q->ntheta > 10) return set_error(M3DRT_EINVAL, "m3drt_quadrature: radau takes 1 to 10 inclinations");
    radau_table(q->ntheta, mu, wmu);
    std::vector<double> zrot(q->ntheta, kPi / 4.);
    nphi_copies(q->nphi, one_d, mu, wmu, zrot.data(), mirror, r);
  } else if (s == "lobatto" || s == "gauss") {
    const bool ok = s == "gauss" ? gauss_table(q->ntheta, mu, wmu) : lobatto_table(q->ntheta, mu, wmu);
    if (!ok) return set_error(M3DRT_EINVAL, "m3drt_quadrature: " + s + " takes 1 to 6 inclinations");
    nphi_copies(q->nphi, one_d, mu, wmu, nullptr, mirror, r);
  } else if (s == "custom") {
    if (q->ntheta < 1 || !q->mu) return set_error(M3DRT_EINVAL, "m3drt_quadrature: custom needs ntheta >= 1 mu values");
    if (mirror) return set_error(M3DRT_EINVAL, "m3drt_quadrature: custom is an output-pass scheme (flux weights)");
    std::vector<double> m(q->mu, q->mu + q->ntheta);
    const int rc = custom(m, q->nphi, one_d, nullptr, r);
    if (rc) return rc;
  } else if (s == "grid+" || s == "gridx" || s == "grid*") {
    // SRC/ray_mod.f90:188-216: the vertical and the rays that reach the next column
    // after one cell in z, four azimuths each, flux weights
    if (mirror) return set_error(M3DRT_EINVAL, "m3drt_quadrature: grid schemes are output-pass schemes");
    if (!(q->ds[0] > 0 && q->ds[1] > 0 && q->ds[2] > 0)) return set_error(M3DRT_EINVAL, "m3drt_quadrature: ds must be > 0");
    const double plus = q->ds[0] / q->ds[2];
    const double cross = std::sqrt(q->ds[0] * q->ds[0] + q->ds[1] * q->ds[1]) / q->ds[2];
    std::vector<double> ang, zr;
    if (s == "grid+") {
      ang = {0.0, std::atan(plus)};
      zr = {0.0, 0.0};
    } else if (s == "gridx") {
      ang = {0.0, std::atan(cross)};
      zr = {0.0, kPi / 4.};
    } else {
      ang = {0.0, std::atan(plus), std::atan(cross)};
      zr = {0.0, 0.0, kPi / 4.};
    }
    for (double &a : ang) a = std::cos(a);
    const int rc = custom(ang, 4, one_d, &zr, r);
    if (rc) return rc;
  } else if (s == "none") {
  } else if (!carlson(s, mirror, r)) {
    return set_error(M3DRT_EINVAL, "m3drt_quadrature: quadrature keyword \"" + s + "\" not recognized");
  }
  *nray = (int32_t)r.n();
  if ((int64_t)r.n() > cap || (r.n() && (!vec || !wts)))
    return set_error(M3DRT_ELIMIT, "m3drt_quadrature: " + std::to_string(r.n()) + " rays, room for " + std::to_string(cap));
  if (r.n()) {
    std::memcpy(vec, r.vec.data(), sizeof(double) * r.vec.size());
    std::memcpy(wts, r.w.data(), sizeof(double) * r.w.size());
  }
  return M3DRT_OK;
}
