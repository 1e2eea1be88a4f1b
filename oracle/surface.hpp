// Oracle (test infrastructure only): the released part's height map.
//
//   * python/precomp/fea/results.py:262-283 (_interpolate_structured: two
//     triangles per cell of the structured top surface, (a, b, c) then
//     (a, c, d); a triangle of zero or negative plan area refuses the surface)
//   * python/precomp/geometry/raycast.py:27-100 (raycast_top: each kept
//     triangle against the grid nodes of its plan bounding box, barycentric
//     coordinates with a 1e-6 tolerance, the highest hit per node)
// The arithmetic keeps numpy's order of operations (no contractions): this
// file is compiled with -ffp-contract=off.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace oracle {

struct HeightGrid {
  double x0, y0, h;
  long long nx, ny;
};

// Surface positions p (ny_s, nx_s, 3). z (ny nx) gets the highest hit or
// -inf, hit the flags; returns the count of triangles with plan area <= 0
// (the caller refuses such a surface; nothing is cast then).
inline long long surface_raycast(const double* p, long long nxs, long long nys, const HeightGrid& g, double* z,
                                 unsigned char* hit) {
  const long long ncell = (nxs - 1) * (nys - 1);
  const long long ntri = 2 * ncell;
  auto corner = [&](long long cx, long long cy, int k) { return p + 3 * (cy * nxs + cx) + k; };
  auto tri = [&](long long t, const double* c[3]) {
    const long long cell = t % ncell;
    const long long cx = cell % (nxs - 1), cy = cell / (nxs - 1);
    const double* a = corner(cx, cy, 0);
    const double* b = corner(cx + 1, cy, 0);
    const double* cc = corner(cx + 1, cy + 1, 0);
    const double* d = corner(cx, cy + 1, 0);
    c[0] = a;
    if (t < ncell) {
      c[1] = b;
      c[2] = cc;
    } else {
      c[1] = cc;
      c[2] = d;
    }
  };
  long long folded = 0;
  for (long long t = 0; t < ntri; ++t) {
    const double* c[3];
    tri(t, c);
    const double e1x = c[1][0] - c[0][0], e1y = c[1][1] - c[0][1];
    const double e2x = c[2][0] - c[0][0], e2y = c[2][1] - c[0][1];
    if (e1x * e2y - e1y * e2x <= 0.0) ++folded;
  }
  for (long long i = 0; i < g.nx * g.ny; ++i) {
    z[i] = -std::numeric_limits<double>::infinity();
    hit[i] = 0;
  }
  if (folded > 0) return folded;
  const double xmax = g.x0 + g.h * static_cast<double>(g.nx - 1);
  const double ymax = g.y0 + g.h * static_cast<double>(g.ny - 1);
  const double eps = 1e-6, tol = 1e-6;
  for (long long t = 0; t < ntri; ++t) {
    const double* c[3];
    tri(t, c);
    const double* a = c[0];
    const double* b = c[1];
    const double* cc = c[2];
    const double v0x = b[0] - a[0], v0y = b[1] - a[1];
    const double v1x = cc[0] - a[0], v1y = cc[1] - a[1];
    const double den = v0x * v1y - v1x * v0y;
    const double scale = std::max(std::max(std::fabs(v0x), std::fabs(v0y)), std::max(std::fabs(v1x), std::fabs(v1y)));
    const double sc = std::max(scale, 1e-300);
    if (!(std::fabs(den) > 1e-12 * (sc * sc))) continue;
    const double xmin_t = std::min(std::min(a[0], b[0]), cc[0]), xmax_t = std::max(std::max(a[0], b[0]), cc[0]);
    const double ymin_t = std::min(std::min(a[1], b[1]), cc[1]), ymax_t = std::max(std::max(a[1], b[1]), cc[1]);
    if (xmax_t < g.x0 || xmin_t > xmax || ymax_t < g.y0 || ymin_t > ymax) continue;
    auto clip = [](double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); };
    const long long i0 = static_cast<long long>(clip(std::ceil((xmin_t - g.x0) / g.h - eps), 0.0, double(g.nx - 1)));
    const long long i1 = static_cast<long long>(clip(std::floor((xmax_t - g.x0) / g.h + eps), -1.0, double(g.nx - 1)));
    const long long j0 = static_cast<long long>(clip(std::ceil((ymin_t - g.y0) / g.h - eps), 0.0, double(g.ny - 1)));
    const long long j1 = static_cast<long long>(clip(std::floor((ymax_t - g.y0) / g.h + eps), -1.0, double(g.ny - 1)));
    for (long long jj = j0; jj <= j1; ++jj)
      for (long long ii = i0; ii <= i1; ++ii) {
        const double px = (g.x0 + static_cast<double>(ii) * g.h) - a[0];
        const double py = (g.y0 + static_cast<double>(jj) * g.h) - a[1];
        const double u = (px * v1y - v1x * py) / den;
        const double v = (v0x * py - px * v0y) / den;
        if (!(u >= -tol && v >= -tol && u + v <= 1.0 + tol)) continue;
        const double zz = (a[2] + u * (b[2] - a[2])) + v * (cc[2] - a[2]);
        const long long k = jj * g.nx + ii;
        if (zz > z[k]) z[k] = zz;
        hit[k] = 1;
      }
  }
  return 0;
}

}  // namespace oracle
