// The released part's height map: the deformed tool-side surface, a
// structured (ny_s, nx_s) grid of nodes cut into two triangles per cell,
// cast along vertical rays through the nodes of a regular grid
// (python/precomp/fea/results.py:262-283, _interpolate_structured, and
// python/precomp/geometry/raycast.py:27-100, raycast_top; the nearest fill of
// the nodes outside the surface stays with the caller).
//
// One thread per triangle tests the grid nodes of its plan bounding box, as
// the reference's (triangle, node) pairs. The highest hit per node is an
// atomic maximum on an order-preserving 64-bit key of the height, kept in
// z's own memory and decoded in place: a maximum does not depend on the
// order of the hits, so the map is deterministic. The arithmetic keeps
// numpy's order of operations (separately rounded products and sums, no
// contraction), so the heights and the inside tests are the reference's bit
// for bit.

#include "surface_kernels.hpp"

#include <cmath>

// numpy's order of operations, each rounded on its own: no product and sum
// below is fused into one multiply-add (the expressions are written in
// numpy's association and this pragma keeps each operation rounded)
#pragma clang fp contract(off)

namespace sbk {

namespace {

constexpr unsigned long long kSign = 0x8000000000000000ull;

// Doubles to unsigned keys of the same order (no NaN here).
__device__ __forceinline__ unsigned long long height_key(double z) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(z));
  return (b & kSign) ? ~b : (b | kSign);
}

__device__ __forceinline__ double key_height(unsigned long long k) {
  const unsigned long long b = (k & kSign) ? (k & ~kSign) : ~k;
  return __longlong_as_double(static_cast<long long>(b));
}

struct Triangle {
  const double* c[3];
};

// Triangle t: (a, b, c) of cell t for t < ncell, else (a, c, d) of cell
// t - ncell, with a, b, c, d the cell's corners counterclockwise from
// (cx, cy).
__device__ __forceinline__ Triangle triangle(const double* p, long long nxs, long long ncell, long long t) {
  const long long cell = t < ncell ? t : t - ncell;
  const long long cx = cell % (nxs - 1), cy = cell / (nxs - 1);
  Triangle r;
  r.c[0] = p + 3 * (cy * nxs + cx);
  if (t < ncell) {
    r.c[1] = p + 3 * (cy * nxs + cx + 1);
    r.c[2] = p + 3 * ((cy + 1) * nxs + cx + 1);
  } else {
    r.c[1] = p + 3 * ((cy + 1) * nxs + cx + 1);
    r.c[2] = p + 3 * ((cy + 1) * nxs + cx);
  }
  return r;
}

__global__ __launch_bounds__(256) void fold_kernel(const double* __restrict__ p, long long nxs, long long ncell,
                                                   unsigned long long* __restrict__ folded) {
  const long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= 2 * ncell) return;
  const Triangle tr = triangle(p, nxs, ncell, t);
  const double e1x = (tr.c[1][0] - tr.c[0][0]), e1y = (tr.c[1][1] - tr.c[0][1]);
  const double e2x = (tr.c[2][0] - tr.c[0][0]), e2y = (tr.c[2][1] - tr.c[0][1]);
  if (((e1x * e2y) - (e1y * e2x)) <= 0.0) atomicAdd(folded, 1ull);
}

__global__ __launch_bounds__(256) void clear_heights_kernel(unsigned long long* __restrict__ key, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) key[i] = height_key(-INFINITY);
}

__device__ __forceinline__ double clip(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ __launch_bounds__(256) void raycast_kernel(const double* __restrict__ p, long long nxs, long long ncell,
                                                      HeightGrid g, unsigned long long* __restrict__ key) {
  const long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= 2 * ncell) return;
  const Triangle tr = triangle(p, nxs, ncell, t);
  const double* a = tr.c[0];
  const double* b = tr.c[1];
  const double* c = tr.c[2];
  const double v0x = (b[0] - a[0]), v0y = (b[1] - a[1]);
  const double v1x = (c[0] - a[0]), v1y = (c[1] - a[1]);
  const double den = ((v0x * v1y) - (v1x * v0y));
  const double scale = fmax(fmax(fabs(v0x), fabs(v0y)), fmax(fabs(v1x), fabs(v1y)));
  const double sc = fmax(scale, 1e-300);
  if (!(fabs(den) > (1e-12 * (sc * sc)))) return;  // seen edge-on from above
  const double xmax = (g.x0 + (g.h * static_cast<double>(g.nx - 1)));
  const double ymax = (g.y0 + (g.h * static_cast<double>(g.ny - 1)));
  const double xlo = fmin(fmin(a[0], b[0]), c[0]), xhi = fmax(fmax(a[0], b[0]), c[0]);
  const double ylo = fmin(fmin(a[1], b[1]), c[1]), yhi = fmax(fmax(a[1], b[1]), c[1]);
  if (xhi < g.x0 || xlo > xmax || yhi < g.y0 || ylo > ymax) return;
  const double eps = 1e-6, tol = 1e-6;
  const long long i0 = static_cast<long long>(
      clip(ceil((((xlo - g.x0) / g.h) - eps)), 0.0, static_cast<double>(g.nx - 1)));
  const long long i1 = static_cast<long long>(
      clip(floor((((xhi - g.x0) / g.h) + eps)), -1.0, static_cast<double>(g.nx - 1)));
  const long long j0 = static_cast<long long>(
      clip(ceil((((ylo - g.y0) / g.h) - eps)), 0.0, static_cast<double>(g.ny - 1)));
  const long long j1 = static_cast<long long>(
      clip(floor((((yhi - g.y0) / g.h) + eps)), -1.0, static_cast<double>(g.ny - 1)));
  const double dz1 = (b[2] - a[2]), dz2 = (c[2] - a[2]);
  for (long long jj = j0; jj <= j1; ++jj)
    for (long long ii = i0; ii <= i1; ++ii) {
      const double px = ((g.x0 + (static_cast<double>(ii) * g.h)) - a[0]);
      const double py = ((g.y0 + (static_cast<double>(jj) * g.h)) - a[1]);
      const double u = (((px * v1y) - (v1x * py)) / den);
      const double v = (((v0x * py) - (px * v0y)) / den);
      if (!(u >= -tol && v >= -tol && (u + v) <= 1.0 + tol)) continue;
      const double zz = ((a[2] + (u * dz1)) + (v * dz2));
      atomicMax(key + (jj * g.nx + ii), height_key(zz));
    }
}

__global__ __launch_bounds__(256) void decode_heights_kernel(unsigned long long* __restrict__ key, long long n,
                                                             unsigned char* __restrict__ hit) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = key[i];
  hit[i] = k != height_key(-INFINITY) ? 1 : 0;
  reinterpret_cast<double*>(key)[i] = key_height(k);
}

unsigned int grid_of(long long n) { return static_cast<unsigned int>((n + 255) / 256); }

}  // namespace

int launch_surface_folds(const double* surface, long long nxs, long long nys, unsigned long long* folded,
                         hipStream_t s) {
  (void)hipGetLastError();
  const long long ncell = (nxs - 1) * (nys - 1);
  if (ncell > 0)
    hipLaunchKernelGGL(fold_kernel, dim3(grid_of(2 * ncell)), dim3(256), 0, s, surface, nxs, ncell, folded);
  return static_cast<int>(hipGetLastError());
}

int launch_surface_raycast(const double* surface, long long nxs, long long nys, const HeightGrid& g, double* z,
                           unsigned char* hit, hipStream_t s) {
  (void)hipGetLastError();
  const long long n = g.nx * g.ny;
  const long long ncell = (nxs - 1) * (nys - 1);
  unsigned long long* key = reinterpret_cast<unsigned long long*>(z);
  hipLaunchKernelGGL(clear_heights_kernel, dim3(grid_of(n)), dim3(256), 0, s, key, n);
  if (ncell > 0)
    hipLaunchKernelGGL(raycast_kernel, dim3(grid_of(2 * ncell)), dim3(256), 0, s, surface, nxs, ncell, g, key);
  hipLaunchKernelGGL(decode_heights_kernel, dim3(grid_of(n)), dim3(256), 0, s, key, n, hit);
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
