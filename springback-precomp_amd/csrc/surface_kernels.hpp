// The released part's height map (the reference's formed_surface before its
// nearest fill): a structured surface cast on a regular grid.
#pragma once

#include <hip/hip_runtime.h>

namespace sbk {

struct HeightGrid {
  double x0, y0, h;
  long long nx, ny;
};

// folded (device, zeroed by the caller) += the triangles of zero or
// negative plan area of the surface (ny_s, nx_s, 3).
int launch_surface_folds(const double* surface, long long nxs, long long nys, unsigned long long* folded,
                         hipStream_t s);
// z (ny nx): the highest hit per grid node, -inf where nothing is hit; hit:
// 1 where hit.
int launch_surface_raycast(const double* surface, long long nxs, long long nys, const HeightGrid& g, double* z,
                           unsigned char* hit, hipStream_t s);

}  // namespace sbk
