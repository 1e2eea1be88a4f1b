// Launch arguments of the element kernels (device pointers, sizes, the
// material and the integration rule by value).
#pragma once

#include <hip/hip_runtime.h>

#include "device_math.hpp"

namespace sbk {

struct RuleArg {
  double xi[kMaxPoints], eta[kMaxPoints], zeta[kMaxPoints], w[kMaxPoints];
  int n;
};

struct ElementArgs {
  // mesh and displacement (node-major xyz; conn: 8 nodes per element, VTK order)
  const double* coords;
  const int* conn;
  const double* u;
  long long ne;
  RuleArg rule;
  int kinematics;  // 0 small strain, 1 finite (Green-Lagrange), 2 logarithmic
  DevMaterial mat;
  // committed history: state_in[f * state_stride + e * rule.n + p]
  const double* state_in;
  long long state_stride;
  const double* alpha_in;  // committed modes, 9 per element, or null (zero)
  // outputs
  double* alpha_out;  // 9 per element
  double* state_out;  // trial history, same layout as state_in
  double* force;      // 24 per element
  double* tangent;    // 576 per element (24 x 24), or null
  int tangent_column_major;  // the tangent's layout: 0 row-major, 1 column-major
  double* energy;     // per element
  int* yielding;      // per element
  int* iterations;    // local Newton iterations per element
  int* status;        // per element: 0 or an ElementStatus
  double* fail_value; // per element: the determinant of a refused Jacobian
  int* local_path;    // per element (optional): the local iteration's phases (kPath* bits)
  // Workspace (device scratch of one call), three regions:
  //   point_ws  [npts][42]: S (6), then the tangent (36, row-major) of each
  //             point's final local pass (written by the local solve)
  //   geom_ws   [npts][44]: reserved (the condensation keeps each point's
  //             F (9), [G | G~] (33) and w at the converged alpha in LDS)
  //   elem_ws   [ne][108]:  K_aa's full-pivot LU factors (81, row-major),
  //             r_a (9), row and column transpositions (9 + 9, as doubles)
  //             of the final local iteration
  // with npts = ne * rule.n, point q = e * rule.n + p.
  double* point_ws;
  double* geom_ws;
  double* elem_ws;
  // and the schedule of the local solve, as int32:
  //   cost      [ne]: the expected work of each element (cost_key_kernel)
  //   order     [ne]: the elements in the order the local solve takes them
  //   counts    [2 * kCostClasses]: per-class counts, then cursors
  int* cost;
  int* order;
  int* counts;
  // elements per wavefront of the local solve (a multiple of 8; 0: 64, or
  // fewer for a launch too small to fill the device)
  int chunk;
  // 1: the local solve takes the elements in `order` (most work first);
  // 0: in index order
  int ordered;
};

constexpr int kPointRecord = 42;
constexpr int kGeomRecord = 44;
constexpr int kElemRecord = 108;
// Element cost classes: sum over the points of return_cost (0, 1 or 2).
constexpr int kCostClasses = 2 * kMaxPoints + 1;
// Doubles of workspace for ne elements of `points` points each.
inline long long workspace_doubles(long long ne, int points) {
  const long long ints = 2 * ne + 2 * kCostClasses;
  return ne * points * (kPointRecord + kGeomRecord) + ne * kElemRecord + (ints + 1) / 2;
}
// Points the regions of ElementArgs into one workspace.
inline void set_workspace(ElementArgs& a, double* ws) {
  const long long npts = a.ne * a.rule.n;
  a.point_ws = ws;
  a.geom_ws = ws + npts * kPointRecord;
  a.elem_ws = ws + npts * (kPointRecord + kGeomRecord);
  int* sched = reinterpret_cast<int*>(a.elem_ws + a.ne * kElemRecord);
  a.cost = sched;
  a.order = sched + a.ne;
  a.counts = sched + 2 * a.ne;
}

int launch_local_solve(const ElementArgs& a, hipStream_t stream);
int launch_condense(const ElementArgs& a, bool tangent, hipStream_t stream);
// incompatible_elastoplastic_stress of every element at state_in / alpha_in:
// kStressFields per element into out; status (if set) per element.
constexpr int kStressFields = 28;
int launch_element_stress(const ElementArgs& a, double* out, hipStream_t stream);

}  // namespace sbk
