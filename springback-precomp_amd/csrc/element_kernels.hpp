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
  double* tangent;    // 576 per element (row-major 24 x 24), or null
  double* energy;     // per element
  int* yielding;      // per element
  int* iterations;    // local Newton iterations per element
  int* status;        // per element: 0 or an ElementStatus
  double* fail_value; // per element: the determinant of a refused Jacobian
};

int launch_local_solve(const ElementArgs& a, hipStream_t stream);
int launch_condense(const ElementArgs& a, bool tangent, hipStream_t stream);

}  // namespace sbk
