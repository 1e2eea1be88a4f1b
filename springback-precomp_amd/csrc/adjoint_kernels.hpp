// Launch arguments of the adjoint (vector-Jacobian product) kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "element_kernels.hpp"

namespace sbk {

// Material parameters of the adjoint: r0, r45, r90, then (C_i, gamma_i) of
// the backstresses 1..4.
constexpr int kAdjointParams = 11;

enum AdjointWrt { kWrtCoords = 1, kWrtDisplacement = 2, kWrtParams = 4 };

struct AdjointArgs {
  ElementArgs e;        // the evaluation's launch record: mesh, u, history, converged alpha, workspace
  const double* v;      // cotangent of the assembled internal force, 3 per node
  int wrt;              // AdjointWrt flags
  int hill;             // dP holds dP/d(r0, r45, r90)
  int backstresses;     // backstresses whose C, gamma are parameters
  double* grad_x;       // 24 per element: v_e^T d f*_e / d X_e
  double* grad_u;       // 24 per element: v_e^T d f*_e / d u_e (= v_e^T K*_e)
  double* grad_p;       // [param][element]: v_e^T d f*_e / d theta
  double dP[3][36];
};

int launch_adjoint(const AdjointArgs& a, hipStream_t stream);

}  // namespace sbk
