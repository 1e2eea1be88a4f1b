// Launch arguments of the batched form's per-problem reductions.
#pragma once

#include <hip/hip_runtime.h>

namespace sbk {

struct ProblemScalarArgs {
  const long long* node_offset;     // num_problems + 1
  const long long* element_offset;  // num_problems + 1
  const double* gross_dof;          // per DOF: sum over elements of |f_e|
  const double* energy;             // per element
  const int* status;                // per element
  const int* yielding;              // per element
  double* energy_out;               // per problem: energy of the non-failed elements
  double* gross_out;                // per problem: |sum_e |f_e||
  long long* yielding_out;          // per problem: yielding points of the non-failed elements
  long long* failed_out;            // per problem: failed elements
  long long* first_failed_out;      // per problem: lowest failed element (local index) or -1
};

int launch_problem_scalars(const ProblemScalarArgs& a, int num_problems, hipStream_t stream);

}  // namespace sbk
