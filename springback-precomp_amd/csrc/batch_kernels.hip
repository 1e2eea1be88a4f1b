// Per-problem scalars of a batched evaluation: the reference's per-system
// Evaluation scalars (src/fem/NonlinearSystem.hpp:27-47, 239-268) for every
// problem of the batch at once, one workgroup per problem.
#include "batch_kernels.hpp"

namespace sbk {

// Each sum runs in a fixed order (thread t takes the segment's items t,
// t + 256, ...; then a fixed tree), so the result does not depend on timing.
__global__ __launch_bounds__(256) void problem_scalars_kernel(ProblemScalarArgs A) {
  __shared__ double se[256], sg[256];
  __shared__ long long sy[256], sf[256], sm[256];
  const int p = blockIdx.x;
  const int t = threadIdx.x;
  const long long e0 = A.element_offset[p], e1 = A.element_offset[p + 1];
  const long long d0 = 3 * A.node_offset[p], d1 = 3 * A.node_offset[p + 1];
  double energy = 0.0, gross = 0.0;
  long long yielding = 0, failed = 0, first = 0x7fffffffffffffffll;
  for (long long e = e0 + t; e < e1; e += 256) {
    if (A.status[e] != 0) {
      ++failed;
      if (e - e0 < first) first = e - e0;
    } else {
      yielding += A.yielding[e];
      energy = energy + A.energy[e];
    }
  }
  for (long long d = d0 + t; d < d1; d += 256) gross = gross + A.gross_dof[d] * A.gross_dof[d];
  se[t] = energy;
  sg[t] = gross;
  sy[t] = yielding;
  sf[t] = failed;
  sm[t] = first;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      se[t] = se[t] + se[t + w];
      sg[t] = sg[t] + sg[t + w];
      sy[t] += sy[t + w];
      sf[t] += sf[t + w];
      sm[t] = sm[t] < sm[t + w] ? sm[t] : sm[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    A.energy_out[p] = se[0];
    A.gross_out[p] = sqrt(sg[0]);
    A.yielding_out[p] = sy[0];
    A.failed_out[p] = sf[0];
    A.first_failed_out[p] = sf[0] > 0 ? sm[0] : -1;
  }
}

int launch_problem_scalars(const ProblemScalarArgs& a, int num_problems, hipStream_t stream) {
  (void)hipGetLastError();
  if (num_problems <= 0) return 0;
  hipLaunchKernelGGL(problem_scalars_kernel, dim3(num_problems), dim3(256), 0, stream, a);
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
