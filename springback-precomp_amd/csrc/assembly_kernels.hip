// Atomic-free assembly on gfx950: the gather form of the reference's
// element-order scatter (src/fem/Assembler.cpp:183-215) and force gather
// (src/fem/NonlinearSystem.hpp:240-262).
//
// assemble_values_kernel: one thread per (3 x 3 block, block row). It sums the
//   block's contributions in their list order (ascending element), so every
//   value is the reference's sum, added in the reference's order, bit for bit.
//   Three consecutive doubles per thread; consecutive threads write
//   consecutive blocks of a row, so the stores coalesce. HBM-bound: each
//   element matrix entry is read once and each value written once.
// gather_force_kernel: one thread per DOF: f_int and the gross |f_e| sum.
// sum_squares_kernel: a fixed-shape tree for the norms (deterministic).
#include "assembly_kernels.hpp"

namespace sbk {

__global__ __launch_bounds__(256) void assemble_values_kernel(AssemblyArgs A) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= 3ll * A.num_blocks) return;
  const int blk = static_cast<int>(t / 3);
  const int j = static_cast<int>(t - 3ll * blk);
  const int c0 = A.block_ptr[blk], c1 = A.block_ptr[blk + 1];
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  for (int c = c0; c < c1; ++c) {
    const int code = A.block_contrib[c];
    const long long e = code >> 6;
    const int al = (code >> 3) & 7, bl = code & 7;
    const double* k = A.ke + e * 576;
    if (!A.column_major) {
      // row 3 al + j, columns 3 bl + 0..2 (contiguous)
      const double* r = k + (3 * al + j) * 24 + 3 * bl;
      v0 = v0 + r[0];
      v1 = v1 + r[1];
      v2 = v2 + r[2];
    } else {
      // the reference's column-major values: column 3 al + j, rows 3 bl + 0..2
      const double* r = k + (3 * bl) * 24 + 3 * al + j;
      v0 = v0 + r[0];
      v1 = v1 + r[24];
      v2 = v2 + r[48];
    }
  }
  double* out = A.values + A.block_pos[blk] + static_cast<long long>(j) * A.block_row_len[blk];
  out[0] = v0;
  out[1] = v1;
  out[2] = v2;
}

__global__ __launch_bounds__(256) void gather_force_kernel(ForceArgs A) {
  const long long dof = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (dof >= 3ll * A.num_nodes) return;
  const long long n = dof / 3;
  const int k = static_cast<int>(dof - 3 * n);
  double f = 0.0, g = 0.0;
  for (int i = A.node_ptr[n]; i < A.node_ptr[n + 1]; ++i) {
    const int code = A.node_incidence[i];
    const double v = A.fe[static_cast<long long>(code >> 3) * 24 + 3 * (code & 7) + k];
    f = f + v;
    g = g + fabs(v);
  }
  A.fint[dof] = f;
  A.gross[dof] = g;
}

// Sums (of squares, or of values) of x[0, n) in a fixed order: each block of
// the grid sums a contiguous slice with a fixed in-block tree and writes a
// partial; one block then sums the partials the same way.
__global__ __launch_bounds__(256) void sum_partials_kernel(const double* __restrict__ x, long long n,
                                                           double* __restrict__ partial, int square) {
  __shared__ double s[256];
  const long long per = (n + gridDim.x - 1) / gridDim.x;
  const long long b0 = static_cast<long long>(blockIdx.x) * per;
  const long long b1 = b0 + per < n ? b0 + per : n;
  double acc = 0.0;
  for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) acc = acc + (square ? x[i] * x[i] : x[i]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void sum_kernel(const double* __restrict__ x, int n, double* __restrict__ out,
                                                  int take_sqrt) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc = acc + x[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(s[0]) : s[0];
}

// Failure summary of an element batch: count, lowest failed element, its
// status; the yielding points' total (integers: order-independent).
__global__ __launch_bounds__(256) void status_init_kernel(long long* summary) {
  if (threadIdx.x == 0) {
    summary[0] = 0;
    summary[1] = 0x7fffffffffffffffll;
    summary[2] = 0;
  }
}
__global__ __launch_bounds__(256) void status_summary_kernel(const int* __restrict__ status,
                                                             const int* __restrict__ yielding, long long n,
                                                             long long* summary) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (status[i] != 0) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&summary[0]), 1ull);
    atomicMin(&summary[1], i);
  } else if (yielding[i] != 0) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&summary[2]), static_cast<unsigned long long>(yielding[i]));
  }
}

int launch_assemble_values(const AssemblyArgs& a, hipStream_t s) {
  (void)hipGetLastError();
  const long long threads = 3ll * a.num_blocks;
  if (threads == 0) return 0;
  const unsigned int grid = static_cast<unsigned int>((threads + 255) / 256);
  hipLaunchKernelGGL(assemble_values_kernel, dim3(grid), dim3(256), 0, s, a);
  return static_cast<int>(hipGetLastError());
}

int launch_gather_force(const ForceArgs& a, hipStream_t s) {
  (void)hipGetLastError();
  const long long threads = 3ll * a.num_nodes;
  if (threads == 0) return 0;
  const unsigned int grid = static_cast<unsigned int>((threads + 255) / 256);
  hipLaunchKernelGGL(gather_force_kernel, dim3(grid), dim3(256), 0, s, a);
  return static_cast<int>(hipGetLastError());
}

int launch_reduce(const double* x, long long n, int square, double* partial, int nparts, double* out,
                  hipStream_t s) {
  (void)hipGetLastError();
  if (nparts < 1) nparts = 1;
  if (nparts > 1024) nparts = 1024;
  hipLaunchKernelGGL(sum_partials_kernel, dim3(nparts), dim3(256), 0, s, x, n, partial, square);
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, partial, nparts, out, square);
  return static_cast<int>(hipGetLastError());
}

int launch_status_summary(const int* status, const int* yielding, long long n, long long* summary, hipStream_t s) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(status_init_kernel, dim3(1), dim3(64), 0, s, summary);
  if (n > 0) {
    const unsigned int grid = static_cast<unsigned int>((n + 255) / 256);
    hipLaunchKernelGGL(status_summary_kernel, dim3(grid), dim3(256), 0, s, status, yielding, n, summary);
  }
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
