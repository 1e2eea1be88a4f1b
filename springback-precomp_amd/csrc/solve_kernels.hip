// Jacobi-preconditioned conjugate gradients on the free-free block of a
// symmetric tangent (K_ff x = b), HBM-bound: each iteration reads the block
// once in its product q = K p and streams six vectors.
//
// Every sum has a fixed order (a wavefront's butterfly per row of the
// product, fixed partial ranges and one block's tree for the dot products),
// so a solve repeats bit for bit. The iteration's scalars stay on the device:
// the host launches iterations in runs and reads r . r between runs only.

#include "solve_kernels.hpp"

#include <cmath>

namespace sbk {

namespace {

constexpr int kRowsPerBlock = 4;  // one wavefront per row

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void spmv_kernel(const int* __restrict__ ptr, const int* __restrict__ index,
                                                   const double* __restrict__ values, long long n,
                                                   const double* __restrict__ x, double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const long long row = static_cast<long long>(blockIdx.x) * kRowsPerBlock + (threadIdx.x >> 6);
  if (row >= n) return;  // the whole wavefront
  const int k0 = ptr[row], k1 = ptr[row + 1];
  double acc = 0.0;
  for (int k = k0 + lane; k < k1; k += 64) acc = fma(values[k], x[index[k]], acc);
  acc = wave_sum(acc);
  if (lane == 0) y[row] = acc;
}

// dinv[i] = 1 / K_ii; a missing, non-positive or non-finite diagonal sets
// the break flag (the block is not positive definite).
__global__ __launch_bounds__(256) void diagonal_kernel(const int* __restrict__ ptr, const int* __restrict__ index,
                                                       const double* __restrict__ values, long long n,
                                                       double* __restrict__ dinv, double* __restrict__ scalars) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k)
    if (index[k] == i) d = values[k];
  if (!(d > 0.0) || !std::isfinite(d)) {
    scalars[kCgBreak] = 1.0;
    dinv[i] = 0.0;
  } else {
    dinv[i] = 1.0 / d;
  }
}

__global__ __launch_bounds__(256) void start_kernel(const double* __restrict__ b, const double* __restrict__ dinv,
                                                    long long n, double* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ z, double* __restrict__ p) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const double bi = b[i];
  x[i] = 0.0;
  r[i] = bi;
  z[i] = dinv[i] * bi;
  p[i] = z[i];
}

// partial[blk] = sum a b and, when c is given, partial[nparts + blk] =
// sum c d, over the block's fixed range.
__global__ __launch_bounds__(256) void dot_partials_kernel(const double* __restrict__ a,
                                                           const double* __restrict__ b,
                                                           const double* __restrict__ c,
                                                           const double* __restrict__ d, long long n,
                                                           double* __restrict__ partial) {
  __shared__ double s1[256], s2[256];
  const long long per = (n + gridDim.x - 1) / gridDim.x;
  const long long i0 = static_cast<long long>(blockIdx.x) * per;
  const long long i1 = i0 + per < n ? i0 + per : n;
  double acc1 = 0.0, acc2 = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    acc1 = fma(a[i], b[i], acc1);
    if (c) acc2 = fma(c[i], d[i], acc2);
  }
  s1[threadIdx.x] = acc1;
  s2[threadIdx.x] = acc2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s1[0];
    if (c) partial[gridDim.x + blockIdx.x] = s2[0];
  }
}

enum ScalarStep : int { kStepStart = 0, kStepAlpha = 1, kStepBeta = 2, kStepResidual = 3 };

// One block: the partial sums in a fixed tree, then the step's scalars.
__global__ __launch_bounds__(256) void scalars_kernel(const double* __restrict__ partial, int nparts, int step,
                                                      double* __restrict__ sc) {
  __shared__ double s1[256], s2[256];
  double acc1 = 0.0, acc2 = 0.0;
  const bool two = step == kStepStart || step == kStepBeta;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    acc1 += partial[i];
    if (two) acc2 += partial[nparts + i];
  }
  s1[threadIdx.x] = acc1;
  s2[threadIdx.x] = acc2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double v1 = s1[0], v2 = s2[0];
  switch (step) {
    case kStepStart:  // r.z, r.r (= b.b)
      sc[kCgRz] = v1;
      sc[kCgRr] = v2;
      sc[kCgBb] = v2;
      sc[kCgAlpha] = 0.0;
      sc[kCgBeta] = 0.0;
      break;
    case kStepAlpha: {  // p.q
      const double rz = sc[kCgRz];
      double alpha = 0.0;
      if (rz != 0.0) {  // r = 0 is converged: no step
        if (!(v1 > 0.0) || !std::isfinite(v1)) {
          sc[kCgBreak] = 1.0;
        } else {
          alpha = rz / v1;
        }
      }
      // after a break every further step is zero
      sc[kCgAlpha] = sc[kCgBreak] != 0.0 ? 0.0 : alpha;
      break;
    }
    case kStepBeta: {  // new r.z, r.r
      const double rz = sc[kCgRz];
      sc[kCgBeta] = (rz != 0.0 && sc[kCgBreak] == 0.0) ? v1 / rz : 0.0;
      sc[kCgRz] = v1;
      sc[kCgRr] = v2;
      break;
    }
    default:  // the true residual's r.r
      sc[kCgRr] = v1;
      break;
  }
}

// x += alpha p, r -= alpha q, z = D^-1 r.
__global__ __launch_bounds__(256) void update_kernel(const double* __restrict__ sc, const double* __restrict__ p,
                                                     const double* __restrict__ q, const double* __restrict__ dinv,
                                                     long long n, double* __restrict__ x, double* __restrict__ r,
                                                     double* __restrict__ z) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const double alpha = sc[kCgAlpha];
  x[i] = fma(alpha, p[i], x[i]);
  const double ri = fma(-alpha, q[i], r[i]);
  r[i] = ri;
  z[i] = dinv[i] * ri;
}

// p = z + beta p.
__global__ __launch_bounds__(256) void direction_kernel(const double* __restrict__ sc, const double* __restrict__ z,
                                                        long long n, double* __restrict__ p) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  p[i] = fma(sc[kCgBeta], p[i], z[i]);
}

// r = b - y.
__global__ __launch_bounds__(256) void difference_kernel(const double* __restrict__ b, const double* __restrict__ y,
                                                         long long n, double* __restrict__ r) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
}

__global__ void clear_break_kernel(double* sc) {
  if (threadIdx.x == 0) sc[kCgBreak] = 0.0;
}

unsigned int grid_of(long long n, long long per) { return static_cast<unsigned int>((n + per - 1) / per); }

}  // namespace

int launch_block_spmv(const int* ptr, const int* index, const double* values, long long n, const double* x,
                      double* y, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0)
    hipLaunchKernelGGL(spmv_kernel, dim3(grid_of(n, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, s, ptr, index,
                       values, n, x, y);
  return static_cast<int>(hipGetLastError());
}

int launch_cg_start(const CgArgs& a, const double* b, double* x, hipStream_t s) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(clear_break_kernel, dim3(1), dim3(64), 0, s, a.scalars);
  if (a.n > 0) {
    const unsigned int g = grid_of(a.n, 256);
    hipLaunchKernelGGL(diagonal_kernel, dim3(g), dim3(256), 0, s, a.ptr, a.index, a.values, a.n, a.dinv, a.scalars);
    hipLaunchKernelGGL(start_kernel, dim3(g), dim3(256), 0, s, b, a.dinv, a.n, x, a.r, a.z, a.p);
  }
  hipLaunchKernelGGL(dot_partials_kernel, dim3(a.nparts), dim3(256), 0, s, a.r, a.z, a.r, a.r, a.n, a.partial);
  hipLaunchKernelGGL(scalars_kernel, dim3(1), dim3(256), 0, s, a.partial, a.nparts, static_cast<int>(kStepStart),
                     a.scalars);
  return static_cast<int>(hipGetLastError());
}

int launch_cg_iterations(const CgArgs& a, double* x, int count, hipStream_t s) {
  (void)hipGetLastError();
  if (a.n == 0) return 0;
  const unsigned int g = grid_of(a.n, 256);
  for (int it = 0; it < count; ++it) {
    hipLaunchKernelGGL(spmv_kernel, dim3(grid_of(a.n, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, s, a.ptr,
                       a.index, a.values, a.n, a.p, a.q);
    hipLaunchKernelGGL(dot_partials_kernel, dim3(a.nparts), dim3(256), 0, s, a.p, a.q,
                       static_cast<const double*>(nullptr), static_cast<const double*>(nullptr), a.n, a.partial);
    hipLaunchKernelGGL(scalars_kernel, dim3(1), dim3(256), 0, s, a.partial, a.nparts, static_cast<int>(kStepAlpha),
                       a.scalars);
    hipLaunchKernelGGL(update_kernel, dim3(g), dim3(256), 0, s, a.scalars, a.p, a.q, a.dinv, a.n, x, a.r, a.z);
    hipLaunchKernelGGL(dot_partials_kernel, dim3(a.nparts), dim3(256), 0, s, a.r, a.z, a.r, a.r, a.n, a.partial);
    hipLaunchKernelGGL(scalars_kernel, dim3(1), dim3(256), 0, s, a.partial, a.nparts, static_cast<int>(kStepBeta),
                       a.scalars);
    hipLaunchKernelGGL(direction_kernel, dim3(g), dim3(256), 0, s, a.scalars, a.z, a.n, a.p);
  }
  return static_cast<int>(hipGetLastError());
}

int launch_cg_residual(const CgArgs& a, const double* b, const double* x, hipStream_t s) {
  (void)hipGetLastError();
  if (a.n > 0) {
    hipLaunchKernelGGL(spmv_kernel, dim3(grid_of(a.n, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, s, a.ptr,
                       a.index, a.values, a.n, x, a.q);
    hipLaunchKernelGGL(difference_kernel, dim3(grid_of(a.n, 256)), dim3(256), 0, s, b, a.q, a.n, a.r);
  }
  hipLaunchKernelGGL(dot_partials_kernel, dim3(a.nparts), dim3(256), 0, s, a.r, a.r,
                     static_cast<const double*>(nullptr), static_cast<const double*>(nullptr), a.n, a.partial);
  hipLaunchKernelGGL(scalars_kernel, dim3(1), dim3(256), 0, s, a.partial, a.nparts, static_cast<int>(kStepResidual),
                     a.scalars);
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
