// Jacobi-preconditioned conjugate gradients on the free-free block of a
// symmetric tangent (K_ff x = b), HBM-bound: each iteration reads the block
// once in its product q = K p and streams the vectors twice.
//
// An iteration is three launches: the product (which also sums p . q per
// block), the update of x, r and z (which sums r . z and r . r per block) and
// the new direction. Each block of the update and of the direction derives
// alpha and beta itself from the partial sums, in the same fixed order, so no
// launch is spent on scalars and every sum has one order: a solve repeats bit
// for bit. The partial sums of r . z alternate between two buffers by
// iteration parity, so the direction reads the old and the new sum while the
// update of the next iteration has not yet written. The host launches
// iterations in runs and reads r . r between runs only.

#include "solve_kernels.hpp"

#include <cmath>
#include <vector>

namespace sbk {

namespace {

constexpr int kWaves = 4;          // wavefronts per product block
constexpr int kLanesPerRow = 16;   // a row's lanes in the product (rows of the block hold up to 81 entries)
constexpr int kGroups = 64 / kLanesPerRow;  // rows per wavefront at a time
constexpr int kSlots = (81 + kLanesPerRow - 1) / kLanesPerRow;  // a lane's entries of a full row

// The sum of v[0, m) in a fixed order, returned to every thread of the
// (256-thread) block; s is the block's scratch.
__device__ double block_sum(const double* __restrict__ v, int m, double* s) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) acc += v[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  const double out = s[0];
  __syncthreads();
  return out;
}

// Two per-thread sums reduced over the block in a fixed tree; thread 0 gets
// the totals.
__device__ void block_reduce2(double& a, double& b, double* s1, double* s2) {
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  a = s1[0];
  b = s2[0];
}

// A block's place in a solve over segments: its problem p, its index j among
// that problem's `count` blocks (the problem's first at `first`) and the
// problem's rows [r0, r0 + n). A plain solve is the one problem of all rows.
struct Part {
  int p, j, count, first;
  long long r0, n;
};

__device__ __forceinline__ long long seg_r0(const CgArgs& A, int p) { return A.seg_row ? A.seg_row[p] : 0; }
__device__ __forceinline__ long long seg_n(const CgArgs& A, int p) {
  return A.seg_row ? A.seg_row[p + 1] - A.seg_row[p] : A.n;
}
__device__ __forceinline__ int vector_first(const CgArgs& A, int p) { return A.vfirst ? A.vfirst[p] : 0; }
__device__ __forceinline__ int vector_count(const CgArgs& A, int p) {
  return A.vfirst ? A.vfirst[p + 1] - A.vfirst[p] : A.nparts;
}
__device__ __forceinline__ int product_first(const CgArgs& A, int p) { return A.pfirst ? A.pfirst[p] : 0; }
__device__ __forceinline__ int product_count(const CgArgs& A, int p) {
  return A.pfirst ? A.pfirst[p + 1] - A.pfirst[p] : A.nrows;
}

// The vector block's place (grid: nparts blocks).
__device__ __forceinline__ Part vector_part(const CgArgs& A) {
  Part q;
  q.p = A.vseg ? A.vseg[blockIdx.x] : 0;
  q.first = vector_first(A, q.p);
  q.count = vector_count(A, q.p);
  q.j = static_cast<int>(blockIdx.x) - q.first;
  q.r0 = seg_r0(A, q.p);
  q.n = seg_n(A, q.p);
  return q;
}

// The product block's place (grid: nrows blocks).
__device__ __forceinline__ Part product_part(const CgArgs& A) {
  Part q;
  q.p = A.pseg ? A.pseg[blockIdx.x] : 0;
  q.first = product_first(A, q.p);
  q.count = product_count(A, q.p);
  q.j = static_cast<int>(blockIdx.x) - q.first;
  q.r0 = seg_r0(A, q.p);
  q.n = seg_n(A, q.p);
  return q;
}

// A problem the launch leaves as it is (converged, failed, or not among
// those a residual or restart is for).
__device__ __forceinline__ bool resting(const CgArgs& A, int p) { return A.active && A.active[p] == 0; }

__device__ __forceinline__ double* scalars_of(const CgArgs& A, int p) { return A.scalars + p * kCgScalars; }

// The block's contiguous range of its problem's rows when they are cut into
// the problem's `count` parts (a plain solve: of [0, n) into gridDim.x parts).
__device__ __forceinline__ void part_range(const Part& q, long long& i0, long long& i1) {
  const long long per = (q.n + q.count - 1) / q.count;
  const long long end = q.r0 + q.n;
  i0 = q.r0 + static_cast<long long>(q.j) * per;
  i1 = i0 + per < end ? i0 + per : end;
}

// The row's sum of K(row, :) x in a fixed order, in every lane of the row's
// group: lane l of the group takes entries l, l + kLanesPerRow, ..., then a
// butterfly within the group. A row beyond n sums to 0 (every lane of the
// wavefront takes part in the butterfly).
__device__ __forceinline__ double row_product(const int* __restrict__ ptr, const int* __restrict__ index,
                                              const double* __restrict__ values, const double* __restrict__ x,
                                              long long row, long long n, int l) {
  double acc = 0.0;
  if (row < n) {
    // the lane's first kSlots entries: every index and value load issued
    // before the gathers of x, and those before the sums (a row of the block
    // holds at most 81 entries; longer rows finish in the loop after)
    const int k0 = ptr[row] + l, k1 = ptr[row + 1];
    int col[kSlots];
    double v[kSlots], xv[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int k = k0 + s * kLanesPerRow;
      col[s] = k < k1 ? index[k] : -1;
      v[s] = k < k1 ? values[k] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < kSlots; ++s) xv[s] = col[s] >= 0 ? x[col[s]] : 0.0;
#pragma unroll
    for (int s = 0; s < kSlots; ++s)
      if (col[s] >= 0) acc = fma(v[s], xv[s], acc);
    for (int k = k0 + kSlots * kLanesPerRow; k < k1; k += kLanesPerRow) acc = fma(values[k], x[index[k]], acc);
  }
#pragma unroll
  for (int off = kLanesPerRow / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  return acc;
}

// out = K in over the block's problem's rows, kGroups rows per wavefront at a
// time, dealt round-robin over the problem's product blocks; partial[block] =
// sum over the block's rows of w out and, when square is given,
// square[block] = sum of out out (each group's rows in order, then the groups
// and wavefronts in order).
__global__ __launch_bounds__(256) void product_kernel(CgArgs A, const double* __restrict__ in,
                                                      double* __restrict__ out, const double* __restrict__ w,
                                                      double* __restrict__ partial, double* __restrict__ square) {
  const Part q = product_part(A);
  if (resting(A, q.p)) return;
  __shared__ double ws[kWaves * kGroups], ws2[kWaves * kGroups];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane / kLanesPerRow, l = lane % kLanesPerRow;
  const long long end = q.r0 + q.n;
  const long long stride = static_cast<long long>(q.count) * kWaves * kGroups;
  double dot = 0.0, sq = 0.0;
  for (long long base = q.r0 + (static_cast<long long>(q.j) * kWaves + wave) * kGroups; base < end;
       base += stride) {
    const long long row = base + g;
    const double o = row_product(A.ptr, A.index, A.values, in, row, end, l);
    if (l == 0 && row < end) {
      out[row] = o;
      dot = fma(w[row], o, dot);
      sq = fma(o, o, sq);
    }
  }
  if (l == 0) {
    ws[wave * kGroups + g] = dot;
    ws2[wave * kGroups + g] = sq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0, sum2 = 0.0;
    for (int i = 0; i < kWaves * kGroups; ++i) {
      sum += ws[i];
      sum2 += ws2[i];
    }
    partial[blockIdx.x] = sum;
    if (square) square[blockIdx.x] = sum2;
  }
}

// alpha = rz / p.q from the problem's partial sums (0 once the iteration
// broke down or r = 0), then x += alpha p, r -= alpha q, z = D^-1 r over the
// block's range, with the block's r . z (into the buffer of iteration k's
// parity) and r . r.
__global__ __launch_bounds__(256) void update_kernel(CgArgs A, double* __restrict__ x, long long k) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double s1[256], s2[256];
  double* sc = scalars_of(A, q.p);
  const double pq = block_sum(A.pq + product_first(A, q.p), product_count(A, q.p), s1);
  const double rz = block_sum(A.rz + ((k + 1) & 1) * A.nparts + q.first, q.count, s1);
  bool broke = sc[kCgBreak] != 0.0;
  double alpha = 0.0;
  if (!broke && rz != 0.0) {
    if (pq > 0.0 && std::isfinite(pq))
      alpha = rz / pq;
    else
      broke = true;  // p . K p <= 0: not positive definite (every block sees the same sums)
  }
  if (q.j == 0 && threadIdx.x == 0) {
    sc[kCgAlpha] = alpha;
    if (broke) sc[kCgBreak] = 1.0;
  }
  long long i0, i1;
  part_range(q, i0, i1);
  double acc_rz = 0.0, acc_rr = 0.0;
  if (A.external) {  // z and r . z come from the preconditioner
    for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
      x[i] = fma(alpha, A.p[i], x[i]);
      A.r[i] = fma(-alpha, A.q[i], A.r[i]);
    }
    return;
  }
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    x[i] = fma(alpha, A.p[i], x[i]);
    const double ri = fma(-alpha, A.q[i], A.r[i]);
    const double zi = A.dinv[i] * ri;
    A.r[i] = ri;
    A.z[i] = zi;
    acc_rz = fma(ri, zi, acc_rz);
    acc_rr = fma(ri, ri, acc_rr);
  }
  block_reduce2(acc_rz, acc_rr, s1, s2);
  if (threadIdx.x == 0) {
    A.rz[(k & 1) * A.nparts + blockIdx.x] = acc_rz;
    A.rr[blockIdx.x] = acc_rr;
  }
}

// After an external preconditioner: the block's r . z into the buffer
// `slot`, its r . r, and p = z when set_p (a start or restart).
__global__ __launch_bounds__(256) void precond_dot_kernel(CgArgs A, int slot, int set_p) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double s1[256], s2[256];
  long long i0, i1;
  part_range(q, i0, i1);
  double acc_rz = 0.0, acc_rr = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double ri = A.r[i], zi = A.z[i];
    if (set_p) A.p[i] = zi;
    acc_rz = fma(ri, zi, acc_rz);
    acc_rr = fma(ri, ri, acc_rr);
  }
  block_reduce2(acc_rz, acc_rr, s1, s2);
  if (threadIdx.x == 0) {
    A.rz[slot * A.nparts + blockIdx.x] = acc_rz;
    A.rr[blockIdx.x] = acc_rr;
  }
}

// beta = r.z (new) / r.z (old) from the problem's partial sums, p = z + beta p
// over the block's range; the problem's first block leaves r . z, r . r and
// beta in its scalars.
__global__ __launch_bounds__(256) void direction_kernel(CgArgs A, long long k) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double s[256];
  double* sc = scalars_of(A, q.p);
  const double rz_new = block_sum(A.rz + (k & 1) * A.nparts + q.first, q.count, s);
  const double rz_old = block_sum(A.rz + ((k + 1) & 1) * A.nparts + q.first, q.count, s);
  const bool broke = sc[kCgBreak] != 0.0;
  const double beta = (!broke && rz_old != 0.0) ? rz_new / rz_old : 0.0;
  if (q.j == 0) {
    const double rr = block_sum(A.rr + q.first, q.count, s);
    if (threadIdx.x == 0) {
      sc[kCgRz] = rz_new;
      sc[kCgRr] = rr;
      sc[kCgBeta] = beta;
    }
  }
  long long i0, i1;
  part_range(q, i0, i1);
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) A.p[i] = fma(beta, A.p[i], A.z[i]);
}

// Conjugate gradients (positive = 1): dinv[i] = 1 / K_ii, and a missing,
// non-positive or non-finite diagonal sets the problem's break flag (its
// block is not positive definite). BiCGStab (positive = 0), whose block need
// not be definite: dinv[i] = 1 / |K_ii|, or 1 where K_ii is zero or not
// finite. Over the vector blocks' ranges.
__global__ __launch_bounds__(256) void diagonal_kernel(CgArgs A, int positive) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  long long i0, i1;
  part_range(q, i0, i1);
  bool bad = false;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    double d = 0.0;
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
      if (A.index[k] == i) d = A.values[k];
    if (!positive) {
      A.dinv[i] = (d != 0.0 && std::isfinite(d)) ? 1.0 / fabs(d) : 1.0;
    } else if (!(d > 0.0) || !std::isfinite(d)) {
      bad = true;
      A.dinv[i] = 0.0;
    } else {
      A.dinv[i] = 1.0 / d;
    }
  }
  if (bad) scalars_of(A, q.p)[kCgBreak] = 1.0;
}

// x = 0, r = b, z = D^-1 b, p = z over the block's range; its r . z goes to
// the buffer iteration 0 reads as the old sum, its r . r to rr.
__global__ __launch_bounds__(256) void start_kernel(CgArgs A, const double* __restrict__ b, double* __restrict__ x) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double s1[256], s2[256];
  long long i0, i1;
  part_range(q, i0, i1);
  double acc_rz = 0.0, acc_rr = 0.0;
  if (A.external) {  // z, p and r . z come from the preconditioner
    for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
      x[i] = 0.0;
      A.r[i] = b[i];
    }
    return;
  }
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double bi = b[i];
    const double zi = A.dinv[i] * bi;
    x[i] = 0.0;
    A.r[i] = bi;
    A.z[i] = zi;
    A.p[i] = zi;
    acc_rz = fma(bi, zi, acc_rz);
    acc_rr = fma(bi, bi, acc_rr);
  }
  block_reduce2(acc_rz, acc_rr, s1, s2);
  if (threadIdx.x == 0) {
    A.rz[A.nparts + blockIdx.x] = acc_rz;
    A.rr[blockIdx.x] = acc_rr;
  }
}

// One block per problem (grid: nseg).
__global__ __launch_bounds__(256) void start_scalars_kernel(CgArgs A) {
  const int p = blockIdx.x;
  if (resting(A, p)) return;
  __shared__ double s[256];
  const int first = vector_first(A, p), count = vector_count(A, p);
  const double rz = block_sum(A.rz + A.nparts + first, count, s);
  const double rr = block_sum(A.rr + first, count, s);
  if (threadIdx.x == 0) {
    double* sc = scalars_of(A, p);
    sc[kCgRz] = rz;
    sc[kCgRr] = rr;
    sc[kCgBb] = rr;
    sc[kCgAlpha] = 0.0;
    sc[kCgBeta] = 0.0;
  }
}

// A restart from the residual in r: z = D^-1 r, p = z over the block's
// range, its r . z into the buffer `slot` (the one the next iteration reads
// as the old sum) and its r . r.
__global__ __launch_bounds__(256) void restart_kernel(CgArgs A, int slot) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double s1[256], s2[256];
  long long i0, i1;
  part_range(q, i0, i1);
  double acc_rz = 0.0, acc_rr = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double ri = A.r[i];
    const double zi = A.dinv[i] * ri;
    A.z[i] = zi;
    A.p[i] = zi;
    acc_rz = fma(ri, zi, acc_rz);
    acc_rr = fma(ri, ri, acc_rr);
  }
  block_reduce2(acc_rz, acc_rr, s1, s2);
  if (threadIdx.x == 0) {
    A.rz[slot * A.nparts + blockIdx.x] = acc_rz;
    A.rr[blockIdx.x] = acc_rr;
  }
}

__global__ __launch_bounds__(256) void restart_scalars_kernel(CgArgs A, int slot) {
  const int p = blockIdx.x;
  if (resting(A, p)) return;
  __shared__ double s[256];
  const int first = vector_first(A, p), count = vector_count(A, p);
  const double rz = block_sum(A.rz + slot * A.nparts + first, count, s);
  const double rr = block_sum(A.rr + first, count, s);
  if (threadIdx.x == 0) {
    double* sc = scalars_of(A, p);
    sc[kCgRz] = rz;
    sc[kCgRr] = rr;
  }
}

// y = K x, kGroups rows per wavefront.
__global__ __launch_bounds__(256) void spmv_kernel(const int* __restrict__ ptr, const int* __restrict__ index,
                                                   const double* __restrict__ values, long long n,
                                                   const double* __restrict__ x, double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const long long base = (static_cast<long long>(blockIdx.x) * kWaves + (threadIdx.x >> 6)) * kGroups;
  if (base >= n) return;  // the whole wavefront
  const long long row = base + lane / kLanesPerRow;
  const double v = row_product(ptr, index, values, x, row, n, lane % kLanesPerRow);
  if (lane % kLanesPerRow == 0 && row < n) y[row] = v;
}

// r = b - q over the block's range and its r . r into partial.
__global__ __launch_bounds__(256) void residual_kernel(CgArgs A, const double* __restrict__ b) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double s1[256], s2[256];
  long long i0, i1;
  part_range(q, i0, i1);
  double acc = 0.0, unused = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double ri = b[i] - A.q[i];
    A.r[i] = ri;
    acc = fma(ri, ri, acc);
  }
  block_reduce2(acc, unused, s1, s2);
  if (threadIdx.x == 0) A.partial[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void residual_scalars_kernel(CgArgs A) {
  const int p = blockIdx.x;
  if (resting(A, p)) return;
  __shared__ double s[256];
  const double rr = block_sum(A.partial + vector_first(A, p), vector_count(A, p), s);
  if (threadIdx.x == 0) scalars_of(A, p)[kCgRr] = rr;
}

// ---------------------------------------------------------------------------
// BiCGStab (van der Vorst), preconditioned with the diagonal, for a block
// that is not symmetric. Iteration k: p = r + beta (p - omega v) (p = r at a
// cycle's start), y = D^-1 p, v = K y, alpha = rho / r^.v, s = r - alpha v,
// x += alpha y, z = D^-1 s, t = K z, omega = t.s / t.t, x += omega z,
// r = s - omega t, rho' = r^ . r. The scalars come from the problem's partial
// sums in every block that needs them, in one order; a zero rho, r^ . v or
// omega with r != 0 is a breakdown (the flag; every later step is zero).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void bicg_start_kernel(CgArgs A, const double* __restrict__ b,
                                                         double* __restrict__ x) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double s1[256], s2[256];
  long long i0, i1;
  part_range(q, i0, i1);
  double acc = 0.0, unused = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double bi = b[i];
    x[i] = 0.0;
    A.r[i] = bi;
    A.rhat[i] = bi;
    A.p[i] = 0.0;
    A.v[i] = 0.0;
    acc = fma(bi, bi, acc);
  }
  block_reduce2(acc, unused, s1, s2);
  if (threadIdx.x == 0) {
    A.rho[blockIdx.x] = acc;  // iteration 0's r^ . r
    A.rr[blockIdx.x] = acc;
  }
}

// rho (the slot's r^ . r) and r . r into the problem's scalars; b . b too at
// the start. One block per problem.
__global__ __launch_bounds__(256) void bicg_scalars_kernel(CgArgs A, int slot, int start) {
  const int p = blockIdx.x;
  if (resting(A, p)) return;
  __shared__ double s[256];
  const int first = vector_first(A, p), count = vector_count(A, p);
  const double rho = block_sum(A.rho + slot * A.nparts + first, count, s);
  const double rr = block_sum(A.rr + first, count, s);
  if (threadIdx.x == 0) {
    double* sc = scalars_of(A, p);
    sc[kCgRz] = rho;
    sc[kCgRr] = rr;
    if (start) {
      sc[kCgBb] = rr;
      sc[kCgAlpha] = 0.0;
      sc[kCgBeta] = 0.0;
    }
  }
}

__global__ __launch_bounds__(256) void bicg_direction_kernel(CgArgs A, long long cycle, long long k) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double sh[256];
  double* sc = scalars_of(A, q.p);
  const int pf = product_first(A, q.p), pc = product_count(A, q.p);
  const double rr = block_sum(A.rr + q.first, q.count, sh);
  const double rho = block_sum(A.rho + (k & 1) * A.nparts + q.first, q.count, sh);
  const double rho_prev = block_sum(A.rho + ((k + 1) & 1) * A.nparts + q.first, q.count, sh);
  const double rv = block_sum(A.pv + pf, pc, sh);
  const double ts = block_sum(A.pts + pf, pc, sh);
  const double tt = block_sum(A.pts + A.nrows + pf, pc, sh);
  const bool first = k == (A.cycles ? A.cycles[q.p] : cycle);
  bool broke = sc[kCgBreak] != 0.0;
  double beta = 0.0, omega = 0.0;
  if (!first) {
    const double alpha = rho_prev / rv;
    omega = ts / tt;
    if (!broke && rr != 0.0) {
      if (rho == 0.0 || omega == 0.0 || !std::isfinite(alpha) || !std::isfinite(omega))
        broke = true;
      else
        beta = (rho / rho_prev) * (alpha / omega);
    }
    if (broke || !std::isfinite(omega)) omega = 0.0;
  }
  if (q.j == 0 && threadIdx.x == 0 && broke) sc[kCgBreak] = 1.0;
  long long i0, i1;
  part_range(q, i0, i1);
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double pi = first ? A.r[i] : fma(beta, fma(-omega, A.v[i], A.p[i]), A.r[i]);
    A.p[i] = pi;
    if (!A.external) A.y[i] = A.dinv[i] * pi;
  }
}

__global__ __launch_bounds__(256) void bicg_half_kernel(CgArgs A, double* __restrict__ x, long long k) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double sh[256];
  double* sc = scalars_of(A, q.p);
  const double rr = block_sum(A.rr + q.first, q.count, sh);
  const double rho = block_sum(A.rho + (k & 1) * A.nparts + q.first, q.count, sh);
  const double rv = block_sum(A.pv + product_first(A, q.p), product_count(A, q.p), sh);
  bool broke = sc[kCgBreak] != 0.0;
  double alpha = 0.0;
  if (!broke && rr != 0.0) {
    if (rv != 0.0 && std::isfinite(rho / rv))
      alpha = rho / rv;
    else
      broke = true;  // every block sees the same sums
  }
  if (q.j == 0 && threadIdx.x == 0) {
    sc[kCgAlpha] = alpha;
    if (broke) sc[kCgBreak] = 1.0;
  }
  long long i0, i1;
  part_range(q, i0, i1);
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double si = fma(-alpha, A.v[i], A.r[i]);
    A.s[i] = si;
    x[i] = fma(alpha, A.y[i], x[i]);
    if (!A.external) A.z[i] = A.dinv[i] * si;
  }
}

__global__ __launch_bounds__(256) void bicg_full_kernel(CgArgs A, double* __restrict__ x, long long k) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double s1[256], s2[256];
  const int pf = product_first(A, q.p), pc = product_count(A, q.p);
  const double ts = block_sum(A.pts + pf, pc, s1);
  const double tt = block_sum(A.pts + A.nrows + pf, pc, s1);
  const bool broke = scalars_of(A, q.p)[kCgBreak] != 0.0;
  const double omega = (!broke && tt != 0.0 && std::isfinite(ts / tt)) ? ts / tt : 0.0;
  long long i0, i1;
  part_range(q, i0, i1);
  double acc_rho = 0.0, acc_rr = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    x[i] = fma(omega, A.z[i], x[i]);
    const double ri = fma(-omega, A.t[i], A.s[i]);
    A.r[i] = ri;
    acc_rho = fma(A.rhat[i], ri, acc_rho);
    acc_rr = fma(ri, ri, acc_rr);
  }
  block_reduce2(acc_rho, acc_rr, s1, s2);
  if (threadIdx.x == 0) {
    A.rho[((k + 1) & 1) * A.nparts + blockIdx.x] = acc_rho;
    A.rr[blockIdx.x] = acc_rr;
  }
}

// A new cycle from the residual in r: r^ = r, its r . r as rho (slot) and rr.
__global__ __launch_bounds__(256) void bicg_restart_kernel(CgArgs A, int slot) {
  const Part q = vector_part(A);
  if (resting(A, q.p)) return;
  __shared__ double s1[256], s2[256];
  long long i0, i1;
  part_range(q, i0, i1);
  double acc = 0.0, unused = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double ri = A.r[i];
    A.rhat[i] = ri;
    acc = fma(ri, ri, acc);
  }
  block_reduce2(acc, unused, s1, s2);
  if (threadIdx.x == 0) {
    A.rho[slot * A.nparts + blockIdx.x] = acc;
    A.rr[blockIdx.x] = acc;
  }
}

// Clears the break flag of every problem the launch is for.
__global__ __launch_bounds__(256) void clear_break_kernel(CgArgs A) {
  for (int p = threadIdx.x; p < A.nseg; p += 256)
    if (!resting(A, p)) scalars_of(A, p)[kCgBreak] = 0.0;
}

// flag[0] = 1 when a row of a problem has a column outside the problem's rows
// (the row's columns ascend: its first and last decide). Over the vector
// blocks' ranges.
__global__ __launch_bounds__(256) void segment_check_kernel(CgArgs A, int* __restrict__ flag) {
  const Part q = vector_part(A);
  long long i0, i1;
  part_range(q, i0, i1);
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    const int k0 = A.ptr[i], k1 = A.ptr[i + 1];
    if (k1 > k0 && (A.index[k0] < q.r0 || A.index[k1 - 1] >= q.r0 + q.n)) flag[0] = 1;
  }
}

unsigned int grid_of(long long n, long long per) { return static_cast<unsigned int>((n + per - 1) / per); }

}  // namespace

int cg_vector_blocks(long long n) {
  // two entries per thread, at most 1024 blocks (each block of the update and
  // the direction sums every partial)
  const long long b = (n + 511) / 512;
  return static_cast<int>(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

int cg_product_blocks(long long n) {
  // about 4 rows per group, at most 2048 blocks (8 wavefronts per SIMD)
  const long long b = (n + 4 * kWaves * kGroups - 1) / (4 * kWaves * kGroups);
  return static_cast<int>(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

long long cg_work_doubles(long long n) {
  return 10 * n + 4ll * cg_product_blocks(n) + 7ll * cg_vector_blocks(n) + kCgScalars;
}

namespace {

// The work's pieces from `work` for a.n rows, a.nparts vector blocks, a.nrows
// product blocks and a.nseg problems.
void lay_out(CgArgs& a, double* work) {
  a.dinv = work;
  a.r = a.dinv + a.n;
  a.z = a.r + a.n;
  a.p = a.z + a.n;
  a.q = a.p + a.n;
  a.pq = a.q + a.n;
  a.rz = a.pq + a.nrows;
  a.rr = a.rz + 2 * a.nparts;
  a.partial = a.rr + a.nparts;
  a.scalars = a.partial + 2 * a.nparts;
  a.rhat = a.scalars + static_cast<long long>(a.nseg) * kCgScalars;
  a.v = a.rhat + a.n;
  a.y = a.v + a.n;
  a.s = a.y + a.n;
  a.t = a.s + a.n;
  a.pv = a.t + a.n;
  a.pts = a.pv + a.nrows;
  a.rho = a.pts + 2 * a.nrows;
}

}  // namespace

void cg_layout(CgArgs& a, double* work) {
  a.nparts = cg_vector_blocks(a.n);
  a.nrows = cg_product_blocks(a.n);
  a.nseg = 1;
  a.seg_row = nullptr;
  a.vseg = a.vfirst = a.pseg = a.pfirst = a.active = nullptr;
  a.cycles = nullptr;
  a.external = 0;
  a.precond = nullptr;
  a.precond_ctx = nullptr;
  lay_out(a, work);
}

CgSegments cg_segments(const std::vector<long long>& rows) {
  CgSegments g;
  g.rows = rows;
  const int nseg = static_cast<int>(rows.size()) - 1;
  g.vfirst.assign(1, 0);
  g.pfirst.assign(1, 0);
  for (int p = 0; p < nseg; ++p) {
    const long long n = rows[p + 1] - rows[p];
    const int v = cg_vector_blocks(n), q = cg_product_blocks(n);
    g.vseg.insert(g.vseg.end(), v, p);
    g.pseg.insert(g.pseg.end(), q, p);
    g.vfirst.push_back(g.vfirst.back() + v);
    g.pfirst.push_back(g.pfirst.back() + q);
  }
  return g;
}

long long cg_segment_work_doubles(const CgSegments& g) {
  const long long n = g.rows.back(), nseg = static_cast<long long>(g.rows.size()) - 1;
  return 10 * n + 4ll * g.pfirst.back() + 7ll * g.vfirst.back() + nseg * kCgScalars;
}

long long cg_segment_meta_bytes(const CgSegments& g) {
  const long long nseg = static_cast<long long>(g.rows.size()) - 1;
  return static_cast<long long>(sizeof(long long)) * (2 * nseg + 1) +
         static_cast<long long>(sizeof(int)) * (g.vseg.size() + g.pseg.size() + 2 * (nseg + 1) + nseg + 1);
}

int cg_segment_layout(CgArgs& a, const CgSegments& g, double* work, void* meta, hipStream_t s) {
  const int nseg = static_cast<int>(g.rows.size()) - 1;
  a.n = g.rows.back();
  a.nseg = nseg;
  a.external = 0;
  a.precond = nullptr;
  a.precond_ctx = nullptr;
  a.nparts = g.vfirst.back();
  a.nrows = g.pfirst.back();
  lay_out(a, work);
  // meta: seg_row (nseg + 1), cycles (nseg), then the int arrays
  long long* ll = static_cast<long long*>(meta);
  a.seg_row = ll;
  a.cycles = ll + nseg + 1;
  int* in = reinterpret_cast<int*>(ll + 2 * nseg + 1);
  int* vseg = in;
  int* vfirst = vseg + g.vseg.size();
  int* pseg = vfirst + nseg + 1;
  int* pfirst = pseg + g.pseg.size();
  int* active = pfirst + nseg + 1;
  a.vseg = vseg;
  a.vfirst = vfirst;
  a.pseg = pseg;
  a.pfirst = pfirst;
  a.active = active;
  std::vector<long long> zeros(nseg, 0);
  std::vector<int> ones(nseg, 1);
  (void)hipGetLastError();
  if (hipMemcpyAsync(ll, g.rows.data(), sizeof(long long) * (nseg + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(ll + nseg + 1, zeros.data(), sizeof(long long) * nseg, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(vseg, g.vseg.data(), sizeof(int) * g.vseg.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(vfirst, g.vfirst.data(), sizeof(int) * (nseg + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(pseg, g.pseg.data(), sizeof(int) * g.pseg.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(pfirst, g.pfirst.data(), sizeof(int) * (nseg + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(active, ones.data(), sizeof(int) * nseg, hipMemcpyHostToDevice, s) != hipSuccess)
    return 1;
  // the host vectors above are released on return: wait for the copies
  return hipStreamSynchronize(s) == hipSuccess ? 0 : 1;
}

int launch_segment_check(const CgArgs& a, int* flag, hipStream_t s) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(segment_check_kernel, dim3(a.nparts), dim3(256), 0, s, a, flag);
  return static_cast<int>(hipGetLastError());
}

int launch_block_spmv(const int* ptr, const int* index, const double* values, long long n, const double* x,
                      double* y, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0)
    hipLaunchKernelGGL(spmv_kernel, dim3(grid_of(n, kWaves * kGroups)), dim3(64 * kWaves), 0, s, ptr, index, values,
                       n, x, y);
  return static_cast<int>(hipGetLastError());
}

int launch_cg_start(const CgArgs& a, const double* b, double* x, hipStream_t s) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(clear_break_kernel, dim3(1), dim3(256), 0, s, a);
  hipLaunchKernelGGL(diagonal_kernel, dim3(a.nparts), dim3(256), 0, s, a, 1);
  hipLaunchKernelGGL(start_kernel, dim3(a.nparts), dim3(256), 0, s, a, b, x);
  if (a.external) {
    if (a.precond(a.precond_ctx, a.r, a.z, s) != 0) return 1;
    hipLaunchKernelGGL(precond_dot_kernel, dim3(a.nparts), dim3(256), 0, s, a, 1, 1);
  }
  hipLaunchKernelGGL(start_scalars_kernel, dim3(a.nseg), dim3(256), 0, s, a);
  return static_cast<int>(hipGetLastError());
}

int launch_cg_iterations(const CgArgs& a, double* x, long long first, int count, hipStream_t s) {
  (void)hipGetLastError();
  if (a.n == 0) return 0;
  for (long long k = first; k < first + count; ++k) {
    hipLaunchKernelGGL(product_kernel, dim3(a.nrows), dim3(64 * kWaves), 0, s, a, a.p, a.q, a.p, a.pq,
                       static_cast<double*>(nullptr));
    hipLaunchKernelGGL(update_kernel, dim3(a.nparts), dim3(256), 0, s, a, x, k);
    if (a.external) {
      if (a.precond(a.precond_ctx, a.r, a.z, s) != 0) return 1;
      hipLaunchKernelGGL(precond_dot_kernel, dim3(a.nparts), dim3(256), 0, s, a, static_cast<int>(k & 1), 0);
    }
    hipLaunchKernelGGL(direction_kernel, dim3(a.nparts), dim3(256), 0, s, a, k);
  }
  return static_cast<int>(hipGetLastError());
}

int launch_bicg_start(const CgArgs& a, const double* b, double* x, hipStream_t s) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(clear_break_kernel, dim3(1), dim3(256), 0, s, a);
  hipLaunchKernelGGL(diagonal_kernel, dim3(a.nparts), dim3(256), 0, s, a, 0);
  hipLaunchKernelGGL(bicg_start_kernel, dim3(a.nparts), dim3(256), 0, s, a, b, x);
  hipLaunchKernelGGL(bicg_scalars_kernel, dim3(a.nseg), dim3(256), 0, s, a, 0, 1);
  return static_cast<int>(hipGetLastError());
}

int launch_bicg_iterations(const CgArgs& a, double* x, long long cycle, long long first, int count,
                           hipStream_t s) {
  (void)hipGetLastError();
  if (a.n == 0) return 0;
  for (long long k = first; k < first + count; ++k) {
    hipLaunchKernelGGL(bicg_direction_kernel, dim3(a.nparts), dim3(256), 0, s, a, cycle, k);
    if (a.external && a.precond(a.precond_ctx, a.p, a.y, s) != 0) return 1;
    hipLaunchKernelGGL(product_kernel, dim3(a.nrows), dim3(64 * kWaves), 0, s, a, a.y, a.v, a.rhat, a.pv,
                       static_cast<double*>(nullptr));
    hipLaunchKernelGGL(bicg_half_kernel, dim3(a.nparts), dim3(256), 0, s, a, x, k);
    if (a.external && a.precond(a.precond_ctx, a.s, a.z, s) != 0) return 1;
    hipLaunchKernelGGL(product_kernel, dim3(a.nrows), dim3(64 * kWaves), 0, s, a, a.z, a.t, a.s, a.pts,
                       a.pts + a.nrows);
    hipLaunchKernelGGL(bicg_full_kernel, dim3(a.nparts), dim3(256), 0, s, a, x, k);
  }
  hipLaunchKernelGGL(bicg_scalars_kernel, dim3(a.nseg), dim3(256), 0, s, a, static_cast<int>((first + count) & 1),
                     0);
  return static_cast<int>(hipGetLastError());
}

int launch_bicg_restart(const CgArgs& a, long long next, hipStream_t s) {
  (void)hipGetLastError();
  const int slot = static_cast<int>(next & 1);
  hipLaunchKernelGGL(clear_break_kernel, dim3(1), dim3(256), 0, s, a);  // a restart after a breakdown
  hipLaunchKernelGGL(bicg_restart_kernel, dim3(a.nparts), dim3(256), 0, s, a, slot);
  hipLaunchKernelGGL(bicg_scalars_kernel, dim3(a.nseg), dim3(256), 0, s, a, slot, 0);
  return static_cast<int>(hipGetLastError());
}

int launch_cg_restart(const CgArgs& a, long long next, hipStream_t s) {
  (void)hipGetLastError();
  const int slot = static_cast<int>((next + 1) & 1);
  if (a.external) {
    if (a.precond(a.precond_ctx, a.r, a.z, s) != 0) return 1;
    hipLaunchKernelGGL(precond_dot_kernel, dim3(a.nparts), dim3(256), 0, s, a, slot, 1);
  } else {
    hipLaunchKernelGGL(restart_kernel, dim3(a.nparts), dim3(256), 0, s, a, slot);
  }
  hipLaunchKernelGGL(restart_scalars_kernel, dim3(a.nseg), dim3(256), 0, s, a, slot);
  return static_cast<int>(hipGetLastError());
}

int launch_cg_residual(const CgArgs& a, const double* b, const double* x, hipStream_t s) {
  (void)hipGetLastError();
  // q = K x (its partial sums of x . q into pq, which the next iteration's
  // product overwrites), r = b - q, r . r
  hipLaunchKernelGGL(product_kernel, dim3(a.nrows), dim3(64 * kWaves), 0, s, a, x, a.q, x, a.pq,
                     static_cast<double*>(nullptr));
  hipLaunchKernelGGL(residual_kernel, dim3(a.nparts), dim3(256), 0, s, a, b);
  hipLaunchKernelGGL(residual_scalars_kernel, dim3(a.nseg), dim3(256), 0, s, a);
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
