// Atomic-free assembly on gfx950: the gather form of the reference's
// element-order scatter (src/fem/Assembler.cpp:183-215) and force gather
// (src/fem/NonlinearSystem.hpp:240-262).
//
// assemble_nodes_kernel: one wavefront per node n. Because the pattern is
//   blocked and structurally symmetric, n's three rows of the row-major
//   values (or n's three columns of the reference's column-major ones) are
//   9 deg(n) contiguous values from row_ptr[3n]. The wavefront adds each
//   incident element's 72 values for n (three rows, or three columns, of K_e)
//   into LDS accumulators that start at 0, element after element in ascending
//   element order, so every value is the reference's sum added in the
//   reference's order, bit for bit; then it writes the node's values with
//   contiguous stores. HBM-bound: each element matrix entry is read once and
//   each value written once, and a wavefront issues the loads of up to four
//   incidences before it adds them. When K_e is stored in the order of the
//   values (row-major K_e for row-major values, column-major K_e, Eigen's
//   storage, for the reference's column-major values) the node's 72 entries of
//   K_e are contiguous; otherwise they are three 24-byte runs per row block.
// gather_force_kernel: one thread per DOF: f_int and the gross |f_e| sum.
// sum_squares_kernel: a fixed-shape tree for the norms (deterministic).
#include <algorithm>

#include "assembly_kernels.hpp"

namespace sbk {

constexpr int kNodeWaves = 4;  // nodes (wavefronts) per block
constexpr int kGrossMaxGroup = 64;  // nodes per workgroup of the stiffness_gross product
constexpr int kBatch = 4;      // incidences loaded ahead (more registers cost more than they hide)
constexpr int kXcds = 8;       // MI355X: workgroups go round-robin to 8 XCDs, each with its own L2

__global__ __launch_bounds__(256) void assemble_nodes_kernel(AssemblyArgs A) {
  extern __shared__ double acc_all[];
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  // Values in the other order than K_e's storage read 24-byte runs of K_e
  // 192 bytes apart, so the nodes of an element share its cache lines. There
  // the mapping is XCD-aware: XCD x takes the x-th contiguous range of node
  // blocks, so the nodes of an element (close in number) mostly meet on one
  // L2. Values in K_e's own order read contiguous runs and keep the plain
  // order.
  const bool across = A.column_major != A.ke_column_major;
  const long long per = gridDim.x / kXcds;  // the grid is a multiple of kXcds
  const long long blk = across ? static_cast<long long>(blockIdx.x % kXcds) * per + blockIdx.x / kXcds
                               : static_cast<long long>(blockIdx.x);
  const long long n = blk * kNodeWaves + w;
  if (n >= A.num_nodes) return;  // the whole wavefront
  double* acc = acc_all + w * 9 * A.max_degree;
  const int base = A.row_ptr[3 * n];
  const int len = A.row_ptr[3 * n + 3] - base;  // 9 deg(n)
  const int deg3 = len / 3;                     // the length of one row
  for (int v = lane; v < len; v += 64) acc[v] = 0.0;
  // input j of an incidence (two per lane: j = lane, and j = 64 + lane for
  // lanes 0-7): line k of n's three (a row of K for row-major values, a
  // column for column-major ones), element block m across it, entry k' of
  // the block; in K_e's storage order j runs along the line, otherwise
  // across the lines
  int k[2], m[2], kp[2];
  for (int r = 0; r < 2; ++r) {
    const int j = lane + 64 * r;
    if (!across) { k[r] = j / 24; m[r] = (j % 24) / 3; kp[r] = j % 3; }
    else { m[r] = j / 9; kp[r] = (j / 3) % 3; k[r] = j % 3; }
  }
  const bool second = lane < 8;
  const int i0 = A.node_ptr[n], i1 = A.node_ptr[n + 1];
  for (int ib = i0; ib < i1; ib += kBatch) {
    double val[kBatch][2];
    unsigned long long pos[kBatch];
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
      const int i = ib + b;
      if (i < i1) {
        const int code = A.node_incidence[i];
        const long long e = code >> 3;
        const int nl = code & 7;
        const double* ke = A.ke + e * 576;
        pos[b] = A.inc_pos[i];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          if (r == 0 || second) {
            // row-major values: K_e(3 nl + k, 3 m + k'); column-major values:
            // K_e(3 m + k', 3 nl + k); at its offset in K_e's storage
            const int off = !across ? (3 * nl + k[r]) * 24 + 3 * m[r] + kp[r]
                                    : (3 * m[r] + kp[r]) * 24 + 3 * nl + k[r];
            val[b][r] = ke[off];
          }
        }
      }
    }
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
      if (ib + b < i1) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          if (r == 0 || second) {
            const int pm = static_cast<int>((pos[b] >> (8 * m[r])) & 0xffull);
            const int v = k[r] * deg3 + 3 * pm + kp[r];
            acc[v] = acc[v] + val[b][r];
          }
        }
        // the next incidence adds into values this one's other lanes wrote
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
  }
  for (int v = lane; v < len; v += 64) A.values[base + v] = acc[v];
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

// stiffness_gross (src/fem/NonlinearSystem.hpp:55-67): product[r] = the sum
// over row r's columns c, in ascending order, of |K(r, c)| |x(c)|, each
// product rounded before it is added (the reference's column loop reaches a
// row's entries in that order, with no fused multiply-add), columns with
// x(c) = 0 adding an exact 0 (the reference skips them; products are never
// negative, so the sum is the same).
//
// A workgroup takes A.group_nodes consecutive nodes. Its threads first form
// every product of their rows, a (block, column) run of three rows at a time,
// into LDS in row-major order: row-major values of rows 3a..3a+2 are
// row_len apart; the reference's column-major ones sit together at
// block_trans[k].x + kk block_trans[k].y (block (a, b), column 3b + kk;
// block_trans[k].y is the length of b's columns). All of these loads and
// the x gathers are independent. Then one thread per row adds its row's
// products from LDS in column order.
template <bool CM>
__global__ __launch_bounds__(256) void abs_product_kernel(AbsProductArgs A) {
  extern __shared__ double prod[];
  __shared__ int blk_start[kGrossMaxGroup + 1];  // the group's nodes' first blocks, then its end
  // XCD-aware for the column-major gather: XCD x takes the x-th contiguous
  // range of groups, so neighbouring nodes' column runs meet on one L2
  const long long per = gridDim.x / kXcds;  // the grid is a multiple of kXcds
  const long long grp = CM ? static_cast<long long>(blockIdx.x % kXcds) * per + blockIdx.x / kXcds
                           : static_cast<long long>(blockIdx.x);
  const long long n0 = grp * A.group_nodes;
  if (n0 >= A.num_nodes) return;  // the whole workgroup
  const int g = static_cast<int>(n0 + A.group_nodes <= A.num_nodes ? A.group_nodes : A.num_nodes - n0);
  const int base = A.row_ptr[3 * n0];
  for (int i = threadIdx.x; i <= g; i += blockDim.x) blk_start[i] = A.row_ptr[3 * (n0 + i)] / 9;
  __syncthreads();
  const int b0 = blk_start[0];
  const int runs = 3 * (blk_start[g] - b0);
  // kR runs per thread at a time: their index loads, then their value and x
  // loads, are issued together
  constexpr int kR = 8;
  for (int i0 = threadIdx.x; i0 < runs; i0 += kR * blockDim.x) {
    int blk[kR], col[kR], start[kR], row_len[kR], bcol[kR];
    int2 t[kR];
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      const int i = i0 + u * static_cast<int>(blockDim.x);
      blk[u] = b0 + (i < runs ? i : runs - 1) / 3;  // past the end: the last run again, unused
      int lo = 0, hi = g - 1;  // the node whose blocks hold blk
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk_start[mid] <= blk[u]) lo = mid; else hi = mid - 1;
      }
      start[u] = 9 * blk_start[lo];  // row_ptr of the node's first row
      row_len[u] = 3 * (blk_start[lo + 1] - blk_start[lo]);
      col[u] = 3 * (blk[u] - blk_start[lo]) + i % 3;  // the column within the row: 3k + kk
    }
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      bcol[u] = A.block_col[blk[u]];
      if constexpr (CM) t[u] = A.block_trans[blk[u]];
    }
    double xc[kR], v[kR][3];
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      const int c = col[u] % 3;
      xc[u] = A.x[3ll * bcol[u] + c];
      if constexpr (CM) {
        const double* src = A.values + t[u].x + static_cast<long long>(c) * t[u].y;
#pragma unroll
        for (int l = 0; l < 3; ++l) v[u][l] = src[l];
      } else {
        const double* src = A.values + start[u] + col[u];
#pragma unroll
        for (int l = 0; l < 3; ++l) v[u][l] = src[l * row_len[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      if (i0 + u * static_cast<int>(blockDim.x) >= runs) break;
      const double x_abs = fabs(xc[u]);
      double* dst = prod + (start[u] - base) + col[u];
#pragma unroll
      for (int l = 0; l < 3; ++l) dst[l * row_len[u]] = x_abs != 0.0 ? __dmul_rn(fabs(v[u][l]), x_abs) : 0.0;
    }
  }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) >= 3 * g) return;
  const int a = static_cast<int>(threadIdx.x) / 3, l = static_cast<int>(threadIdx.x) % 3;
  const int row_len = 3 * (blk_start[a + 1] - blk_start[a]);
  const double* p = prod + (9 * blk_start[a] - base) + l * row_len;
  // row_len is a multiple of 3: nine products read ahead of their adds
  double acc = 0.0;
  int q = 0;
  for (; q + 9 <= row_len; q += 9) {
    double t[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) t[j] = p[q + j];
#pragma unroll
    for (int j = 0; j < 9; ++j) acc = __dadd_rn(acc, t[j]);
  }
  for (; q < row_len; q += 3) {
    const double t0 = p[q], t1 = p[q + 1], t2 = p[q + 2];
    acc = __dadd_rn(__dadd_rn(__dadd_rn(acc, t0), t1), t2);
  }
  A.product[3 * n0 + threadIdx.x] = acc;
}

// Sums (of squares, or of values) of x[0, n) in a fixed order: each block of
// the grid sums a contiguous slice with a fixed in-block tree and writes a
// partial; one block then sums the partials the same way.
// With idx, element i is x[idx[i]] (NaN for an index outside [0, limit)):
// the same sums as over the gathered vector, without writing it.
__global__ __launch_bounds__(256) void sum_partials_kernel(const double* __restrict__ x, long long n,
                                                           double* __restrict__ partial, int square,
                                                           const int* __restrict__ idx, long long limit) {
  __shared__ double s[256];
  const long long per = (n + gridDim.x - 1) / gridDim.x;
  const long long b0 = static_cast<long long>(blockIdx.x) * per;
  const long long b1 = b0 + per < n ? b0 + per : n;
  double acc = 0.0;
  for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    double v;
    if (idx) {
      const int k = idx[i];
      v = (k >= 0 && k < limit) ? x[k] : __builtin_nan("");
    } else {
      v = x[i];
    }
    acc = acc + (square ? v * v : v);
  }
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
  if (a.num_nodes == 0) return 0;
  const long long blocks = (a.num_nodes + kNodeWaves - 1) / kNodeWaves;
  const unsigned int grid = static_cast<unsigned int>((blocks + kXcds - 1) / kXcds * kXcds);
  const std::size_t lds = sizeof(double) * kNodeWaves * 9 * static_cast<std::size_t>(a.max_degree);
  hipLaunchKernelGGL(assemble_nodes_kernel, dim3(grid), dim3(64 * kNodeWaves), lds, s, a);
  return static_cast<int>(hipGetLastError());
}

int launch_stiffness_gross(const AbsProductArgs& a, hipStream_t s) {
  (void)hipGetLastError();
  if (a.num_nodes > 0) {
    // nodes per workgroup: their products (at most 9 max_degree each) within
    // 16 KB of LDS (8 nodes of degree 27: measured best against 4 and 16 on
    // the 100 x 100 x 4 plate), at most kGrossMaxGroup (3 rows per node <= 256
    // threads)
    AbsProductArgs b = a;
    const long long per_node = 9ll * (a.max_degree > 0 ? a.max_degree : 1) * static_cast<long long>(sizeof(double));
    b.group_nodes = static_cast<int>(std::max(1ll, std::min<long long>(kGrossMaxGroup, 16384 / per_node)));
    const long long groups = (a.num_nodes + b.group_nodes - 1) / b.group_nodes;
    const unsigned int grid = static_cast<unsigned int>((groups + kXcds - 1) / kXcds * kXcds);
    const std::size_t lds = static_cast<std::size_t>(per_node) * b.group_nodes;
    if (a.column_major)
      hipLaunchKernelGGL(abs_product_kernel<true>, dim3(grid), dim3(256), lds, s, b);
    else
      hipLaunchKernelGGL(abs_product_kernel<false>, dim3(grid), dim3(256), lds, s, b);
  }
  const int parts = static_cast<int>(a.num_free / 256 + 1 < 1024 ? a.num_free / 256 + 1 : 1024);
  const int rc = static_cast<int>(hipGetLastError());
  if (rc != 0) return rc;
  return launch_reduce_gather(a.product, a.free, 3 * a.num_nodes, a.num_free, 1, a.partial, parts, a.out, s);
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
  return launch_reduce_gather(x, nullptr, 0, n, square, partial, nparts, out, s);
}

int launch_reduce_gather(const double* x, const int* idx, long long limit, long long n, int square, double* partial,
                         int nparts, double* out, hipStream_t s) {
  (void)hipGetLastError();
  if (nparts < 1) nparts = 1;
  if (nparts > 1024) nparts = 1024;
  hipLaunchKernelGGL(sum_partials_kernel, dim3(nparts), dim3(256), 0, s, x, n, partial, square, idx, limit);
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, partial, nparts, out, square);
  return static_cast<int>(hipGetLastError());
}

// The largest of x[i * stride] over i in [0, n) (max is exact in any order):
// per-block maxima, then one block over them; 0 for n = 0 (the values are
// non-negative: equivalent plastic strains).
__global__ __launch_bounds__(256) void max_partials_kernel(const double* __restrict__ x, long long n,
                                                           double* __restrict__ partial) {
  __shared__ double s[256];
  const long long per = (n + gridDim.x - 1) / gridDim.x;
  const long long b0 = static_cast<long long>(blockIdx.x) * per;
  const long long b1 = b0 + per < n ? b0 + per : n;
  double acc = 0.0;
  for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) acc = fmax(acc, x[i]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void max_kernel(const double* __restrict__ x, int n, double* __restrict__ out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc = fmax(acc, x[i]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

int launch_reduce_max(const double* x, long long n, double* partial, int nparts, double* out, hipStream_t s) {
  (void)hipGetLastError();
  if (nparts < 1) nparts = 1;
  if (nparts > 1024) nparts = 1024;
  hipLaunchKernelGGL(max_partials_kernel, dim3(nparts), dim3(256), 0, s, x, n, partial);
  hipLaunchKernelGGL(max_kernel, dim3(1), dim3(256), 0, s, partial, nparts, out);
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
