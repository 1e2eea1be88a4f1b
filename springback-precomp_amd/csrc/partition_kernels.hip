// The free-free block of the assembled matrix on the device: the pattern of
// PartitionedFactor::build (src/fem/Forming.cpp:208-236) for a list of free
// DOFs, and the copy of PartitionedFactor::assemble (:89-98).
//
// The system's pattern is symmetric, so its row pointers and column indices
// are also its column pointers and row indices: the reference walks the
// columns of its column-major matrix, this walks the rows of the same
// arrays, and position q means the same entry of either. Row d = 3 a + l of
// node a holds its neighbour blocks in ascending order, three columns each,
// so entry j of the row is column 3 block_col[blocks before a + j / 3] + j % 3.
// The block keeps the layout of the values it is copied from: column-major
// values give the reference's kff_ (its columns), row-major ones its rows.
//
// Every step is deterministic (counts, block offsets, in-block scans, and a
// ballot within each row's wavefront): no atomics.

#include "partition_kernels.hpp"

namespace sbk {

namespace {

constexpr int kRowWaves = 4;  // free rows per 256-thread block in the row kernels

// map = -1 everywhere; bad = 0.
__global__ __launch_bounds__(256) void partition_clear_kernel(int* __restrict__ map, long long n,
                                                              int* __restrict__ bad) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) map[i] = -1;
  if (i == 0) bad[0] = 0;
}

// map[free[i]] = i; a DOF outside [0, n) or a list that does not strictly
// ascend sets bad (every such thread stores the same 1).
__global__ __launch_bounds__(256) void partition_map_kernel(const int* __restrict__ free_dofs, long long nf,
                                                            long long n, int* __restrict__ map,
                                                            int* __restrict__ bad) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nf) return;
  const int f = free_dofs[i];
  const bool ok = f >= 0 && f < n && (i == 0 || free_dofs[i - 1] < f);
  if (ok)
    map[f] = static_cast<int>(i);
  else
    bad[0] = 1;
}

// The row's entry j: its column.
__device__ __forceinline__ int entry_column(const int* __restrict__ block_col, int blk0, int j) {
  return 3 * block_col[blk0 + j / 3] + j % 3;
}

// One wavefront per free row: counts[c] = the row's entries in free columns.
__global__ __launch_bounds__(256) void partition_count_kernel(PartitionArgs A) {
  const int lane = threadIdx.x & 63;
  const long long c = static_cast<long long>(blockIdx.x) * kRowWaves + (threadIdx.x >> 6);
  if (c >= A.num_free) return;  // the whole wavefront
  const int d = A.free_dofs[c];
  int count = 0;
  if (d >= 0 && d < 3 * A.num_nodes) {  // else the list is refused; stay in bounds
    const int first = A.row_ptr[d];
    const int len = A.row_ptr[d + 1] - first;
    const int blk0 = A.row_ptr[3 * (d / 3)] / 9;
    for (int j0 = 0; j0 < len; j0 += 64) {
      const int j = j0 + lane;
      const bool keep = j < len && A.map[entry_column(A.block_col, blk0, j)] >= 0;
      count += __popcll(__ballot(keep));
    }
  }
  if (lane == 0) A.counts[c] = count;
}

// block_sums[b] = the counts of rows [256 b, 256 b + 256).
__global__ __launch_bounds__(256) void partition_block_sums_kernel(const int* __restrict__ counts, long long nf,
                                                                   int* __restrict__ block_sums) {
  __shared__ int s[256];
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  s[threadIdx.x] = i < nf ? counts[i] : 0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = s[0];
}

// Exclusive offsets of the nb block sums in place and the total in
// block_sums[nb] (one block; each thread scans a contiguous run).
__global__ __launch_bounds__(256) void partition_offsets_kernel(int* __restrict__ block_sums, int nb) {
  __shared__ int part[256];
  const int per = (nb + 255) / 256;
  const int b0 = static_cast<int>(threadIdx.x) * per;
  const int b1 = b0 + per < nb ? b0 + per : nb;
  int run = 0;
  for (int b = b0; b < b1; ++b) run += block_sums[b];
  part[threadIdx.x] = run;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < 256; ++k) {
      const int v = part[k];
      part[k] = acc;
      acc += v;
    }
    block_sums[nb] = acc;
  }
  __syncthreads();
  run = part[threadIdx.x];
  for (int b = b0; b < b1; ++b) {
    const int v = block_sums[b];
    block_sums[b] = run;
    run += v;
  }
}

// block_ptr[c + 1] = the block offset + the inclusive scan of the counts
// within the block; block_ptr[0] = 0.
__global__ __launch_bounds__(256) void partition_ptr_kernel(const int* __restrict__ counts, long long nf,
                                                            const int* __restrict__ block_sums,
                                                            int* __restrict__ block_ptr) {
  __shared__ int scan[256];
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  scan[threadIdx.x] = i < nf ? counts[i] : 0;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int add = static_cast<int>(threadIdx.x) >= off ? scan[threadIdx.x - off] : 0;
    __syncthreads();
    scan[threadIdx.x] += add;
    __syncthreads();
  }
  if (i < nf) block_ptr[i + 1] = block_sums[blockIdx.x] + scan[threadIdx.x];
  if (i == 0) block_ptr[0] = 0;
}

// One wavefront per free row: its kept entries in order, at block_ptr[c].
__global__ __launch_bounds__(256) void partition_fill_kernel(PartitionArgs A) {
  const int lane = threadIdx.x & 63;
  const long long c = static_cast<long long>(blockIdx.x) * kRowWaves + (threadIdx.x >> 6);
  if (c >= A.num_free) return;
  const int d = A.free_dofs[c];
  if (d < 0 || d >= 3 * A.num_nodes) return;  // a refused list; nothing written
  const int first = A.row_ptr[d];
  const int len = A.row_ptr[d + 1] - first;
  const int blk0 = A.row_ptr[3 * (d / 3)] / 9;
  int base = A.block_ptr[c];
  for (int j0 = 0; j0 < len; j0 += 64) {
    const int j = j0 + lane;
    const int r = j < len ? A.map[entry_column(A.block_col, blk0, j)] : -1;
    const unsigned long long kept = __ballot(r >= 0);
    const int before = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned int>(kept >> 32),
                                                 __builtin_amdgcn_mbcnt_lo(static_cast<unsigned int>(kept), 0u));
    if (r >= 0) {
      A.block_index[base + before] = r;
      A.source[base + before] = first + j;
    }
    base += __popcll(kept);
  }
}

__global__ __launch_bounds__(256) void free_block_values_kernel(const double* __restrict__ values,
                                                                const int* __restrict__ source, long long nnz,
                                                                double* __restrict__ out) {
  const long long p = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (p < nnz) out[p] = values[source[p]];
}

unsigned int blocks_of(long long n, long long per) { return static_cast<unsigned int>((n + per - 1) / per); }

}  // namespace

int launch_partition_count(const PartitionArgs& a, hipStream_t s) {
  (void)hipGetLastError();
  const long long n = 3 * a.num_nodes;
  const long long nf = a.num_free;
  const long long nb = (nf + 255) / 256;
  hipLaunchKernelGGL(partition_clear_kernel, dim3(blocks_of(n > 0 ? n : 1, 256)), dim3(256), 0, s, a.map, n, a.bad);
  if (nf > 0) {
    hipLaunchKernelGGL(partition_map_kernel, dim3(blocks_of(nf, 256)), dim3(256), 0, s, a.free_dofs, nf, n, a.map,
                       a.bad);
    hipLaunchKernelGGL(partition_count_kernel, dim3(blocks_of(nf, kRowWaves)), dim3(64 * kRowWaves), 0, s, a);
    hipLaunchKernelGGL(partition_block_sums_kernel, dim3(static_cast<unsigned int>(nb)), dim3(256), 0, s, a.counts,
                       nf, a.block_sums);
  }
  hipLaunchKernelGGL(partition_offsets_kernel, dim3(1), dim3(256), 0, s, a.block_sums, static_cast<int>(nb));
  hipLaunchKernelGGL(partition_ptr_kernel, dim3(static_cast<unsigned int>(nb > 0 ? nb : 1)), dim3(256), 0, s,
                     a.counts, nf, a.block_sums, a.block_ptr);
  return static_cast<int>(hipGetLastError());
}

int launch_partition_fill(const PartitionArgs& a, hipStream_t s) {
  (void)hipGetLastError();
  if (a.num_free > 0)
    hipLaunchKernelGGL(partition_fill_kernel, dim3(blocks_of(a.num_free, kRowWaves)), dim3(64 * kRowWaves), 0, s, a);
  return static_cast<int>(hipGetLastError());
}

int launch_free_block_values(const double* values, const int* source, long long nnz, double* out, hipStream_t s) {
  (void)hipGetLastError();
  if (nnz > 0)
    hipLaunchKernelGGL(free_block_values_kernel, dim3(blocks_of(nnz, 256)), dim3(256), 0, s, values, source, nnz,
                       out);
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
