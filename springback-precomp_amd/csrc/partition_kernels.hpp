// The free-free block of the assembled matrix (the reference's
// PartitionedFactor, src/fem/Forming.cpp:59-276): its pattern for a list of
// free DOFs, built on the device, and the per-iteration copy of its values.
#pragma once

#include <hip/hip_runtime.h>

namespace sbk {

struct PartitionArgs {
  const int* row_ptr;     // 3 num_nodes + 1 (DOF rows of the symmetric pattern)
  const int* block_col;   // per block (row-major block order): its column node
  long long num_nodes;
  const int* free_dofs;   // num_free, strictly ascending in [0, 3 num_nodes)
  long long num_free;
  int* map;               // 3 num_nodes: DOF -> free position, -1 if not free
  int* counts;            // num_free: the block's entries per free row
  int* block_sums;        // ceil(num_free / 256) + 1: per-block offsets, then the total
  int* bad;               // 1 word: set when the free list is not strictly ascending in range
  int* block_ptr;         // num_free + 1
  int* block_index;       // the block's column (row) indices, renumbered to free positions
  int* source;            // per block entry: its position in the full values
};

// The map, the per-row counts and their offsets: block_ptr and
// block_sums[nb] (the block's nonzero count) and bad[0] on the device.
int launch_partition_count(const PartitionArgs& a, hipStream_t s);
// The block's indices and sources (after block_ptr is known).
int launch_partition_fill(const PartitionArgs& a, hipStream_t s);
// out[p] = values[source[p]] for p < nnz.
int launch_free_block_values(const double* values, const int* source, long long nnz, double* out, hipStream_t s);

}  // namespace sbk
