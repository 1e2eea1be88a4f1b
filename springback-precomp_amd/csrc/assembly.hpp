// Host-side sparsity pattern and the gather maps of the assembly kernels.
//
// The pattern is the reference's (src/fem/Assembler.cpp:86-173): three DOFs
// per node, a dense 3 x 3 block for every pair of nodes that share an
// element, each node's neighbourhood sorted ascending. It is structurally
// symmetric, so the reference's column-major (outer, inner) arrays are also
// the row-major (row_ptr, col_idx) arrays of the same matrix, bit for bit.
//
// The gather maps turn the reference's element-order scatter
// (Assembler.cpp:183-215) into an atomic-free gather: every 3 x 3 block of the
// matrix lists its contributions (element, local row node, local column node)
// in ascending element order, and every node lists its (element, local node)
// incidences in ascending order, so a gather that sums them in list order
// adds the same numbers in the same order as the serial scatter.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sbk {

struct HostPattern {
  int64_t num_nodes = 0, num_elements = 0;
  std::vector<int32_t> row_ptr;   // 3 num_nodes + 1 (== the reference's outer)
  std::vector<int32_t> col_idx;   // nnz (== the reference's inner)
  // blocks in row-major order: block i = (row node a, k-th neighbour b)
  std::vector<int32_t> block_row_node;   // a
  std::vector<int32_t> block_pos;        // row_ptr[3a] + 3k: the first value of the block's first row
  std::vector<int32_t> block_row_len;    // 3 deg(a): the stride between the block's rows
  std::vector<int32_t> block_ptr;        // contributions of block i: [block_ptr[i], block_ptr[i+1])
  std::vector<int32_t> block_contrib;    // e * 64 + a_local * 8 + b_local, ascending e
  // node incidences for f_int: [node_ptr[n], node_ptr[n+1]) of e * 8 + a_local, ascending e
  std::vector<int32_t> node_ptr;
  std::vector<int32_t> node_incidence;
  // for every incidence (e, n): byte m of inc_pos is the position of element
  // node m within n's sorted neighbourhood (the node-wise assembly's map)
  std::vector<uint64_t> inc_pos;
  int32_t max_degree = 0;                // the largest neighbourhood
  // per block (row node a, k-th neighbour b), row-major block order: b, and
  // the index of a within b's neighbourhood (where block (a, b)'s rows sit in
  // the reference's column-major values)
  std::vector<int32_t> block_col_node;
  std::vector<int32_t> block_mirror;
  // the diagonal block of every node (contact blocks)
  std::vector<int32_t> diag_block;       // node -> block index of (node, node)
  std::vector<int32_t> block_offset;     // the reference's per-element (a, b) offsets, 64 per element
};

// Builds the pattern and the maps; returns false (why set) for invalid input.
bool build_pattern(int64_t num_nodes, int64_t num_elements, const int32_t* conn, HostPattern& out, std::string* why);

}  // namespace sbk
