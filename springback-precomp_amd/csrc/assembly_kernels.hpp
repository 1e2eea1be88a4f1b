// Launch arguments of the assembly kernels (device pointers).
#pragma once

#include <hip/hip_runtime.h>

namespace sbk {

struct AssemblyArgs {
  const double* ke;             // 576 per element, row-major or column-major (ke_column_major)
  const int* row_ptr;           // 3 num_nodes + 1: node n's values are [row_ptr[3n], row_ptr[3n + 3])
  const int* node_ptr;          // num_nodes + 1
  const int* node_incidence;    // e * 8 + a_local, ascending e
  const unsigned long long* inc_pos;  // per incidence: byte m = position of element node m in n's neighbourhood
  long long num_nodes;
  int max_degree;               // the largest neighbourhood (sizes the LDS accumulators)
  int column_major;             // 0: row-major values of K; 1: the reference's column-major values
  int ke_column_major;          // 0: K_e row-major; 1: K_e column-major (Eigen's storage, the reference's Ke)
  double* values;               // nnz
};

struct ForceArgs {
  const double* fe;            // 24 per element
  const int* node_ptr;         // num_nodes + 1
  const int* node_incidence;   // e * 8 + a_local
  long long num_nodes;
  double* fint;                // 3 num_nodes
  double* gross;               // 3 num_nodes: sum_e |f_e| per DOF
};

// stiffness_gross: product = |K| |x| (3 num_nodes), out[0] = the 2-norm of
// product over the free DOFs.
struct AbsProductArgs {
  const int* row_ptr;
  const int* block_col;         // per block (row-major block order): its column node
  const int2* block_trans;      // per block (a, b): {where row 3a sits in column 3b of the
                                // column-major values, the length of b's columns}
  long long num_nodes;
  int max_degree;               // the largest neighbourhood (sizes the LDS staging)
  int group_nodes;              // set by the launch: nodes per workgroup
  const double* values;         // nnz, row-major or the reference's column-major
  int column_major;
  const double* x;              // 3 num_nodes
  const int* free;              // num_free DOF indices
  long long num_free;
  double* product;              // 3 num_nodes
  double* partial;              // 1024
  double* out;                  // 1
};

int launch_assemble_values(const AssemblyArgs& a, hipStream_t s);
int launch_stiffness_gross(const AbsProductArgs& a, hipStream_t s);
int launch_gather_force(const ForceArgs& a, hipStream_t s);
// out[0] = sum x (square = 0) or sqrt(sum x^2) (square = 1), fixed-order tree
// over nparts contiguous slices (nparts <= 1024 partial sums).
// out[0] = max(0, x[0..n)) by per-block maxima (nparts blocks, at most 1024).
int launch_reduce_max(const double* x, long long n, double* partial, int nparts, double* out, hipStream_t s);
int launch_reduce(const double* x, long long n, int square, double* partial, int nparts, double* out,
                  hipStream_t s);
// The same over x[idx[i]], i in [0, n) (NaN for an index outside [0, limit)).
int launch_reduce_gather(const double* x, const int* idx, long long limit, long long n, int square, double* partial,
                         int nparts, double* out, hipStream_t s);
// summary[0] failed elements, [1] lowest failed element, [2] yielding points
int launch_status_summary(const int* status, const int* yielding, long long n, long long* summary, hipStream_t s);

}  // namespace sbk
