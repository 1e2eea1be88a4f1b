// The smoothed-aggregation hierarchy of a free-free block: built on the host
// from the block's pattern (once per partition), set up on the device from
// the block's values (once per solve), applied as a V-cycle (once or twice
// per Krylov iteration).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "amg_kernels.hpp"

namespace sbk {

class AmgHierarchy {
 public:
  AmgHierarchy() = default;
  AmgHierarchy(const AmgHierarchy&) = delete;
  AmgHierarchy& operator=(const AmgHierarchy&) = delete;
  ~AmgHierarchy();

  // The hierarchy of a block of n rows (block_ptr, block_index: its device
  // pattern, nnz entries; the rows of one node share their columns) whose
  // row i is the full DOF free_dofs[i] (host, ascending; node = DOF / 3,
  // component = DOF % 3) at nodes with coordinates coords (host, 3 per node):
  // the near-nullspace is the six rigid-body modes. An empty string, or why
  // no hierarchy could be built (nothing is kept then).
  std::string build(const int* block_ptr, const int* block_index, long long n, long long nnz,
                    const std::vector<long long>& free_dofs, const std::vector<double>& coords, hipStream_t s);
  // The numeric set-up for the block's values (device, nnz); symmetrise: the
  // hierarchy of the block's symmetric part (A + A^T) / 2 (a block that is
  // not symmetric, for BiCGStab: its Chebyshev smoothing needs a real
  // spectrum). 0 on success.
  int setup(const double* block_values, bool symmetrise, hipStream_t s);
  // Whether the last set-up met a zero pivot on the coarsest level (waits
  // for the stream).
  bool singular(hipStream_t s) const;
  // x = M^-1 b (device vectors of n).
  int apply(const double* b, double* x, hipStream_t s) const;
  bool empty() const { return lv_.empty(); }
  // Rows and nonzeros of every level.
  std::vector<long long> level_rows() const;
  std::vector<long long> level_nnz() const { return nnz_; }

 private:
  std::vector<AmgLevel> lv_;
  const int* transpose_ = nullptr;  // level 0: each entry's mirror (j, i), -1 when absent
  double* symmetric_ = nullptr;     // level 0: (A + A^T) / 2
  long long nnz0_ = 0;
  std::vector<long long> nnz_;
  std::vector<void*> owned_;
  template <class T>
  T* upload(const std::vector<T>& v, hipStream_t s, bool& ok);
  template <class T>
  T* alloc(std::size_t count, bool& ok);
  void release();
};

// The callback form of apply for the block solve.
int amg_apply(const void* hierarchy, const double* b, double* x, hipStream_t s);

}  // namespace sbk
