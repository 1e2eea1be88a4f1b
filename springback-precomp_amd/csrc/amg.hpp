// The smoothed-aggregation hierarchy of a free-free block: built on the host
// from the block's pattern (once per partition), set up on the device from
// the block's values (once per solve), applied as a V-cycle (once or twice
// per Krylov iteration).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "amg_kernels.hpp"

namespace sbk {

class AmgBatch;

// One problem's hierarchy.
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
  // The same from the block's pattern on the host (ptr: n + 1, index), which
  // the hierarchy uploads and keeps.
  std::string build_host(std::vector<int> ptr, std::vector<int> index, const std::vector<long long>& free_dofs,
                         const std::vector<double>& coords, hipStream_t s);
  bool empty() const { return lv_.empty(); }
  // Rows and nonzeros of every level.
  std::vector<long long> level_rows() const;
  std::vector<long long> level_nnz() const { return nnz_; }

 private:
  friend class AmgBatch;
  std::string build_levels(std::vector<int> ptr, std::vector<int> index, const int* block_ptr,
                           const int* block_index, const std::vector<long long>& free_dofs,
                           const std::vector<double>& coords, hipStream_t s);
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

// The hierarchies of a set of problems (a plain solve's set holds one), set
// up and applied together: every launch covers one level of every problem,
// each with the block decomposition of its own, so that each problem's
// set-up and V-cycles are those it would have alone, bit for bit. Problem p's
// rows sit at row_offset in the vectors the V-cycle reads and writes, its
// block's values at value_offset in the values the set-up reads.
class AmgBatch {
 public:
  AmgBatch() = default;
  AmgBatch(const AmgBatch&) = delete;
  AmgBatch& operator=(const AmgBatch&) = delete;
  ~AmgBatch();

  void add(std::unique_ptr<AmgHierarchy> h, long long row_offset, long long value_offset);
  // The device descriptors and group tables, once every problem is added. 0 on success.
  int finalize();
  // The numeric set-up for the block's values (device); symmetrise: every
  // hierarchy of its block's symmetric part (A + A^T) / 2 (a block that is
  // not symmetric, for BiCGStab: its Chebyshev smoothing needs a real
  // spectrum). 0 on success.
  int setup(const double* block_values, bool symmetrise, hipStream_t s);
  // Per problem: whether the last set-up met a zero pivot on its coarsest
  // level (waits for the stream).
  std::vector<int> singular(hipStream_t s) const;
  // x = M^-1 b (device vectors over every problem's rows) for the problems
  // whose active flag is set (set_active: device flags, null for all).
  int apply(const double* b, double* x, hipStream_t s) const;
  void set_active(const int* active) { active_ = active; }
  int problems() const { return static_cast<int>(problems_.size()); }
  // Rows and nonzeros of every level of the first problem.
  std::vector<long long> level_rows() const;
  std::vector<long long> level_nnz() const;

 private:
  struct Problem {
    std::unique_ptr<AmgHierarchy> h;
    long long row_offset, value_offset;
    int base;  // its first level in the level array
  };
  std::vector<Problem> problems_;
  std::vector<AmgLevel> host_levels_;
  std::vector<AmgGroupEntry> entries_;
  AmgLevel* d_levels_ = nullptr;
  AmgGroupEntry* d_entries_ = nullptr;
  int* d_maps_ = nullptr;
  AmgTables tables_;
  const int* active_ = nullptr;
  void release();
};

// The callback form of apply for the block solve (ctx: an AmgBatch).
int amg_apply(const void* batch, const double* b, double* x, hipStream_t s);

}  // namespace sbk
