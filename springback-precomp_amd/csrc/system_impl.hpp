// The device nonlinear system's state, shared by the plain system
// (system.cpp) and the batched form (batch.cpp).
#pragma once

#include "sparlab_gpu.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "amg.hpp"
#include "assembly.hpp"
#include "assembly_kernels.hpp"
#include "element_kernels.hpp"

namespace sbk {
int capi_fail(int code, const std::string& message);
int capi_prepare_elements(const sparlab_material* material, const sparlab_rule* rule, int kinematics,
                          ElementArgs& a);
}  // namespace sbk

struct sparlab_system {
  sparlab_material material;
  sparlab_rule rule;
  int kinematics = 2;
  int points = 0;
  int fields = 0;
  int64_t nn = 0, ne = 0;
  sbk::HostPattern pattern;
  sbk::ElementArgs args{};
  // device memory
  double* coords = nullptr;
  int32_t* conn = nullptr;
  double* state[2] = {nullptr, nullptr};  // committed = state[cur], trial = state[1 - cur]
  double* alpha[2] = {nullptr, nullptr};
  int cur = 0;
  double *fe = nullptr, *ke = nullptr, *energy = nullptr, *fail_value = nullptr, *gross = nullptr;
  int32_t *yielding = nullptr, *iterations = nullptr, *status = nullptr;
  int32_t* row_ptr = nullptr;
  uint64_t* inc_pos = nullptr;
  int32_t *node_ptr = nullptr, *node_incidence = nullptr, *diag_block = nullptr;
  int32_t* block_col = nullptr;    // per block: its column node
  int2* block_trans = nullptr;     // per block: its place in the column-major values (stiffness_gross)
  double* kproduct = nullptr;      // stiffness_gross scratch (3 nn)
  // the free-free block of the last partition (part_free < 0: none)
  int32_t *part_map = nullptr, *part_counts = nullptr, *part_sums = nullptr, *part_bad = nullptr;
  int32_t *part_ptr = nullptr, *part_index = nullptr, *part_source = nullptr;
  int64_t part_free = -1, part_nnz = 0;
  double* part_work = nullptr;  // the block solve's vectors, partial sums and scalars
  // the block solve's preconditioner (SPARLAB_PRECONDITIONER_*) and the
  // partition's multigrid hierarchy (built at its first solve); for the
  // segmented solve, one hierarchy per segment (built at the first solve
  // with these segment rows)
  int preconditioner = 0;
  sbk::AmgBatch* amg = nullptr;
  sbk::AmgBatch* amg_seg = nullptr;
  std::vector<long long> amg_seg_rows;
  // the segmented block solve's work and metadata (grown as needed)
  double* seg_work = nullptr;
  std::size_t seg_work_doubles = 0;
  unsigned char* seg_meta = nullptr;
  std::size_t seg_meta_bytes = 0;
  double* ws = nullptr;         // the element kernels' workspace
  double* partial = nullptr;    // 1024 partial sums
  double* scalars = nullptr;    // [0] energy, [1] gross
  long long* summary = nullptr; // failed count, lowest failed element, yielding points
  bool evaluated = false;
  bool last_tangent = false;
  bool ke_column_major = false;  // the layout of the last evaluation's element tangents
  // the adjoint: per-element products and whether the workspace still holds
  // the last evaluation (cleared by commit and by new coordinates, material
  // or history)
  double *grad_x = nullptr, *grad_u = nullptr, *grad_p = nullptr;
  bool adjoint_ready = false;
};

namespace sbk {
// dP/d(r0, r45, r90) of a Hill48 material calibrated by r-values (false and
// zeros otherwise).
bool hill_r_derivatives(const sparlab_material& m, double dP[3][36]);
// The launches of one evaluation of the system on `stream` (what
// sparlab_system_evaluate_device queues).
int system_enqueue(sparlab_system* s, const double* u, int want_tangent, int column_major, double* fint,
                   double* values, hipStream_t stream);
// The node-wise assembly of element matrices ke into the system's pattern.
AssemblyArgs assembly_args(const sparlab_system* s, const double* ke, int column_major, double* values,
                           int ke_column_major);
}  // namespace sbk
