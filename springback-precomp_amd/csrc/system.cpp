// The nonlinear system on the device: the drop-in for the element loop of
// detail::NonlinearSystem (src/fem/NonlinearSystem.hpp:192-268) with its
// committed history (:119-182) and the Assembler's pattern and scatter
// (src/fem/Assembler.cpp:86-215, 271-292).
//
// Device-resident: the mesh, the pattern and gather maps, the committed and
// the trial history (swapped on commit), and the per-element outputs that the
// assembly gathers. One evaluation = the element kernels, the force gather,
// the value gather (with a tangent) and three small reductions.
#include "sparlab_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "adjoint_kernels.hpp"
#include "assembly_kernels.hpp"
#include "partition_kernels.hpp"
#include "solve_kernels.hpp"
#include "system_impl.hpp"

namespace {

template <class T>
bool dalloc(T*& p, std::size_t count) {
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)) == hipSuccess;
}
template <class T>
bool upload(T*& p, const std::vector<T>& v) {
  if (!dalloc(p, v.size())) return false;
  return v.empty() || hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Per block (a, b) in row-major block order: where row 3a's entry of column
// 3b sits in the column-major values (column 3b's start plus 3 times a's
// place in b's neighbourhood), and the length of b's columns (3 deg(b)).
std::vector<int2> block_trans(const sbk::HostPattern& p) {
  std::vector<int2> t(p.block_mirror.size());
  for (std::size_t k = 0; k < t.size(); ++k) {
    const int32_t b = p.block_col_node[k];
    const int32_t start = p.row_ptr[3 * static_cast<std::size_t>(b)];
    t[k] = make_int2(start + 3 * p.block_mirror[k], p.row_ptr[3 * static_cast<std::size_t>(b) + 1] - start);
  }
  return t;
}

// 24 x 24 element tangents, column-major to row-major, in place (host).
void transpose_tangents(double* ke, int64_t count) {
  for (int64_t e = 0; e < count; ++e) {
    double* m = ke + 576 * e;
    for (int i = 0; i < 24; ++i)
      for (int j = i + 1; j < 24; ++j) std::swap(m[i * 24 + j], m[j * 24 + i]);
  }
}

void release_multigrid(sparlab_system* s) {
  delete s->amg;
  s->amg = nullptr;
  delete s->amg_seg;
  s->amg_seg = nullptr;
  s->amg_seg_rows.clear();
}

void release_partition(sparlab_system* s) {
  release_multigrid(s);
  dfree(s->part_counts); dfree(s->part_sums); dfree(s->part_bad);
  dfree(s->part_ptr); dfree(s->part_index); dfree(s->part_source); dfree(s->part_work);
  s->part_free = -1;
  s->part_nnz = 0;
}

void release(sparlab_system* s) {
  release_partition(s);
  dfree(s->part_map);
  dfree(s->coords); dfree(s->conn);
  for (int i = 0; i < 2; ++i) { dfree(s->state[i]); dfree(s->alpha[i]); }
  dfree(s->fe); dfree(s->ke); dfree(s->energy); dfree(s->fail_value); dfree(s->gross);
  dfree(s->yielding); dfree(s->iterations); dfree(s->status);
  dfree(s->row_ptr); dfree(s->inc_pos);
  dfree(s->node_ptr); dfree(s->node_incidence); dfree(s->diag_block);
  dfree(s->block_col); dfree(s->block_trans); dfree(s->kproduct);
  dfree(s->ws); dfree(s->partial); dfree(s->scalars); dfree(s->summary);
  dfree(s->grad_x); dfree(s->grad_u); dfree(s->grad_p);
  dfree(s->seg_work); dfree(s->seg_meta);
  s->seg_work_doubles = s->seg_meta_bytes = 0;
}

}  // namespace

namespace sbk {

AssemblyArgs assembly_args(const sparlab_system* s, const double* ke, int column_major, double* values,
                           int ke_column_major) {
  AssemblyArgs g{};
  g.ke = ke;
  g.row_ptr = s->row_ptr;
  g.node_ptr = s->node_ptr;
  g.node_incidence = s->node_incidence;
  g.inc_pos = reinterpret_cast<const unsigned long long*>(s->inc_pos);
  g.num_nodes = s->nn;
  g.max_degree = s->pattern.max_degree;
  g.column_major = column_major ? 1 : 0;
  g.ke_column_major = ke_column_major ? 1 : 0;
  g.values = values;
  return g;
}

int system_enqueue(sparlab_system* s, const double* u, int want_tangent, int column_major, double* fint,
                   double* values, hipStream_t st) {
  ElementArgs a = s->args;
  a.coords = s->coords;
  a.conn = s->conn;
  a.u = u;
  a.ne = s->ne;
  a.state_in = s->state[s->cur];
  a.state_stride = s->ne * s->points;
  a.alpha_in = s->alpha[s->cur];
  a.alpha_out = s->alpha[1 - s->cur];
  a.state_out = s->state[1 - s->cur];
  a.force = s->fe;
  a.tangent = want_tangent ? s->ke : nullptr;
  // the element tangents in the layout of the assembled values (the
  // assembly then copies node blocks straight through)
  a.tangent_column_major = column_major ? 1 : 0;
  a.energy = s->energy;
  a.yielding = s->yielding;
  a.iterations = s->iterations;
  a.status = s->status;
  a.fail_value = s->fail_value;
  set_workspace(a, s->ws);
  if (launch_local_solve(a, st) != 0 || launch_condense(a, want_tangent != 0, st) != 0)
    return capi_fail(SPARLAB_E_DEVICE, "element kernel launch failed");
  ForceArgs f{s->fe, s->node_ptr, s->node_incidence, static_cast<long long>(s->nn), fint, s->gross};
  if (launch_gather_force(f, st) != 0) return capi_fail(SPARLAB_E_DEVICE, "force gather launch failed");
  if (want_tangent) {
    if (launch_assemble_values(assembly_args(s, s->ke, column_major, values, column_major), st) != 0)
      return capi_fail(SPARLAB_E_DEVICE, "assembly launch failed");
  }
  const int parts = static_cast<int>(std::min<int64_t>(1024, (3 * s->nn + 4095) / 4096 + 1));
  if (launch_reduce(s->gross, 3 * s->nn, 1, s->partial, parts, s->scalars + 1, st) != 0 ||
      launch_reduce(s->energy, s->ne, 0, s->partial, static_cast<int>(std::min<int64_t>(1024, s->ne / 4096 + 1)),
                         s->scalars, st) != 0 ||
      launch_status_summary(s->status, s->yielding, s->ne, s->summary, st) != 0)
    return capi_fail(SPARLAB_E_DEVICE, "reduction launch failed");
  s->evaluated = true;
  s->last_tangent = want_tangent != 0;
  if (want_tangent) s->ke_column_major = column_major != 0;
  s->adjoint_ready = true;
  return SPARLAB_OK;
}

}  // namespace sbk

extern "C" {

int sparlab_system_create(const double* coords, int64_t num_nodes, const int32_t* conn, int64_t num_elements,
                          const sparlab_material* material, const sparlab_rule* rule, int kinematics,
                          sparlab_system** out) {
  if (out == nullptr || coords == nullptr || conn == nullptr) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  *out = nullptr;
  auto* s = new sparlab_system();
  s->material = *material;
  s->rule = *rule;
  s->kinematics = kinematics;
  s->nn = num_nodes;
  s->ne = num_elements;
  int st = sbk::capi_prepare_elements(material, rule, kinematics, s->args);
  if (st != SPARLAB_OK) { delete s; return st; }
  s->points = s->args.rule.n;
  s->fields = sbk::state_fields(s->args.mat.kin_fields);
  std::string why;
  bool built = false;
  try {
    built = sbk::build_pattern(num_nodes, num_elements, conn, s->pattern, &why);
  } catch (const std::exception& ex) {
    delete s;
    return sbk::capi_fail(SPARLAB_E_DEVICE, std::string("the pattern: host failure: ") + ex.what());
  }
  if (!built) {
    delete s;
    return sbk::capi_fail(SPARLAB_E_MESH, why);
  }
  const std::size_t ne = static_cast<std::size_t>(num_elements), nn = static_cast<std::size_t>(num_nodes);
  const std::size_t hist = static_cast<std::size_t>(s->fields) * ne * static_cast<std::size_t>(s->points);
  bool ok = dalloc(s->coords, 3 * nn) && dalloc(s->conn, 8 * ne) && dalloc(s->state[0], hist) &&
            dalloc(s->state[1], hist) && dalloc(s->alpha[0], 9 * ne) && dalloc(s->alpha[1], 9 * ne) &&
            dalloc(s->fe, 24 * ne) && dalloc(s->ke, 576 * ne) && dalloc(s->energy, ne) &&
            dalloc(s->fail_value, ne) && dalloc(s->gross, 3 * nn) && dalloc(s->yielding, ne) &&
            dalloc(s->iterations, ne) && dalloc(s->status, ne) &&
            dalloc(s->ws, static_cast<std::size_t>(sbk::workspace_doubles(num_elements, s->points))) &&
            dalloc(s->partial, 1024) &&
            dalloc(s->scalars, 3) && dalloc(s->summary, 3) && upload(s->row_ptr, s->pattern.row_ptr) &&
            upload(s->inc_pos, s->pattern.inc_pos) && upload(s->node_ptr, s->pattern.node_ptr) &&
            upload(s->node_incidence, s->pattern.node_incidence) && upload(s->diag_block, s->pattern.diag_block) &&
            upload(s->block_col, s->pattern.block_col_node) && upload(s->block_trans, block_trans(s->pattern)) &&
            dalloc(s->kproduct, 3 * nn);
  ok = ok && hipMemcpy(s->coords, coords, 3 * nn * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->conn, conn, 8 * ne * sizeof(int32_t), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemset(s->state[0], 0, hist * sizeof(double)) == hipSuccess &&
       hipMemset(s->alpha[0], 0, 9 * ne * sizeof(double)) == hipSuccess;
  if (!ok) {
    release(s);
    delete s;
    return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation or upload failed");
  }
  *out = s;
  return SPARLAB_OK;
}

void sparlab_system_destroy(sparlab_system* s) {
  if (!s) return;
  release(s);
  delete s;
}

int64_t sparlab_system_nnz(const sparlab_system* s) { return s ? static_cast<int64_t>(s->pattern.col_idx.size()) : -1; }

int sparlab_system_pattern(const sparlab_system* s, int32_t* row_ptr, int32_t* col_idx) {
  if (!s || !row_ptr || !col_idx) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  std::memcpy(row_ptr, s->pattern.row_ptr.data(), s->pattern.row_ptr.size() * sizeof(int32_t));
  std::memcpy(col_idx, s->pattern.col_idx.data(), s->pattern.col_idx.size() * sizeof(int32_t));
  return SPARLAB_OK;
}

int64_t sparlab_pattern_build(int64_t num_nodes, int64_t num_elements, const int32_t* conn, int32_t* row_ptr,
                              int32_t* col_idx, int32_t* block_offset) {
  if (conn == nullptr && num_elements > 0) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null connectivity");
  sbk::HostPattern p;
  std::string why;
  if (!sbk::build_pattern(num_nodes, num_elements, conn, p, &why)) return sbk::capi_fail(SPARLAB_E_MESH, why);
  if (row_ptr && col_idx) {
    std::memcpy(row_ptr, p.row_ptr.data(), p.row_ptr.size() * sizeof(int32_t));
    std::memcpy(col_idx, p.col_idx.data(), p.col_idx.size() * sizeof(int32_t));
  }
  if (block_offset) std::memcpy(block_offset, p.block_offset.data(), p.block_offset.size() * sizeof(int32_t));
  return static_cast<int64_t>(p.col_idx.size());
}

int sparlab_system_diagonal_blocks(const sparlab_system* s, int32_t* first_value) {
  if (!s || !first_value) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  for (std::size_t n = 0; n < s->pattern.diag_block.size(); ++n)
    first_value[n] = s->pattern.block_pos[static_cast<std::size_t>(s->pattern.diag_block[n])];
  return SPARLAB_OK;
}

int sparlab_system_evaluate_device(sparlab_system* s, const double* u, int want_tangent, int column_major,
                                   double* fint, double* values, void* stream) {
  if (!s || !u || !fint) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (want_tangent && !values) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "want_tangent needs a values buffer");
  return sbk::system_enqueue(s, u, want_tangent, column_major, fint, values, static_cast<hipStream_t>(stream));
}

int sparlab_system_assemble_device(sparlab_system* s, const double* fe, const double* ke, int layout,
                                   double* fint, double* values, double* gross, void* stream) {
  if (!s || (!fe && fint) || (!ke && values)) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (layout & ~(SPARLAB_VALUES_COLUMN_MAJOR | SPARLAB_KE_COLUMN_MAJOR))
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "unknown layout flags");
  const int column_major = (layout & SPARLAB_VALUES_COLUMN_MAJOR) != 0;
  const int ke_column_major = (layout & SPARLAB_KE_COLUMN_MAJOR) != 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (fint) {
    sbk::ForceArgs f{fe, s->node_ptr, s->node_incidence, static_cast<long long>(s->nn), fint, s->gross};
    if (sbk::launch_gather_force(f, st) != 0) return sbk::capi_fail(SPARLAB_E_DEVICE, "force gather launch failed");
    const int parts = static_cast<int>(std::min<int64_t>(1024, (3 * s->nn + 4095) / 4096 + 1));
    if (gross && sbk::launch_reduce(s->gross, 3 * s->nn, 1, s->partial, parts, gross, st) != 0)
      return sbk::capi_fail(SPARLAB_E_DEVICE, "reduction launch failed");
  }
  if (values) {
    if (sbk::launch_assemble_values(sbk::assembly_args(s, ke, column_major, values, ke_column_major), st) != 0)
      return sbk::capi_fail(SPARLAB_E_DEVICE, "assembly launch failed");
  }
  return SPARLAB_OK;
}

int sparlab_system_stiffness_gross_device(sparlab_system* s, const double* values, int column_major, const double* x,
                                         const int32_t* free_dofs, int64_t num_free, double* out, double* product,
                                         void* stream) {
  if (!s || !values || !x || !out || (num_free > 0 && !free_dofs)) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (num_free < 0 || num_free > 3 * s->nn) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "bad free DOF count");
  sbk::AbsProductArgs a{};
  a.row_ptr = s->row_ptr;
  a.block_col = s->block_col;
  a.block_trans = s->block_trans;
  a.num_nodes = s->nn;
  a.max_degree = s->pattern.max_degree;
  a.values = values;
  a.column_major = column_major ? 1 : 0;
  a.x = x;
  a.free = free_dofs;
  a.num_free = num_free;
  a.product = product ? product : s->kproduct;
  a.partial = s->partial;
  a.out = out;
  if (sbk::launch_stiffness_gross(a, static_cast<hipStream_t>(stream)) != 0)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "stiffness_gross launch failed");
  return SPARLAB_OK;
}

int sparlab_system_partition_device(sparlab_system* s, const int32_t* free_dofs, int64_t num_free,
                                    int64_t* block_nnz, void* stream) {
  if (!s || !block_nnz || (num_free > 0 && !free_dofs)) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  const int64_t n = 3 * s->nn;
  if (num_free < 0 || num_free > n) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "bad free DOF count");
  if (s->pattern.col_idx.size() > static_cast<std::size_t>(INT32_MAX))
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the pattern is too large for 32-bit block sources");
  hipStream_t st = static_cast<hipStream_t>(stream);
  // the previous partition goes first (its buffers are freed with the
  // stream's work on them finished)
  if (hipStreamSynchronize(st) != hipSuccess) return sbk::capi_fail(SPARLAB_E_DEVICE, "stream synchronization failed");
  release_partition(s);
  *block_nnz = 0;
  const std::size_t nf = static_cast<std::size_t>(num_free);
  const std::size_t nb = (nf + 255) / 256;
  if ((!s->part_map && !dalloc(s->part_map, static_cast<std::size_t>(n))) || !dalloc(s->part_counts, nf) ||
      !dalloc(s->part_sums, nb + 1) || !dalloc(s->part_bad, 1) || !dalloc(s->part_ptr, nf + 1)) {
    release_partition(s);
    return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation failed");
  }
  sbk::PartitionArgs a{};
  a.row_ptr = s->row_ptr;
  a.block_col = s->block_col;
  a.num_nodes = s->nn;
  a.free_dofs = free_dofs;
  a.num_free = num_free;
  a.map = s->part_map;
  a.counts = s->part_counts;
  a.block_sums = s->part_sums;
  a.bad = s->part_bad;
  a.block_ptr = s->part_ptr;
  int32_t head[2] = {0, 0};  // the block's nonzero count, the refusal flag
  if (sbk::launch_partition_count(a, st) != 0 ||
      hipMemcpyAsync(&head[0], s->part_sums + nb, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&head[1], s->part_bad, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    release_partition(s);
    return sbk::capi_fail(SPARLAB_E_DEVICE, "the partition's count failed");
  }
  if (head[1] != 0) {
    release_partition(s);
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the free DOFs are not strictly ascending in [0, 3 num_nodes)");
  }
  const std::size_t nnz = static_cast<std::size_t>(head[0]);
  if (!dalloc(s->part_index, nnz) || !dalloc(s->part_source, nnz) ||
      !dalloc(s->part_work, static_cast<std::size_t>(sbk::cg_work_doubles(num_free)))) {
    release_partition(s);
    return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation failed");
  }
  a.block_index = s->part_index;
  a.source = s->part_source;
  if (sbk::launch_partition_fill(a, st) != 0) {
    release_partition(s);
    return sbk::capi_fail(SPARLAB_E_DEVICE, "the partition's fill failed");
  }
  s->part_free = num_free;
  s->part_nnz = static_cast<int64_t>(nnz);
  *block_nnz = s->part_nnz;
  return SPARLAB_OK;
}

int sparlab_system_partition_pattern_device(sparlab_system* s, int32_t* block_ptr, int32_t* block_index,
                                            void* stream) {
  if (!s || !block_ptr || (s->part_nnz > 0 && !block_index))
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (s->part_free < 0) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "no partition");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(block_ptr, s->part_ptr, static_cast<std::size_t>(s->part_free + 1) * sizeof(int32_t),
                     hipMemcpyDeviceToDevice, st) != hipSuccess ||
      (s->part_nnz > 0 &&
       hipMemcpyAsync(block_index, s->part_index, static_cast<std::size_t>(s->part_nnz) * sizeof(int32_t),
                      hipMemcpyDeviceToDevice, st) != hipSuccess))
    return sbk::capi_fail(SPARLAB_E_DEVICE, "copying the block's pattern failed");
  return SPARLAB_OK;
}

int sparlab_system_free_block_device(sparlab_system* s, const double* values, double* block_values, void* stream) {
  if (!s || !values || (s->part_nnz > 0 && !block_values)) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (s->part_free < 0) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "no partition");
  if (sbk::launch_free_block_values(values, s->part_source, s->part_nnz, block_values,
                                    static_cast<hipStream_t>(stream)) != 0)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "free block launch failed");
  return SPARLAB_OK;
}

namespace {

// How one problem's block solve ended.
enum SolveEnd : int32_t { kSolved = 0, kNotDefinite = 1, kBrokeDown = 2, kNotConverged = 3 };

// The block solve of a.nseg problems (a plain solve: one) on stream st, each
// problem through the same steps as it would take alone: runs of iterations,
// the true residual once the iteration's residual meets the tolerance, a
// restart from it when it does not (residual replacement), BiCGStab's
// restarts after a breakdown. Every launch acts on the problems that need it
// (the active flags; a plain solve has none and needs every launch). Per
// problem: iterations, relative residual and how it ended. False when a
// device call failed.
bool run_block_solve(sbk::CgArgs& a, int symmetric, const double* b, double* x, double tolerance,
                     int64_t max_iterations, hipStream_t st, int64_t* iterations, double* relative_residual,
                     int32_t* end) {
  // iterations between two reads of the residual: 16 with the diagonal, 4
  // with the multigrid cycle (an iteration costs a V-cycle or two, a read a
  // synchronisation; a solve then overshoots its tolerance by fewer cycles)
  const int64_t kRun = a.external ? 4 : 16;
  const int nseg = a.nseg;
  std::vector<double> sc(static_cast<std::size_t>(nseg) * sbk::kCgScalars);
  auto read = [&]() {
    return hipMemcpyAsync(sc.data(), a.scalars, sc.size() * sizeof(double), hipMemcpyDeviceToHost, st) ==
               hipSuccess &&
           hipStreamSynchronize(st) == hipSuccess;
  };
  auto at = [&](int p, int k) { return sc[static_cast<std::size_t>(p) * sbk::kCgScalars + k]; };
  std::vector<int> on(nseg, 1);  // the device's active flags (all set by the layout)
  auto select = [&](const std::vector<int>& want) {
    if (!a.active || want == on) return true;
    on = want;
    return hipMemcpyAsync(a.active, on.data(), on.size() * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess &&
           hipStreamSynchronize(st) == hipSuccess;
  };
  std::vector<int64_t> cycle(nseg, 0);  // each problem's first iteration since its last (re)start
  auto upload_cycles = [&]() {
    return !a.cycles || (hipMemcpyAsync(a.cycles, cycle.data(), cycle.size() * sizeof(int64_t),
                                        hipMemcpyHostToDevice, st) == hipSuccess &&
                         hipStreamSynchronize(st) == hipSuccess);
  };
  const int started = symmetric ? sbk::launch_cg_start(a, b, x, st) : sbk::launch_bicg_start(a, b, x, st);
  if (started != 0 || !read()) return false;
  std::vector<double> bb(nseg), goal(nseg);
  std::vector<char> converged(nseg), stopped(nseg);
  std::vector<int> restarts(nseg, 0);
  for (int p = 0; p < nseg; ++p) {
    bb[p] = at(p, sbk::kCgBb);
    goal[p] = tolerance * tolerance * bb[p];
    converged[p] = at(p, sbk::kCgRr) <= goal[p];  // r = b at the start: exact
    stopped[p] = converged[p] || at(p, sbk::kCgBreak) != 0.0;
    iterations[p] = 0;
    relative_residual[p] = 0.0;
  }
  int64_t done = 0;
  // The iteration's residual drifts from b - K x; when it meets the
  // tolerance the true residual is computed, and the iteration restarts from
  // it (residual replacement) unless that meets the tolerance too.
  // BiCGStab's shadow residual can become orthogonal to the residual in
  // floating point (r^ . r = 0: a breakdown); the iteration then restarts from
  // the true residual with r^ = r, at most kRestarts times. A break of
  // conjugate gradients (p . K p <= 0: round-off near the solution, or an
  // indefinite block) left every step after it zero, so x is the last good
  // iterate: the solution if its true residual meets the tolerance.
  constexpr int kRestarts = 16;
  enum Next : char { kNone, kBreakCg, kBreakBicg, kMet };
  std::vector<char> next(nseg);
  std::vector<int> run_set(nseg), sub(nseg);
  while (done < max_iterations) {
    bool any = false;
    for (int p = 0; p < nseg; ++p) any |= (run_set[p] = !stopped[p]) != 0;
    if (!any) break;
    if (!select(run_set)) return false;
    const int run = static_cast<int>(max_iterations - done < kRun ? max_iterations - done : kRun);
    const int ran = symmetric ? sbk::launch_cg_iterations(a, x, done, run, st)
                              : sbk::launch_bicg_iterations(a, x, cycle[0], done, run, st);
    if (ran != 0 || !read()) return false;
    done += run;
    bool residual = false;
    for (int p = 0; p < nseg; ++p) {
      next[p] = kNone;
      if (!run_set[p]) continue;
      iterations[p] = done;
      const bool broke = at(p, sbk::kCgBreak) != 0.0;
      if (broke && symmetric)
        next[p] = kBreakCg;
      else if (broke && restarts[p] < kRestarts)
        next[p] = kBreakBicg;
      else if (broke)
        stopped[p] = 1;  // BiCGStab out of restarts
      else if (at(p, sbk::kCgRr) <= goal[p])
        next[p] = kMet;
      residual |= next[p] != kNone;
    }
    if (!residual) continue;
    for (int p = 0; p < nseg; ++p) sub[p] = next[p] != kNone;
    if (!select(sub) || sbk::launch_cg_residual(a, b, x, st) != 0 || !read()) return false;
    bool restart = false;
    for (int p = 0; p < nseg; ++p) {
      sub[p] = 0;
      if (next[p] == kNone) continue;
      converged[p] = at(p, sbk::kCgRr) <= goal[p];
      if (converged[p] || next[p] == kBreakCg) {
        stopped[p] = 1;
        continue;
      }
      sub[p] = 1;
      cycle[p] = done;
      if (next[p] == kBreakBicg) ++restarts[p];
      restart = true;
    }
    if (!restart) continue;
    if (!select(sub) || (!symmetric && !upload_cycles())) return false;
    const int restarted = symmetric ? sbk::launch_cg_restart(a, done, st) : sbk::launch_bicg_restart(a, done, st);
    if (restarted != 0 || !read()) return false;
  }
  // the ends; a solve out of iterations gets its true residual
  bool residual = false;
  for (int p = 0; p < nseg; ++p) {
    sub[p] = 0;
    if (converged[p]) {
      end[p] = kSolved;
    } else if (at(p, sbk::kCgBreak) != 0.0) {
      end[p] = symmetric ? kNotDefinite : kBrokeDown;
    } else {
      end[p] = kNotConverged;
      sub[p] = 1;
      residual = true;
    }
  }
  if (residual && (!select(sub) || sbk::launch_cg_residual(a, b, x, st) != 0 || !read())) return false;
  for (int p = 0; p < nseg; ++p) {
    if (end[p] == kNotDefinite || end[p] == kBrokeDown) continue;
    // after a converged run the scalars hold the true residual's r . r
    const double rr = at(p, sbk::kCgRr);
    relative_residual[p] = bb[p] > 0.0 ? std::sqrt(rr / bb[p]) : std::sqrt(rr);
  }
  return true;
}

// The partition's free DOFs (ascending) and the node coordinates, from the
// device.
bool read_partition(sparlab_system* s, std::vector<long long>& free_dofs, std::vector<double>& coords,
                    hipStream_t st) {
  const std::size_t ndof = 3 * static_cast<std::size_t>(s->nn);
  std::vector<int32_t> map(ndof);
  coords.resize(ndof);
  if (hipMemcpyAsync(map.data(), s->part_map, ndof * sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(coords.data(), s->coords, ndof * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return false;
  free_dofs.clear();
  free_dofs.reserve(static_cast<std::size_t>(s->part_free));
  for (std::size_t d = 0; d < ndof; ++d)
    if (map[d] >= 0) free_dofs.push_back(static_cast<long long>(d));
  return true;
}

// Set up a batch for the block's values; the first singular problem, or -1.
int set_up_multigrid(sbk::AmgBatch& batch, const double* block_values, int symmetric, hipStream_t st, int& bad) {
  bad = -1;
  if (batch.setup(block_values, !symmetric, st) != 0) return sbk::capi_fail(SPARLAB_E_DEVICE, "the multigrid set-up failed");
  const std::vector<int> sing = batch.singular(st);
  for (std::size_t p = 0; p < sing.size(); ++p)
    if (sing[p]) {
      bad = static_cast<int>(p);
      break;
    }
  return SPARLAB_OK;
}

// The partition's multigrid hierarchy (built at the first solve after the
// partition or new coordinates) set up for the block's values.
int prepare_multigrid(sparlab_system* s, const double* block_values, int symmetric, hipStream_t st) {
  if (!s->amg) {
    std::vector<long long> free_dofs;
    std::vector<double> coords;
    if (!read_partition(s, free_dofs, coords, st)) return sbk::capi_fail(SPARLAB_E_DEVICE, "reading the partition failed");
    auto h = std::make_unique<sbk::AmgHierarchy>();
    const std::string why = h->build(s->part_ptr, s->part_index, s->part_free, s->part_nnz, free_dofs, coords, st);
    if (!why.empty()) return sbk::capi_fail(SPARLAB_E_SOLVER, "the multigrid hierarchy could not be built: " + why);
    auto* b = new sbk::AmgBatch();
    b->add(std::move(h), 0, 0);
    if (b->finalize() != 0) {
      delete b;
      return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation failed");
    }
    s->amg = b;
  }
  int bad = -1;
  const int rc = set_up_multigrid(*s->amg, block_values, symmetric, st, bad);
  if (rc != SPARLAB_OK) return rc;
  if (bad >= 0) return sbk::capi_fail(SPARLAB_E_SOLVER, "the multigrid hierarchy's coarsest level is singular");
  return SPARLAB_OK;
}

// One hierarchy per segment of the partition's block (each built from its
// rows alone, as a problem's own block), set up for the block's values.
int prepare_segment_multigrid(sparlab_system* s, const std::vector<long long>& rows, const double* block_values,
                              int symmetric, hipStream_t st) {
  if (!s->amg_seg || s->amg_seg_rows != rows) {
    delete s->amg_seg;
    s->amg_seg = nullptr;
    s->amg_seg_rows.clear();
    std::vector<long long> free_dofs;
    std::vector<double> coords;
    std::vector<int32_t> ptr(static_cast<std::size_t>(s->part_free) + 1), index(static_cast<std::size_t>(s->part_nnz));
    if (!read_partition(s, free_dofs, coords, st) ||
        hipMemcpyAsync(ptr.data(), s->part_ptr, ptr.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        (!index.empty() && hipMemcpyAsync(index.data(), s->part_index, index.size() * sizeof(int32_t),
                                          hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess)
      return sbk::capi_fail(SPARLAB_E_DEVICE, "reading the partition failed");
    // the segments' hierarchies are independent: built by a pool of host
    // threads, each with a stream of its own on the caller's device
    const int nseg = static_cast<int>(rows.size()) - 1;
    std::vector<std::unique_ptr<sbk::AmgHierarchy>> built(static_cast<std::size_t>(nseg));
    std::vector<std::string> why(static_cast<std::size_t>(nseg));
    std::atomic<int> next{0};
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return sbk::capi_fail(SPARLAB_E_DEVICE, "no current device");
    auto work = [&]() {
      hipStream_t ws = nullptr;
      const bool ok = hipSetDevice(device) == hipSuccess &&
                      hipStreamCreateWithFlags(&ws, hipStreamNonBlocking) == hipSuccess;
      for (int p; (p = next++) < nseg;) {
        if (!ok) {
          why[static_cast<std::size_t>(p)] = "no stream for the build";
          continue;
        }
        const long long r0 = rows[p], r1 = rows[p + 1];
        const int v0 = ptr[static_cast<std::size_t>(r0)], v1 = ptr[static_cast<std::size_t>(r1)];
        try {  // nothing may escape a worker thread
          // the segment's rows alone: pointers from 0, columns from r0 (the
          // segment check has shown that none leaves the segment)
          std::vector<int> sp(static_cast<std::size_t>(r1 - r0) + 1), si(static_cast<std::size_t>(v1 - v0));
          for (long long i = r0; i <= r1; ++i)
            sp[static_cast<std::size_t>(i - r0)] = ptr[static_cast<std::size_t>(i)] - v0;
          for (int k = v0; k < v1; ++k)
            si[static_cast<std::size_t>(k - v0)] = index[static_cast<std::size_t>(k)] - static_cast<int>(r0);
          std::vector<long long> sf(free_dofs.begin() + r0, free_dofs.begin() + r1);
          auto h = std::make_unique<sbk::AmgHierarchy>();
          why[static_cast<std::size_t>(p)] = h->build_host(std::move(sp), std::move(si), sf, coords, ws);
          built[static_cast<std::size_t>(p)] = std::move(h);
        } catch (const std::exception& ex) {
          why[static_cast<std::size_t>(p)] = std::string("host failure: ") + ex.what();
        }
      }
      if (ws) (void)hipStreamDestroy(ws);
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const int nthreads = std::max(1, std::min({nseg, 16, hw > 0 ? static_cast<int>(hw) : 1}));
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) {
      try {
        pool.emplace_back(work);
      } catch (const std::exception&) {
        break;  // fewer threads: the others take the remaining segments
      }
    }
    work();
    for (auto& t : pool) t.join();
    auto* b = new sbk::AmgBatch();
    for (int p = 0; p < nseg; ++p) {
      if (!why[static_cast<std::size_t>(p)].empty()) {
        delete b;
        return sbk::capi_fail(SPARLAB_E_SOLVER, "segment " + std::to_string(p) +
                                                    ": the multigrid hierarchy could not be built: " +
                                                    why[static_cast<std::size_t>(p)]);
      }
      b->add(std::move(built[static_cast<std::size_t>(p)]), rows[p], ptr[static_cast<std::size_t>(rows[p])]);
    }
    if (b->finalize() != 0) {
      delete b;
      return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation failed");
    }
    s->amg_seg = b;
    s->amg_seg_rows = rows;
  }
  int bad = -1;
  const int rc = set_up_multigrid(*s->amg_seg, block_values, symmetric, st, bad);
  if (rc != SPARLAB_OK) return rc;
  if (bad >= 0)
    return sbk::capi_fail(SPARLAB_E_SOLVER,
                          "segment " + std::to_string(bad) + ": the multigrid hierarchy's coarsest level is singular");
  return SPARLAB_OK;
}

const char* solve_end_message(int32_t end, int symmetric) {
  if (end == kNotDefinite) return "the free-free block is not positive definite (conjugate gradients)";
  if (end == kBrokeDown) return "the block solve broke down (BiCGStab)";
  (void)symmetric;
  return "the block solve did not converge";
}

}  // namespace

int sparlab_system_solve_free_block_device(sparlab_system* s, const double* block_values, int symmetric,
                                           const double* b, double* x, double tolerance, int64_t max_iterations,
                                           int64_t* iterations, double* relative_residual, void* stream) {
  if (!s || !iterations || !relative_residual) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (s->part_free < 0) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "no partition");
  const int64_t n = s->part_free;
  if ((s->part_nnz > 0 && !block_values) || (n > 0 && (!b || !x)))
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (n > 0 && b < x + n && x < b + n) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "b and x overlap");
  if (!(tolerance > 0.0) || !std::isfinite(tolerance) || max_iterations < 0)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "bad tolerance or iteration limit");
  hipStream_t st = static_cast<hipStream_t>(stream);
  sbk::CgArgs a{};
  a.ptr = s->part_ptr;
  a.index = s->part_index;
  a.values = block_values;
  a.n = n;
  sbk::cg_layout(a, s->part_work);
  if (s->preconditioner == SPARLAB_PRECONDITIONER_MULTIGRID) {
    int prepared;
    try {
      prepared = prepare_multigrid(s, block_values, symmetric, st);
    } catch (const std::exception& ex) {
      return sbk::capi_fail(SPARLAB_E_DEVICE, std::string("the multigrid hierarchy: host failure: ") + ex.what());
    }
    if (prepared != SPARLAB_OK) return prepared;
    a.external = 1;
    a.precond = sbk::amg_apply;
    s->amg->set_active(nullptr);
    a.precond_ctx = s->amg;
  }
  int32_t end = kSolved;
  if (!run_block_solve(a, symmetric, b, x, tolerance, max_iterations, st, iterations, relative_residual, &end))
    return sbk::capi_fail(SPARLAB_E_DEVICE, "the block solve's launches failed");
  if (end != kSolved) return sbk::capi_fail(SPARLAB_E_SOLVER, solve_end_message(end, symmetric));
  return SPARLAB_OK;
}

int sparlab_system_set_block_preconditioner(sparlab_system* s, int kind) {
  if (!s) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (kind != SPARLAB_PRECONDITIONER_DIAGONAL && kind != SPARLAB_PRECONDITIONER_MULTIGRID)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "unknown preconditioner");
  s->preconditioner = kind;
  return SPARLAB_OK;
}

int sparlab_system_block_preconditioner_levels(sparlab_system* s, int32_t* count, int64_t* rows,
                                               int64_t* nonzeros) {
  if (!s || !count) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  *count = 0;
  if (!s->amg) return SPARLAB_OK;
  const std::vector<long long> r = s->amg->level_rows(), z = s->amg->level_nnz();
  *count = static_cast<int32_t>(r.size());
  for (std::size_t l = 0; l < r.size() && l < SPARLAB_MULTIGRID_MAX_LEVELS; ++l) {
    if (rows) rows[l] = r[l];
    if (nonzeros) nonzeros[l] = z[l];
  }
  return SPARLAB_OK;
}

int sparlab_system_solve_free_block_segments_device(sparlab_system* s, const double* block_values, int symmetric,
                                                    int32_t num_segments, const int64_t* segment_rows,
                                                    const double* b, double* x, double tolerance,
                                                    int64_t max_iterations, int64_t* iterations,
                                                    double* relative_residual, int32_t* status, void* stream) {
  if (!s || !segment_rows || !iterations || !relative_residual || !status)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (s->part_free < 0) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "no partition");
  const int64_t n = s->part_free;
  if ((s->part_nnz > 0 && !block_values) || (n > 0 && (!b || !x)))
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (n > 0 && b < x + n && x < b + n) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "b and x overlap");
  if (!(tolerance > 0.0) || !std::isfinite(tolerance) || max_iterations < 0)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "bad tolerance or iteration limit");
  if (num_segments < 1) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "at least one segment");
  std::vector<long long> rows(segment_rows, segment_rows + num_segments + 1);
  if (rows.front() != 0 || rows.back() != n)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the segments must cover the block's rows from 0 to num_free");
  for (int p = 0; p < num_segments; ++p)
    if (rows[p + 1] < rows[p]) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the segment rows must ascend");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const sbk::CgSegments g = sbk::cg_segments(rows);
  const std::size_t work = static_cast<std::size_t>(sbk::cg_segment_work_doubles(g));
  const std::size_t meta = static_cast<std::size_t>(sbk::cg_segment_meta_bytes(g)) + sizeof(int);
  if (work > s->seg_work_doubles) {
    dfree(s->seg_work);
    s->seg_work_doubles = 0;
    if (!dalloc(s->seg_work, work)) return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation failed");
    s->seg_work_doubles = work;
  }
  if (meta > s->seg_meta_bytes) {
    dfree(s->seg_meta);
    s->seg_meta_bytes = 0;
    if (!dalloc(s->seg_meta, meta)) return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation failed");
    s->seg_meta_bytes = meta;
  }
  sbk::CgArgs a{};
  a.ptr = s->part_ptr;
  a.index = s->part_index;
  a.values = block_values;
  if (sbk::cg_segment_layout(a, g, s->seg_work, s->seg_meta, st) != 0)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "copying the segments failed");
  // the block must be block-diagonal over the segments (its pattern's check:
  // a flag after the metadata)
  int* flag = reinterpret_cast<int*>(static_cast<char*>(static_cast<void*>(s->seg_meta)) + meta - sizeof(int));
  int bad = 0;
  if (hipMemsetAsync(flag, 0, sizeof(int), st) != hipSuccess || sbk::launch_segment_check(a, flag, st) != 0 ||
      hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "the segments' check failed");
  if (bad) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the free-free block couples two segments");
  if (s->preconditioner == SPARLAB_PRECONDITIONER_MULTIGRID) {
    int prepared;
    try {
      prepared = prepare_segment_multigrid(s, rows, block_values, symmetric, st);
    } catch (const std::exception& ex) {
      return sbk::capi_fail(SPARLAB_E_DEVICE, std::string("the multigrid hierarchies: host failure: ") + ex.what());
    }
    if (prepared != SPARLAB_OK) return prepared;
    a.external = 1;
    a.precond = sbk::amg_apply;
    s->amg_seg->set_active(a.active);  // a resting problem's V-cycles do nothing
    a.precond_ctx = s->amg_seg;
  }
  if (!run_block_solve(a, symmetric, b, x, tolerance, max_iterations, st, iterations, relative_residual, status))
    return sbk::capi_fail(SPARLAB_E_DEVICE, "the block solve's launches failed");
  for (int p = 0; p < num_segments; ++p)
    if (status[p] != kSolved)
      return sbk::capi_fail(SPARLAB_E_SOLVER, "segment " + std::to_string(p) + ": " +
                                                  solve_end_message(status[p], symmetric));
  return SPARLAB_OK;
}

int sparlab_system_stress_device(sparlab_system* s, const double* u, double* stress, int32_t* status,
                                 void* stream) {
  if (!s || !u || !stress) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  sbk::ElementArgs a = s->args;
  a.coords = s->coords;
  a.conn = s->conn;
  a.u = u;
  a.ne = s->ne;
  a.state_in = s->state[s->cur];
  a.state_stride = s->ne * s->points;
  a.alpha_in = s->alpha[s->cur];
  a.status = status;
  a.fail_value = nullptr;
  if (sbk::launch_element_stress(a, stress, static_cast<hipStream_t>(stream)) != 0)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "stress kernel launch failed");
  return SPARLAB_OK;
}

int sparlab_system_result(sparlab_system* s, sparlab_evaluation* out, void* stream) {
  if (!s || !out) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (!s->evaluated) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "no evaluation to read");
  hipStream_t st = static_cast<hipStream_t>(stream);
  double sc[2];
  long long sum[3];
  if (hipMemcpyAsync(sc, s->scalars, sizeof(sc), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(sum, s->summary, sizeof(sum), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "reading the evaluation failed");
  out->energy = sc[0];
  out->gross = sc[1];
  out->failed_elements = sum[0];
  out->yielding_points = sum[2];
  out->first_failed_element = sum[0] > 0 ? sum[1] : -1;
  out->first_failure_status = 0;
  out->first_failure_value = 0.0;
  if (sum[0] > 0) {
    int32_t code = 0;
    double v = 0.0;
    if (hipMemcpy(&code, s->status + sum[1], sizeof(code), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&v, s->fail_value + sum[1], sizeof(v), hipMemcpyDeviceToHost) != hipSuccess)
      return sbk::capi_fail(SPARLAB_E_DEVICE, "reading the first failure failed");
    out->first_failure_status = code;
    out->first_failure_value = v;
  }
  return sum[0] > 0 ? SPARLAB_E_SOLVER : SPARLAB_OK;
}

int sparlab_system_evaluate(sparlab_system* s, const double* u_host, int want_tangent, int column_major,
                            double* fint_host, double* values_host, sparlab_evaluation* out) {
  if (!s || !u_host || !fint_host || !out) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  const std::size_t nd = 3 * static_cast<std::size_t>(s->nn);
  const std::size_t nnz = s->pattern.col_idx.size();
  double *du = nullptr, *df = nullptr, *dv = nullptr;
  if (!dalloc(du, nd) || !dalloc(df, nd) || (want_tangent && !dalloc(dv, nnz))) {
    dfree(du); dfree(df); dfree(dv);
    return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation failed");
  }
  int rc = SPARLAB_OK;
  if (hipMemcpy(du, u_host, nd * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    rc = sbk::capi_fail(SPARLAB_E_DEVICE, "upload of u failed");
  } else {
    rc = sparlab_system_evaluate_device(s, du, want_tangent, column_major, df, dv, nullptr);
    if (rc == SPARLAB_OK) {
      rc = sparlab_system_result(s, out, nullptr);
      if (hipMemcpy(fint_host, df, nd * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
          (want_tangent && values_host &&
           hipMemcpy(values_host, dv, nnz * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
        rc = sbk::capi_fail(SPARLAB_E_DEVICE, "download failed");
      else if (rc == SPARLAB_E_SOLVER)
        sbk::capi_fail(SPARLAB_E_SOLVER, "Hex8 element " + std::to_string(out->first_failed_element) +
                                             " failed with status " + std::to_string(out->first_failure_status));
    }
  }
  dfree(du); dfree(df); dfree(dv);
  return rc;
}

int sparlab_system_commit(sparlab_system* s) {
  if (!s || !s->evaluated) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "commit needs an evaluation");
  s->cur = 1 - s->cur;
  s->evaluated = false;
  s->adjoint_ready = false;
  return SPARLAB_OK;
}

int sparlab_system_history(const sparlab_system* s, int trial, double* states, double* alpha) {
  if (!s) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null system");
  const int k = trial ? 1 - s->cur : s->cur;
  const std::size_t hist = static_cast<std::size_t>(s->fields) * s->ne * s->points;
  if (hipDeviceSynchronize() != hipSuccess) return sbk::capi_fail(SPARLAB_E_DEVICE, "synchronize failed");
  if ((states && hipMemcpy(states, s->state[k], hist * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
      (alpha && hipMemcpy(alpha, s->alpha[k], 9 * s->ne * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
    return sbk::capi_fail(SPARLAB_E_DEVICE, "download failed");
  return SPARLAB_OK;
}

int sparlab_system_set_history(sparlab_system* s, const double* states, const double* alpha) {
  if (!s) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null system");
  const std::size_t hist = static_cast<std::size_t>(s->fields) * s->ne * s->points;
  if ((states && hipMemcpy(s->state[s->cur], states, hist * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) ||
      (alpha && hipMemcpy(s->alpha[s->cur], alpha, 9 * s->ne * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
    return sbk::capi_fail(SPARLAB_E_DEVICE, "upload failed");
  s->adjoint_ready = false;
  return SPARLAB_OK;
}

int sparlab_system_set_coords(sparlab_system* s, const double* coords) {
  if (!s || !coords) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (hipMemcpy(s->coords, coords, 3 * static_cast<std::size_t>(s->nn) * sizeof(double), hipMemcpyHostToDevice) !=
      hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "upload failed");
  s->adjoint_ready = false;
  release_multigrid(s);  // their rigid-body modes were the old coordinates'
  return SPARLAB_OK;
}

int sparlab_system_set_material(sparlab_system* s, const sparlab_material* material) {
  if (!s || !material) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  sbk::ElementArgs a = s->args;
  const int st = sbk::capi_prepare_elements(material, &s->rule, s->kinematics, a);
  if (st != SPARLAB_OK) return st;
  if (sbk::state_fields(a.mat.kin_fields) != s->fields)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the new material needs another history layout (kinematic terms)");
  a.chunk = s->args.chunk;
  s->args = a;
  s->material = *material;
  s->adjoint_ready = false;
  return SPARLAB_OK;
}

int sparlab_system_vjp_device(sparlab_system* s, const double* u, const double* v, int wrt, double* grad_coords,
                              double* grad_disp, double* grad_params, void* stream) {
  if (!s || !u || !v) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (!s->adjoint_ready)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the adjoint needs the last evaluation (before commit or new inputs)");
  if (((wrt & SPARLAB_VJP_COORDS) && !grad_coords) || ((wrt & SPARLAB_VJP_DISPLACEMENT) && !grad_disp) ||
      ((wrt & SPARLAB_VJP_PARAMETERS) && !grad_params))
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "a requested gradient has no buffer");
  const std::size_t ne = static_cast<std::size_t>(s->ne);
  if ((!s->grad_x && !dalloc(s->grad_x, 24 * ne)) || (!s->grad_u && !dalloc(s->grad_u, 24 * ne)) ||
      (!s->grad_p && !dalloc(s->grad_p, static_cast<std::size_t>(sbk::kAdjointParams) * ne)))
    return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation for the adjoint failed");
  hipStream_t st = static_cast<hipStream_t>(stream);
  sbk::AdjointArgs q{};
  sbk::ElementArgs& a = q.e;
  a = s->args;
  a.coords = s->coords;
  a.conn = s->conn;
  a.u = u;
  a.ne = s->ne;
  a.state_in = s->state[s->cur];
  a.state_stride = s->ne * s->points;
  a.alpha_in = s->alpha[s->cur];
  a.alpha_out = s->alpha[1 - s->cur];
  a.state_out = s->state[1 - s->cur];
  a.status = s->status;
  sbk::set_workspace(a, s->ws);
  q.v = v;
  q.wrt = wrt;
  q.hill = sbk::hill_r_derivatives(s->material, q.dP) ? 1 : 0;
  q.backstresses = s->material.num_backstresses;
  q.grad_x = s->grad_x;
  q.grad_u = s->grad_u;
  q.grad_p = s->grad_p;
  if (sbk::launch_adjoint(q, st) != 0) return sbk::capi_fail(SPARLAB_E_DEVICE, "adjoint launch failed");
  if (wrt & SPARLAB_VJP_COORDS) {
    sbk::ForceArgs f{s->grad_x, s->node_ptr, s->node_incidence, static_cast<long long>(s->nn), grad_coords, s->gross};
    if (sbk::launch_gather_force(f, st) != 0) return sbk::capi_fail(SPARLAB_E_DEVICE, "gather launch failed");
  }
  if (wrt & SPARLAB_VJP_DISPLACEMENT) {
    sbk::ForceArgs f{s->grad_u, s->node_ptr, s->node_incidence, static_cast<long long>(s->nn), grad_disp, s->gross};
    if (sbk::launch_gather_force(f, st) != 0) return sbk::capi_fail(SPARLAB_E_DEVICE, "gather launch failed");
  }
  if (wrt & SPARLAB_VJP_PARAMETERS) {
    const int parts = static_cast<int>(std::min<int64_t>(1024, s->ne / 4096 + 1));
    for (int k = 0; k < sbk::kAdjointParams; ++k)
      if (sbk::launch_reduce(s->grad_p + static_cast<std::size_t>(k) * ne, s->ne, 0, s->partial, parts,
                             grad_params + k, st) != 0)
        return sbk::capi_fail(SPARLAB_E_DEVICE, "reduction launch failed");
  }
  return SPARLAB_OK;
}

int sparlab_system_element_outputs(const sparlab_system* s, double* fe, double* ke) {
  if (!s) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null system");
  if (hipDeviceSynchronize() != hipSuccess) return sbk::capi_fail(SPARLAB_E_DEVICE, "synchronize failed");
  if ((fe && hipMemcpy(fe, s->fe, 24 * s->ne * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
      (ke && hipMemcpy(ke, s->ke, 576 * s->ne * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
    return sbk::capi_fail(SPARLAB_E_DEVICE, "download failed");
  if (ke && s->ke_column_major) transpose_tangents(ke, s->ne);
  return SPARLAB_OK;
}

int sparlab_system_info_get(const sparlab_system* s, sparlab_system_info* out) {
  if (!s || !out) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  *out = sparlab_system_info{};
  out->num_nodes = s->nn;
  out->num_elements = s->ne;
  out->nnz = static_cast<int64_t>(s->pattern.col_idx.size());
  out->points = s->points;
  out->state_fields = s->fields;
  out->kinematic_fields = s->args.mat.kin_fields;
  out->general = s->args.mat.general;
  // NonlinearSystem::initialise_history (src/fem/NonlinearSystem.hpp:273-311):
  // one material, so every element is elastoplastic or none is; a recovering
  // backstress makes the tangent non-symmetric; the incompatible-mode element
  // averages no dilatation.
  out->plastic = s->args.mat.enabled;
  out->elastoplastic = s->args.mat.enabled;
  out->symmetric = s->args.mat.symmetric;
  out->averaged = 0;
  out->modes = 9;
  out->kinematics = s->kinematics;
  return SPARLAB_OK;
}

int sparlab_system_max_plastic_strain(sparlab_system* s, double* out, void* stream) {
  if (!s || !out) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  *out = 0.0;
  if (!s->args.mat.enabled || s->ne == 0) return SPARLAB_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long npts = static_cast<long long>(s->ne) * s->points;
  const double* abar = s->state[s->cur] + static_cast<std::size_t>(6 + 6 * s->args.mat.kin_fields) * npts;
  const int parts = static_cast<int>(std::min<long long>(1024, npts / 4096 + 1));
  if (sbk::launch_reduce_max(abar, npts, s->partial, parts, s->scalars + 2, st) != 0 ||
      hipMemcpyAsync(out, s->scalars + 2, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "the plastic strain maximum failed");
  return SPARLAB_OK;
}

int sparlab_system_history_range(const sparlab_system* s, int trial, int64_t first, int64_t count, double* states,
                                 double* alpha) {
  if (!s) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null system");
  if (first < 0 || count < 0 || first > s->ne || count > s->ne - first)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the element range lies outside the mesh");
  if (count == 0) return SPARLAB_OK;
  const int k = trial ? 1 - s->cur : s->cur;
  const std::size_t npts = static_cast<std::size_t>(s->ne) * s->points;
  const std::size_t width = static_cast<std::size_t>(count) * s->points;
  if (hipDeviceSynchronize() != hipSuccess) return sbk::capi_fail(SPARLAB_E_DEVICE, "synchronize failed");
  if (states) {
    // field f of the range's points is one contiguous run of the field's row
    std::vector<double> rows(static_cast<std::size_t>(s->fields) * width);
    if (hipMemcpy2D(rows.data(), width * sizeof(double), s->state[k] + static_cast<std::size_t>(first) * s->points,
                    npts * sizeof(double), width * sizeof(double), static_cast<std::size_t>(s->fields),
                    hipMemcpyDeviceToHost) != hipSuccess)
      return sbk::capi_fail(SPARLAB_E_DEVICE, "download failed");
    for (std::size_t q = 0; q < width; ++q)
      for (int f = 0; f < s->fields; ++f) states[q * s->fields + f] = rows[static_cast<std::size_t>(f) * width + q];
  }
  if (alpha && hipMemcpy(alpha, s->alpha[k] + 9 * first, 9 * static_cast<std::size_t>(count) * sizeof(double),
                         hipMemcpyDeviceToHost) != hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "download failed");
  return SPARLAB_OK;
}

int sparlab_system_element_outputs_range(const sparlab_system* s, int64_t first, int64_t count, double* fe,
                                         double* ke) {
  if (!s) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null system");
  if (first < 0 || count < 0 || first > s->ne || count > s->ne - first)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the element range lies outside the mesh");
  if (count == 0) return SPARLAB_OK;
  if (hipDeviceSynchronize() != hipSuccess) return sbk::capi_fail(SPARLAB_E_DEVICE, "synchronize failed");
  if ((fe && hipMemcpy(fe, s->fe + 24 * first, 24 * count * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
      (ke && hipMemcpy(ke, s->ke + 576 * first, 576 * count * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
    return sbk::capi_fail(SPARLAB_E_DEVICE, "download failed");
  if (ke && s->ke_column_major) transpose_tangents(ke, count);
  return SPARLAB_OK;
}

}  // extern "C"
