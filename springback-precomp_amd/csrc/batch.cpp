// The batched form: many independent forming problems advanced together.
//
// Problems share no nodes, so the union of their meshes is one block-diagonal
// system: its pattern is every problem's pattern (Assembler::build_pattern,
// src/fem/Assembler.cpp:86-173) with rows and columns shifted by the
// problem's first DOF, and problem p's CSR values are one contiguous slice of
// the union's values. One evaluation of the union is therefore one launch of
// each element, gather and assembly kernel over every problem
// (NonlinearSystem::evaluate, src/fem/NonlinearSystem.hpp:192-268, for all of
// them), plus one per-problem reduction of the scalars. The launch sequence is
// captured once into a HIP graph per (buffers, tangent, value order,
// committed side) and replayed every Newton iteration. The local solve's
// elements per wavefront are picked by timing the candidates on the batch.
#include "sparlab_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "batch_kernels.hpp"
#include "system_impl.hpp"

struct sparlab_batch {
  int num_problems = 0;
  std::vector<int64_t> node_offset, element_offset, value_offset;
  sparlab_system* system = nullptr;
  long long* d_node_offset = nullptr;
  long long* d_element_offset = nullptr;
  double *energy = nullptr, *gross = nullptr;
  long long *yielding = nullptr, *failed = nullptr, *first_failed = nullptr;
  hipStream_t capture = nullptr;
  struct Graph {
    const double* u;
    double* fint;
    double* values;
    int tangent, column_major, cur;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  bool use_graphs = true;
};

namespace {

void drop_graphs(sparlab_batch* b) {
  for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
  b->graphs.clear();
}

void release(sparlab_batch* b) {
  drop_graphs(b);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  if (b->system) sparlab_system_destroy(b->system);
  for (void* p : {static_cast<void*>(b->d_node_offset), static_cast<void*>(b->d_element_offset),
                  static_cast<void*>(b->energy), static_cast<void*>(b->gross), static_cast<void*>(b->yielding),
                  static_cast<void*>(b->failed), static_cast<void*>(b->first_failed)})
    if (p) (void)hipFree(p);
}

// The launches of one batched evaluation.
int enqueue(sparlab_batch* b, const double* u, int want_tangent, int column_major, double* fint, double* values,
            hipStream_t st) {
  const int rc = sbk::system_enqueue(b->system, u, want_tangent, column_major, fint, values, st);
  if (rc != SPARLAB_OK) return rc;
  sbk::ProblemScalarArgs a{b->d_node_offset, b->d_element_offset, b->system->gross, b->system->energy,
                           b->system->status, b->system->yielding, b->energy, b->gross, b->yielding, b->failed,
                           b->first_failed};
  if (sbk::launch_problem_scalars(a, b->num_problems, st) != 0)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "per-problem reduction launch failed");
  return SPARLAB_OK;
}

}  // namespace

extern "C" {

int sparlab_batch_create(int32_t num_problems, const int64_t* node_offset, const int64_t* element_offset,
                         const double* coords, const int32_t* conn, const sparlab_material* material,
                         const sparlab_rule* rule, int kinematics, sparlab_batch** out) {
  if (out == nullptr || node_offset == nullptr || element_offset == nullptr || coords == nullptr ||
      conn == nullptr || material == nullptr || rule == nullptr)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  *out = nullptr;
  if (num_problems <= 0) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "a batch needs at least one problem");
  if (node_offset[0] != 0 || element_offset[0] != 0)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the offsets must start at 0");
  for (int p = 0; p < num_problems; ++p)
    if (node_offset[p + 1] < node_offset[p] || element_offset[p + 1] < element_offset[p])
      return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the offsets must not decrease");
  const int64_t nn = node_offset[num_problems], ne = element_offset[num_problems];
  // problem-local node indices -> union indices; a node outside its problem is a mesh error
  std::vector<int32_t> uconn(static_cast<std::size_t>(8 * ne));
  for (int p = 0; p < num_problems; ++p) {
    const int64_t n0 = node_offset[p], n1 = node_offset[p + 1];
    for (int64_t e = element_offset[p]; e < element_offset[p + 1]; ++e)
      for (int a = 0; a < 8; ++a) {
        const int64_t local = conn[8 * e + a];
        if (local < 0 || n0 + local >= n1)
          return sbk::capi_fail(SPARLAB_E_MESH, "problem " + std::to_string(p) + ", element " +
                                                    std::to_string(e - element_offset[p]) +
                                                    ": node index outside the problem");
        uconn[static_cast<std::size_t>(8 * e + a)] = static_cast<int32_t>(n0 + local);
      }
  }
  auto* b = new sparlab_batch();
  b->num_problems = num_problems;
  b->node_offset.assign(node_offset, node_offset + num_problems + 1);
  b->element_offset.assign(element_offset, element_offset + num_problems + 1);
  int st = sparlab_system_create(coords, nn, uconn.data(), ne, material, rule, kinematics, &b->system);
  if (st != SPARLAB_OK) {
    delete b;
    return st;
  }
  // the first value of each problem's rows: rows of problem p are [3 n0, 3 n1)
  b->value_offset.resize(static_cast<std::size_t>(num_problems) + 1);
  for (int p = 0; p <= num_problems; ++p)
    b->value_offset[static_cast<std::size_t>(p)] = b->system->pattern.row_ptr[static_cast<std::size_t>(3 * node_offset[p])];
  const std::size_t np = static_cast<std::size_t>(num_problems);
  std::vector<long long> no(node_offset, node_offset + num_problems + 1),
      eo(element_offset, element_offset + num_problems + 1);
  bool ok = hipMalloc(reinterpret_cast<void**>(&b->d_node_offset), (np + 1) * sizeof(long long)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&b->d_element_offset), (np + 1) * sizeof(long long)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&b->energy), np * sizeof(double)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&b->gross), np * sizeof(double)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&b->yielding), np * sizeof(long long)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&b->failed), np * sizeof(long long)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&b->first_failed), np * sizeof(long long)) == hipSuccess &&
            hipMemcpy(b->d_node_offset, no.data(), (np + 1) * sizeof(long long), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(b->d_element_offset, eo.data(), (np + 1) * sizeof(long long), hipMemcpyHostToDevice) ==
                hipSuccess &&
            hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) == hipSuccess;
  if (!ok) {
    release(b);
    delete b;
    return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation for the batch failed");
  }
  *out = b;
  return SPARLAB_OK;
}

void sparlab_batch_destroy(sparlab_batch* b) {
  if (!b) return;
  release(b);
  delete b;
}

sparlab_system* sparlab_batch_system(sparlab_batch* b) { return b ? b->system : nullptr; }

int sparlab_batch_value_offsets(const sparlab_batch* b, int64_t* value_offset) {
  if (!b || !value_offset) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  std::copy(b->value_offset.begin(), b->value_offset.end(), value_offset);
  return SPARLAB_OK;
}

int sparlab_batch_set_graphs(sparlab_batch* b, int enabled) {
  if (!b) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  b->use_graphs = enabled != 0;
  if (!b->use_graphs) drop_graphs(b);
  return SPARLAB_OK;
}

int sparlab_batch_evaluate_device(sparlab_batch* b, const double* u, int want_tangent, int column_major,
                                  double* fint, double* values, void* stream) {
  if (!b || !u || !fint) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (want_tangent && !values) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "want_tangent needs a values buffer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!b->use_graphs) return enqueue(b, u, want_tangent, column_major, fint, values, st);
  const int tangent = want_tangent != 0, cm = column_major != 0, cur = b->system->cur;
  double* vals = tangent ? values : nullptr;
  for (auto& g : b->graphs)
    if (g.u == u && g.fint == fint && g.values == vals && g.tangent == tangent && g.column_major == cm &&
        g.cur == cur) {
      if (hipGraphLaunch(g.exec, st) != hipSuccess)
        return sbk::capi_fail(SPARLAB_E_DEVICE, "launching the batch's graph failed");
      b->system->evaluated = true;
      b->system->last_tangent = tangent != 0;
      return SPARLAB_OK;
    }
  // capture the launch sequence once for these buffers and this committed side
  hipGraph_t graph = nullptr;
  if (hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "graph capture could not start");
  const int rc = enqueue(b, u, tangent, cm, fint, values, b->capture);
  const hipError_t end = hipStreamEndCapture(b->capture, &graph);
  if (rc != SPARLAB_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (end != hipSuccess || graph == nullptr) return sbk::capi_fail(SPARLAB_E_DEVICE, "graph capture failed");
  hipGraphExec_t exec = nullptr;
  const hipError_t inst = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (inst != hipSuccess) return sbk::capi_fail(SPARLAB_E_DEVICE, "graph instantiation failed");
  b->graphs.push_back({u, fint, vals, tangent, cm, cur, exec});
  if (hipGraphLaunch(exec, st) != hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "launching the batch's graph failed");
  b->system->evaluated = true;
  b->system->last_tangent = tangent != 0;
  return SPARLAB_OK;
}

int sparlab_batch_result(sparlab_batch* b, double* energy, double* gross, int64_t* yielding, int64_t* failed,
                         int64_t* first_failed, void* stream) {
  if (!b) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (!b->system->evaluated) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "no evaluation to read");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const std::size_t np = static_cast<std::size_t>(b->num_problems);
  std::vector<long long> y(np), f(np), ff(np);
  std::vector<double> en(np), gr(np);
  if (hipMemcpyAsync(en.data(), b->energy, np * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(gr.data(), b->gross, np * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(y.data(), b->yielding, np * sizeof(long long), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(f.data(), b->failed, np * sizeof(long long), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(ff.data(), b->first_failed, np * sizeof(long long), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "reading the batch's scalars failed");
  int64_t failures = 0;
  for (std::size_t p = 0; p < np; ++p) {
    if (energy) energy[p] = en[p];
    if (gross) gross[p] = gr[p];
    if (yielding) yielding[p] = y[p];
    if (failed) failed[p] = f[p];
    if (first_failed) first_failed[p] = ff[p];
    failures += f[p];
  }
  return failures > 0 ? SPARLAB_E_SOLVER : SPARLAB_OK;
}

int sparlab_batch_autotune(sparlab_batch* b, const double* u, int* chosen, void* stream) {
  if (!b || !u) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  sparlab_system* s = b->system;
  sbk::ElementArgs a = s->args;
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
  a.tangent = s->ke;
  a.energy = s->energy;
  a.yielding = s->yielding;
  a.iterations = s->iterations;
  a.status = s->status;
  a.fail_value = s->fail_value;
  sbk::set_workspace(a, s->ws);
  const int candidates[] = {8, 16, 32, 64, 128, 256};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "event creation failed");
  int best = s->args.chunk > 0 ? s->args.chunk : 64;
  float best_ms = 0.0f;
  bool have = false, failed = false;
  for (int c : candidates) {
    a.chunk = c;
    if (sbk::launch_local_solve(a, st) != 0) { failed = true; break; }  // warm-up
    if (hipEventRecord(t0, st) != hipSuccess) { failed = true; break; }
    for (int r = 0; r < 2; ++r)
      if (sbk::launch_local_solve(a, st) != 0) { failed = true; break; }
    if (failed || hipEventRecord(t1, st) != hipSuccess || hipEventSynchronize(t1) != hipSuccess) {
      failed = true;
      break;
    }
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, t0, t1) != hipSuccess) { failed = true; break; }
    if (!have || ms < best_ms) { best_ms = ms; best = c; have = true; }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (failed) return sbk::capi_fail(SPARLAB_E_DEVICE, "timing the local solve failed");
  if (s->args.chunk != best) {
    s->args.chunk = best;
    drop_graphs(b);  // the captured launches carry the old value
  }
  if (chosen) *chosen = best;
  return SPARLAB_OK;
}

}  // extern "C"
