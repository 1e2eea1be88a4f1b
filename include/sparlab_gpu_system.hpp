// sparlab_gpu_system.hpp - the device nonlinear system of sparlab_gpu.h in the
// shape of the reference's detail::NonlinearSystem
// (src/fem/NonlinearSystem.hpp:85-334): the members the forming driver calls
// (src/fem/Forming.cpp:693-714, 770-783, 1005, 1046, 1202, 1211, 1261,
// 1274-1277, 1339-1340), on host std::vector data, with the reference's
// PlasticState record and its exceptions. Header-only C++17 over the C ABI;
// a maintainer maps the vectors to Eigen (Vector, the SparseMatrix's values)
// in the thin shim that sits in the reference.
//
// Device memory: the system keeps device copies of u, f_int and the tangent's
// values, and downloads what each member returns. Every member runs on the
// system's stream and returns when its results are on the host.
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sparlab_gpu.h"

namespace sparlab_gpu {

// The reference's exception types (include/sparlab/core/Exceptions.hpp).
struct SolverError : std::runtime_error {
  explicit SolverError(const std::string& m) : std::runtime_error(m) {}
};
struct MeshError : std::runtime_error {
  explicit MeshError(const std::string& m) : std::runtime_error(m) {}
};
struct ConfigError : std::runtime_error {
  explicit ConfigError(const std::string& m) : std::runtime_error(m) {}
};

// A C-ABI return code as the reference's exception.
inline void check(int code) {
  if (code >= 0) return;
  const std::string m = sparlab_last_error();
  switch (code) {
    case SPARLAB_E_SOLVER: throw SolverError(m);
    case SPARLAB_E_MESH: throw MeshError(m);
    case SPARLAB_E_ARGUMENT: throw ConfigError(m);
    default: throw std::runtime_error("sparlab_gpu: " + m);
  }
}

// PlasticState (include/sparlab/material/Plasticity.hpp:185-201), field for
// field: plastic strain (engineering shears), the whole back stress beta,
// the general return's backstresses alpha_i (Prager's last), the equivalent
// plastic strain, the plane-stress thickness strain, the loading flag.
struct PlasticState {
  std::array<double, 6> plastic_strain{};
  std::array<double, 6> back_stress{};
  std::array<std::array<double, 6>, 4> back_stresses{};
  double equivalent_plastic_strain = 0.0;
  double thickness_strain = 0.0;
  bool loading = false;
};

// One point's library fields (sparlab_state_fields of them) -> PlasticState:
// the general return keeps alpha_i (beta is their sum), the radial return
// beta alone.
inline PlasticState to_plastic_state(const double* f, const sparlab_system_info& info) {
  PlasticState p;
  for (int i = 0; i < 6; ++i) p.plastic_strain[i] = f[i];
  if (info.general) {
    for (int t = 0; t < info.kinematic_fields; ++t)
      for (int i = 0; i < 6; ++i) {
        p.back_stresses[t][i] = f[6 + 6 * t + i];
        p.back_stress[i] += f[6 + 6 * t + i];
      }
  } else {
    for (int i = 0; i < 6; ++i) p.back_stress[i] = f[6 + i];
  }
  p.equivalent_plastic_strain = f[6 + 6 * info.kinematic_fields];
  p.loading = f[7 + 6 * info.kinematic_fields] != 0.0;
  return p;
}

inline void from_plastic_state(const PlasticState& p, const sparlab_system_info& info, double* f) {
  for (int i = 0; i < 6; ++i) f[i] = p.plastic_strain[i];
  if (info.general) {
    for (int t = 0; t < info.kinematic_fields; ++t)
      for (int i = 0; i < 6; ++i) f[6 + 6 * t + i] = p.back_stresses[t][i];
  } else {
    for (int i = 0; i < 6; ++i) f[6 + i] = p.back_stress[i];
  }
  f[6 + 6 * info.kinematic_fields] = p.equivalent_plastic_strain;
  f[7 + 6 * info.kinematic_fields] = p.loading ? 1.0 : 0.0;
}

// Evaluation (src/fem/NonlinearSystem.hpp:27-47) of a forming model (no
// applied load: f_ext = q = 0, residual = internal). The trial states and
// modes stay on the device until System::trial_states / commit.
struct Evaluation {
  std::vector<double> internal;        // f_int, 3 per node [N]
  std::vector<double> tangent_values;  // K's values in the pattern's column-major order (empty without)
  double energy = 0.0;
  double gross = 0.0;
  int64_t yielding_points = 0;
};

class System {
 public:
  // A Hex8 mesh (coords 3 per node, conn 8 per element, VTK order), its
  // material (finalized here), integration rule and kinematics; the history
  // starts virgin (NonlinearSystem.hpp:95-101, 273-311).
  System(const std::vector<double>& coords, const std::vector<int32_t>& conn, sparlab_material material,
         sparlab_rule rule, int kinematics, hipStream_t stream = nullptr)
      : stream_(stream) {
    if (coords.size() % 3 != 0 || conn.size() % 8 != 0) throw ConfigError("coords or conn has a partial record");
    check(sparlab_material_finalize(&material));
    check(sparlab_system_create(coords.data(), static_cast<int64_t>(coords.size() / 3), conn.data(),
                                static_cast<int64_t>(conn.size() / 8), &material, &rule, kinematics, &h_));
    check(sparlab_system_info_get(h_, &info_));
    outer_.resize(static_cast<std::size_t>(3 * info_.num_nodes + 1));
    inner_.resize(static_cast<std::size_t>(info_.nnz));
    check(sparlab_system_pattern(h_, outer_.data(), inner_.data()));
    const std::size_t n = static_cast<std::size_t>(3 * info_.num_nodes);
    alloc(&d_u_, n);
    alloc(&d_fint_, n);
    alloc(&d_values_, static_cast<std::size_t>(info_.nnz));
  }
  ~System() {
    (void)hipFree(d_u_);
    (void)hipFree(d_fint_);
    (void)hipFree(d_values_);
    if (h_) sparlab_system_destroy(h_);
  }
  System(const System&) = delete;
  System& operator=(const System&) = delete;

  const sparlab_system_info& info() const { return info_; }
  sparlab_system* handle() const { return h_; }
  // The pattern: the column-major outer and inner arrays of the tangent
  // (structurally symmetric, so also its CSR row pointers and columns).
  const std::vector<int32_t>& outer() const { return outer_; }
  const std::vector<int32_t>& inner() const { return inner_; }

  bool symmetric() const { return info_.symmetric != 0; }
  bool plastic() const { return info_.plastic != 0; }
  bool averaged(int64_t e) const {
    element(e);
    return info_.averaged != 0;
  }
  bool elastoplastic(int64_t e) const {
    element(e);
    return info_.elastoplastic != 0;
  }

  // evaluate(u, lambda, want_tangent) (NonlinearSystem.hpp:192-268): f_int,
  // the tangent's column-major values, energy, gross and yielding points;
  // the trial history is kept for commit. A failed element throws the
  // reference's SolverError (or MeshError) for the lowest failing element.
  Evaluation evaluate(const std::vector<double>& u, double /*lambda*/, bool want_tangent) const {
    const std::size_t n = static_cast<std::size_t>(3 * info_.num_nodes);
    if (u.size() != n) throw ConfigError("u has " + std::to_string(u.size()) + " entries; the model has " +
                                         std::to_string(n) + " DOFs");
    device(hipMemcpyAsync(d_u_, u.data(), n * sizeof(double), hipMemcpyHostToDevice, stream_));
    check(sparlab_system_evaluate_device(h_, d_u_, want_tangent ? 1 : 0, 1, d_fint_, want_tangent ? d_values_ : nullptr,
                                         stream_));
    sparlab_evaluation ev{};
    check(sparlab_system_result(h_, &ev, stream_));  // throws for a failed element
    Evaluation out;
    out.internal.resize(n);
    device(hipMemcpyAsync(out.internal.data(), d_fint_, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (want_tangent) {
      out.tangent_values.resize(static_cast<std::size_t>(info_.nnz));
      device(hipMemcpyAsync(out.tangent_values.data(), d_values_, out.tangent_values.size() * sizeof(double),
                            hipMemcpyDeviceToHost, stream_));
    }
    device(hipStreamSynchronize(stream_));
    out.energy = ev.energy;
    out.gross = ev.gross;
    out.yielding_points = ev.yielding_points;
    return out;
  }

  // commit(Evaluation&) (NonlinearSystem.hpp:119-129): the last evaluation's
  // states and modes become the committed ones.
  void commit() { check(sparlab_system_commit(h_)); }

  // committed(e) / internal(e) (NonlinearSystem.hpp:108-137).
  std::vector<PlasticState> committed(int64_t e) const { return states(0, element(e), 1)[0]; }
  std::array<double, 9> internal(int64_t e) const { return modes(0, element(e), 1)[0]; }
  // Evaluation::states / modes of the last evaluation, elements first ..
  // first + count - 1.
  std::vector<std::vector<PlasticState>> trial_states(int64_t first, int64_t count) const {
    return states(1, first, count);
  }
  std::vector<std::array<double, 9>> trial_modes(int64_t first, int64_t count) const { return modes(1, first, count); }

  // committed_all / set_committed, internal_all / set_internal
  // (NonlinearSystem.hpp:138-179): the history a restart carries; the sizes
  // are checked as the reference checks them (ConfigError).
  std::vector<std::vector<PlasticState>> committed_all() const { return states(0, 0, info_.num_elements); }
  void set_committed(const std::vector<std::vector<PlasticState>>& all) {
    if (static_cast<int64_t>(all.size()) != info_.num_elements)
      throw ConfigError("the saved plastic history has " + std::to_string(all.size()) + " element(s); the model has " +
                        std::to_string(info_.num_elements));
    const std::size_t nf = static_cast<std::size_t>(info_.state_fields);
    const std::size_t npts = static_cast<std::size_t>(info_.num_elements) * info_.points;
    std::vector<double> soa(nf * npts), point(nf);
    for (std::size_t e = 0; e < all.size(); ++e) {
      if (static_cast<int>(all[e].size()) != info_.points)
        throw ConfigError("the saved plastic history of element " + std::to_string(e) + " has " +
                          std::to_string(all[e].size()) + " point(s); the model keeps " + std::to_string(info_.points));
      for (int q = 0; q < info_.points; ++q) {
        from_plastic_state(all[e][static_cast<std::size_t>(q)], info_, point.data());
        for (std::size_t f = 0; f < nf; ++f) soa[f * npts + e * info_.points + q] = point[f];
      }
    }
    check(sparlab_system_set_history(h_, soa.data(), nullptr));
  }
  std::vector<std::array<double, 9>> internal_all() const { return modes(0, 0, info_.num_elements); }
  void set_internal(const std::vector<std::array<double, 9>>& all) {
    if (static_cast<int64_t>(all.size()) != info_.num_elements)
      throw ConfigError("the saved internal-mode history has " + std::to_string(all.size()) +
                        " element(s); the model keeps " + std::to_string(info_.num_elements));
    std::vector<double> flat(9 * all.size());
    for (std::size_t e = 0; e < all.size(); ++e)
      for (int i = 0; i < 9; ++i) flat[9 * e + i] = all[e][static_cast<std::size_t>(i)];
    check(sparlab_system_set_history(h_, nullptr, flat.data()));
  }

  // max_plastic_strain() (NonlinearSystem.hpp:181-188).
  double max_plastic_strain() const {
    double v = 0.0;
    check(sparlab_system_max_plastic_strain(h_, &v, stream_));
    return v;
  }

  // elastoplastic_stress of every element at u from committed(e),
  // averaged(e) and internal(e) (src/fem/Forming.cpp:1263-1281):
  // SPARLAB_STRESS_FIELDS per element; a failed element throws.
  std::vector<double> stress(const std::vector<double>& u) const {
    const std::size_t n = static_cast<std::size_t>(3 * info_.num_nodes);
    if (u.size() != n) throw ConfigError("u has the wrong size");
    const std::size_t ne = static_cast<std::size_t>(info_.num_elements);
    double* d_stress = nullptr;
    int32_t* d_status = nullptr;
    alloc(&d_stress, ne * SPARLAB_STRESS_FIELDS);
    if (hipMalloc(reinterpret_cast<void**>(&d_status), ne * sizeof(int32_t)) != hipSuccess) {
      (void)hipFree(d_stress);
      throw std::bad_alloc();
    }
    std::vector<double> out(ne * SPARLAB_STRESS_FIELDS);
    std::vector<int32_t> status(ne);
    int rc = hipMemcpyAsync(d_u_, u.data(), n * sizeof(double), hipMemcpyHostToDevice, stream_) == hipSuccess
                 ? sparlab_system_stress_device(h_, d_u_, d_stress, d_status, stream_)
                 : SPARLAB_E_DEVICE;
    if (rc == SPARLAB_OK &&
        (hipMemcpyAsync(out.data(), d_stress, out.size() * sizeof(double), hipMemcpyDeviceToHost, stream_) !=
             hipSuccess ||
         hipMemcpyAsync(status.data(), d_status, ne * sizeof(int32_t), hipMemcpyDeviceToHost, stream_) != hipSuccess ||
         hipStreamSynchronize(stream_) != hipSuccess))
      rc = SPARLAB_E_DEVICE;
    (void)hipFree(d_stress);
    (void)hipFree(d_status);
    check(rc);
    for (std::size_t e = 0; e < ne; ++e) {
      if (status[e] == SPARLAB_ELEMENT_OK) continue;
      const std::string where = "Hex8 element " + std::to_string(e);
      if (status[e] <= SPARLAB_ELEMENT_POINT_JACOBIAN) throw MeshError(where + ": det J <= 0");
      throw SolverError(where + ": the stress recovery failed (status " + std::to_string(status[e]) + ")");
    }
    return out;
  }

 private:
  int64_t element(int64_t e) const {
    if (e < 0 || e >= info_.num_elements) throw std::out_of_range("element " + std::to_string(e));
    return e;
  }
  static void device(hipError_t err) {
    if (err != hipSuccess) throw std::runtime_error(std::string("sparlab_gpu: ") + hipGetErrorString(err));
  }
  static void alloc(double** p, std::size_t n) {
    if (hipMalloc(reinterpret_cast<void**>(p), (n > 0 ? n : 1) * sizeof(double)) != hipSuccess) throw std::bad_alloc();
  }
  std::vector<std::vector<PlasticState>> states(int trial, int64_t first, int64_t count) const {
    const std::size_t nf = static_cast<std::size_t>(info_.state_fields);
    std::vector<double> flat(static_cast<std::size_t>(count) * info_.points * nf);
    check(sparlab_system_history_range(h_, trial, first, count, flat.data(), nullptr));
    std::vector<std::vector<PlasticState>> out(static_cast<std::size_t>(count));
    for (std::size_t e = 0; e < out.size(); ++e) {
      out[e].resize(static_cast<std::size_t>(info_.points));
      for (int q = 0; q < info_.points; ++q)
        out[e][static_cast<std::size_t>(q)] = to_plastic_state(flat.data() + (e * info_.points + q) * nf, info_);
    }
    return out;
  }
  std::vector<std::array<double, 9>> modes(int trial, int64_t first, int64_t count) const {
    std::vector<double> flat(9 * static_cast<std::size_t>(count));
    check(sparlab_system_history_range(h_, trial, first, count, nullptr, flat.data()));
    std::vector<std::array<double, 9>> out(static_cast<std::size_t>(count));
    for (std::size_t e = 0; e < out.size(); ++e)
      for (int i = 0; i < 9; ++i) out[e][static_cast<std::size_t>(i)] = flat[9 * e + i];
    return out;
  }

  sparlab_system* h_ = nullptr;
  sparlab_system_info info_{};
  std::vector<int32_t> outer_, inner_;
  hipStream_t stream_ = nullptr;
  double* d_u_ = nullptr;
  double* d_fint_ = nullptr;
  double* d_values_ = nullptr;
};

}  // namespace sparlab_gpu
