// Oracle (test infrastructure only): a C ABI over the restatement, for the
// tests (ctypes), the smoke check and the CPU-baseline leg of the benchmark.
//
// State record: 39 doubles per integration point, the fields of
// PlasticState (include/sparlab/material/Plasticity.hpp:185-201):
//   [0, 6) plastic strain (engineering shears), [6, 12) back stress beta,
//   [12, 36) backstresses alpha_1..alpha_4 (6 each), 36 equivalent plastic
//   strain, 37 thickness strain, 38 loading flag (0/1).
// Material record: 54 doubles (see read_material).
// Status codes: 0 ok, 1 SolverError, 2 MeshError, 3 ConfigError, 4 other.
#include "assembly.hpp"
#include "element.hpp"
#include "surface.hpp"

#include <algorithm>
#include <cstring>
#include <omp.h>

using namespace oracle;

namespace {

constexpr int kStateRecord = 39;

MaterialData read_material(const double* m) {
  MaterialData d;
  d.E = m[0];
  d.nu = m[1];
  d.sy0 = m[2];
  d.H = m[3];
  d.Hkin = m[4];
  d.Q = m[5];
  d.delta = m[6];
  d.criterion = static_cast<int>(m[7]);
  d.nback = static_cast<int>(m[8]);
  for (int i = 0; i < 4; ++i) { d.C[i] = m[9 + i]; d.gamma[i] = m[13 + i]; }
  d.integration = static_cast<int>(m[17]);
  for (int i = 0; i < 36; ++i) d.P[i] = m[18 + i];
  return d;
}

template <class T> PlasticState<T> read_state(const double* s) {
  PlasticState<T> st;
  for (int i = 0; i < 6; ++i) { st.ep(i) = s[i]; st.beta(i) = s[6 + i]; }
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 6; ++i) st.alpha[k](i) = s[12 + 6 * k + i];
  st.abar = s[36];
  st.thickness_strain = s[37];
  st.loading = s[38] != 0.0;
  return st;
}

template <class T> void write_state(const PlasticState<T>& st, double* s) {
  for (int i = 0; i < 6; ++i) { s[i] = val(st.ep(i)); s[6 + i] = val(st.beta(i)); }
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 6; ++i) s[12 + 6 * k + i] = val(st.alpha[k](i));
  s[36] = val(st.abar);
  s[37] = val(st.thickness_strain);
  s[38] = st.loading ? 1.0 : 0.0;
}

void put_message(const std::exception& ex, char* msg, int len) {
  if (msg == nullptr || len <= 0) return;
  std::strncpy(msg, ex.what(), static_cast<std::size_t>(len - 1));
  msg[len - 1] = '\0';
}

template <class F> int guarded(char* msg, int len, F&& f) {
  try {
    f();
    return 0;
  } catch (const SolverError& ex) {
    put_message(ex, msg, len);
    return 1;
  } catch (const MeshError& ex) {
    put_message(ex, msg, len);
    return 2;
  } catch (const ConfigError& ex) {
    put_message(ex, msg, len);
    return 3;
  } catch (const std::exception& ex) {
    put_message(ex, msg, len);
    return 4;
  }
}

// One element evaluation (IncompatibleModes.cpp:577-600) on scalar type T.
// scalars_out: energy, yielding points, local iterations, symmetric, passes, path.
template <class T>
int element_impl(const double* x0, const double* ue, const double* states_in, const double* alpha_in,
                 const MaterialData& mat, const Rule& rule, int kinematics, int want_tangent,
                 double* f, double* k, double* alpha_out, double* states_out, double* scalars,
                 char* msg, int msglen) {
  return guarded(msg, msglen, [&] {
    IncompatibleHex8<T> el(mat, rule, kinematics);
    T uet[24];
    for (int i = 0; i < 24; ++i) uet[i] = T(ue[i]);
    el.set_up(x0, uet, 0);
    std::vector<PlasticState<T>> committed(static_cast<std::size_t>(rule.n));
    for (int q = 0; q < rule.n; ++q) committed[static_cast<std::size_t>(q)] = read_state<T>(states_in + q * kStateRecord);
    T at[9];
    if (alpha_in != nullptr)
      for (int i = 0; i < 9; ++i) at[i] = T(alpha_in[i]);
    const ElementResult<T> r = el.evaluate(committed, alpha_in != nullptr ? at : nullptr, want_tangent != 0);
    for (int i = 0; i < 24; ++i) f[i] = val(r.force(i));
    if (k != nullptr)
      for (int i = 0; i < 576; ++i) k[i] = val(r.tangent.a[i]);
    if (alpha_out != nullptr)
      for (int i = 0; i < 9; ++i) alpha_out[i] = val(r.internal(i));
    if (states_out != nullptr)
      for (int q = 0; q < rule.n; ++q) write_state(r.states[static_cast<std::size_t>(q)], states_out + q * kStateRecord);
    if (scalars != nullptr) {
      scalars[0] = val(r.energy);
      scalars[1] = r.yielding_points;
      scalars[2] = r.internal_iterations;
      scalars[3] = r.symmetric ? 1.0 : 0.0;
      scalars[4] = r.passes;
      scalars[5] = r.path;
    }
  });
}

}  // namespace

extern "C" {

int oracle_state_record() { return kStateRecord; }

// Tests only: plants one of the mistakes of the Plant list (0: none).
void oracle_set_plant(int which) { g_plant = which; }

// The hardening laws at a (Material.cpp:73-109): out = yield stress, its
// slope and the isotropic hardening energy.
void oracle_hardening(const double* mat, double a, double* out) {
  const MaterialData m = read_material(mat);
  out[0] = yield(m, a);
  out[1] = yield_slope(m, a);
  out[2] = isotropic_energy(m, a);
}

int oracle_rule_points(int in_plane, int thickness, int axis) {
  try {
    return make_rule(in_plane, thickness, axis).n;
  } catch (...) {
    return -1;
  }
}

int oracle_rule(int in_plane, int thickness, int axis, double* xi, double* eta, double* zeta, double* w) {
  try {
    const Rule r = make_rule(in_plane, thickness, axis);
    for (int q = 0; q < r.n; ++q) { xi[q] = r.xi[q]; eta[q] = r.eta[q]; zeta[q] = r.zeta[q]; w[q] = r.w[q]; }
    return r.n;
  } catch (...) {
    return -1;
  }
}

// Hill48 yield matrix (Material.cpp:135-217); coeffs_out: F, G, H, L, M, N.
int oracle_hill_yield_matrix(int calibration, const double* values /*r0 r45 r90 s45 s90 sb L M F G H N*/,
                             const double* rd, const double* nd, double* p, double* coeffs_out,
                             char* msg, int msglen) {
  return guarded(msg, msglen, [&] {
    HillInput h;
    h.calibration = calibration;
    h.r0 = values[0]; h.r45 = values[1]; h.r90 = values[2];
    h.sigma45 = values[3]; h.sigma90 = values[4]; h.sigma_b = values[5];
    h.shear_l = values[6]; h.shear_m = values[7];
    h.F = values[8]; h.G = values[9]; h.Hh = values[10]; h.N = values[11];
    h.L = values[6]; h.M = values[7];
    for (int i = 0; i < 3; ++i) { h.rd[i] = rd[i]; h.nd[i] = nd[i]; }
    derive_hill(h, p);
    if (coeffs_out != nullptr) {
      coeffs_out[0] = h.F; coeffs_out[1] = h.G; coeffs_out[2] = h.Hh;
      coeffs_out[3] = h.L; coeffs_out[4] = h.M; coeffs_out[5] = h.N;
    }
  });
}

void oracle_von_mises_yield_matrix(double* p) { von_mises_yield_matrix(p); }

// plastic_return (Plasticity.cpp:513-539), three-dimensional.
// scalars_out: energy, yielding, symmetric.
int oracle_plastic_return(const double* mat, const double* strain, const double* state_in, int want_tangent,
                          double* stress, double* tangent, double* state_out, double* scalars,
                          char* msg, int msglen) {
  const MaterialData m = read_material(mat);
  return guarded(msg, msglen, [&] {
    Vec<double, 6> e;
    for (int i = 0; i < 6; ++i) e(i) = strain[i];
    const PlasticResponse<double> r = plastic_return(m, e, read_state<double>(state_in), want_tangent != 0);
    for (int i = 0; i < 6; ++i) stress[i] = r.stress(i);
    for (int i = 0; i < 36; ++i) tangent[i] = r.tangent.a[i];
    write_state(r.state, state_out);
    scalars[0] = r.energy;
    scalars[1] = r.yielding ? 1.0 : 0.0;
    scalars[2] = r.symmetric ? 1.0 : 0.0;
  });
}

// logarithmic_strain / logarithmic_stress / logarithmic_curvature
// (LogarithmicStrain.cpp:109-197). t (6) may be null.
int oracle_log_strain(const double* green, const double* t, double* axes, double* mu, double* strain,
                      double* projection, double* first, double* s_out, double* curvature,
                      char* msg, int msglen) {
  return guarded(msg, msglen, [&] {
    Vec<double, 6> g;
    for (int i = 0; i < 6; ++i) g(i) = green[i];
    const LogStrain<double> ls = logarithmic_strain(g);
    for (int i = 0; i < 9; ++i) { axes[i] = ls.axes.a[i]; first[i] = ls.first.a[i]; }
    for (int i = 0; i < 3; ++i) mu[i] = ls.green(i);
    for (int i = 0; i < 6; ++i) strain[i] = ls.strain(i);
    for (int i = 0; i < 36; ++i) projection[i] = ls.projection.a[i];
    if (t != nullptr) {
      Vec<double, 6> tv;
      for (int i = 0; i < 6; ++i) tv(i) = t[i];
      const Vec<double, 6> s = logarithmic_stress(ls, tv);
      for (int i = 0; i < 6; ++i) s_out[i] = s(i);
      const Mat<double, 6, 6> c = logarithmic_curvature(ls, tv);
      for (int i = 0; i < 36; ++i) curvature[i] = c.a[i];
    }
  });
}

// The element (IncompatibleModes.cpp:577-600). x0: 8 x 3 node-major; ue 24;
// states_in: points x 39; alpha_in: 9 or null; scalars_out: 5.
int oracle_element(const double* x0, const double* ue, const double* states_in, const double* alpha_in,
                   const double* mat, int in_plane, int thickness, int axis, int kinematics,
                   int want_tangent, double* f, double* k, double* alpha_out, double* states_out,
                   double* scalars, char* msg, int msglen) {
  Rule rule;
  const int st = guarded(msg, msglen, [&] { rule = make_rule(in_plane, thickness, axis); });
  if (st != 0) return st;
  return element_impl<double>(x0, ue, states_in, alpha_in, read_material(mat), rule, kinematics,
                              want_tangent, f, k, alpha_out, states_out, scalars, msg, msglen);
}

// The same element on the op-counting scalar: returns the FP64 operations of
// one evaluation (negative on a failure).
long long oracle_element_flops(const double* x0, const double* ue, const double* states_in,
                               const double* alpha_in, const double* mat, int in_plane, int thickness,
                               int axis, int kinematics, int want_tangent) {
  const Rule rule = make_rule(in_plane, thickness, axis);
  std::vector<double> f(24), k(576), a(9), so(static_cast<std::size_t>(rule.n) * kStateRecord);
  double sc[6];
  flop_counter() = 0;
  const int st = element_impl<Counted>(x0, ue, states_in, alpha_in, read_material(mat), rule, kinematics,
                                       want_tangent, f.data(), k.data(), a.data(), so.data(), sc, nullptr, 0);
  const long long n = static_cast<long long>(flop_counter());
  return st == 0 ? n : -1;
}

// The same count split at the end of the local iteration: local[0] = the
// FLOPs up to convergence of the modes, local[1] = the coupling blocks and
// the condensation. Returns the total or -1.
long long oracle_element_flops_split(const double* x0, const double* ue, const double* states_in,
                                     const double* alpha_in, const double* mat, int in_plane, int thickness,
                                     int axis, int kinematics, int want_tangent, long long* parts) {
  const Rule rule = make_rule(in_plane, thickness, axis);
  std::vector<double> f(24), k(576), a(9), so(static_cast<std::size_t>(rule.n) * kStateRecord);
  double sc[6];
  flop_counter() = 0;
  flop_mark() = 0;
  const int st = element_impl<Counted>(x0, ue, states_in, alpha_in, read_material(mat), rule, kinematics,
                                       want_tangent, f.data(), k.data(), a.data(), so.data(), sc, nullptr, 0);
  const long long n = static_cast<long long>(flop_counter());
  parts[0] = static_cast<long long>(flop_mark());
  parts[1] = n - parts[0];
  return st == 0 ? n : -1;
}

// A batch of independent elements, OpenMP schedule(static) over elements as
// NonlinearSystem.hpp:208-237 (the CPU baseline). Element e reads
// x0[e*24], ue[e*24], states_in[e*P*39], alpha_in[e*9] (alpha_in may be null)
// and writes f[e*24], k[e*576] (k may be null), alpha_out[e*9],
// states_out[e*P*39], scalars[e*6], status[e]. Returns the failures.
int oracle_element_batch(long long ne, const double* x0, const double* ue, const double* states_in,
                         const double* alpha_in, const double* mat, int in_plane, int thickness, int axis,
                         int kinematics, int want_tangent, int threads, double* f, double* k,
                         double* alpha_out, double* states_out, double* scalars, int* status) {
  const MaterialData m = read_material(mat);
  const Rule rule = make_rule(in_plane, thickness, axis);
  const long long per = static_cast<long long>(rule.n) * kStateRecord;
  int failures = 0;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static) reduction(+ : failures)
  for (long long e = 0; e < ne; ++e) {
    const int st = element_impl<double>(
        x0 + e * 24, ue + e * 24, states_in + e * per, alpha_in ? alpha_in + e * 9 : nullptr, m, rule,
        kinematics, want_tangent, f + e * 24, k ? k + e * 576 : nullptr, alpha_out + e * 9,
        states_out + e * per, scalars + e * 6, nullptr, 0);
    status[e] = st;
    if (st != 0) ++failures;
  }
  return failures;
}

// incompatible_elastoplastic_stress (IncompatibleModes.cpp:602-664) for a
// batch of elements at their committed states and modes (alpha_in may be
// null: zero), OpenMP over elements. out[e*28]: cauchy 6, piola_kirchhoff 6,
// kirchhoff 6, logarithmic_strain 6 (tensorial Voigt), von_mises,
// max_equivalent_plastic_strain, max_strain, min_jacobian; status[e] as
// oracle_element. Returns the failures.
int oracle_element_stress_batch(long long ne, const double* x0, const double* ue, const double* states_in,
                                const double* alpha_in, const double* mat, int in_plane, int thickness, int axis,
                                int kinematics, int threads, double* out, int* status) {
  const MaterialData m = read_material(mat);
  const Rule rule = make_rule(in_plane, thickness, axis);
  const long long per = static_cast<long long>(rule.n) * kStateRecord;
  int failures = 0;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static) reduction(+ : failures)
  for (long long e = 0; e < ne; ++e) {
    double* o = out + e * 28;
    const int st = guarded(nullptr, 0, [&] {
      IncompatibleHex8<double> el(m, rule, kinematics);
      el.set_up(x0 + e * 24, ue + e * 24, e);
      std::vector<PlasticState<double>> committed(static_cast<std::size_t>(rule.n));
      for (int q = 0; q < rule.n; ++q)
        committed[static_cast<std::size_t>(q)] = read_state<double>(states_in + e * per + q * kStateRecord);
      const ElementStress<double> r = el.stress(committed, alpha_in ? alpha_in + e * 9 : nullptr);
      for (int i = 0; i < 6; ++i) {
        o[i] = r.cauchy(i);
        o[6 + i] = r.piola_kirchhoff(i);
        o[12 + i] = r.kirchhoff(i);
        o[18 + i] = r.logarithmic_strain(i);
      }
      o[24] = r.von_mises;
      o[25] = r.max_equivalent_plastic_strain;
      o[26] = r.max_strain;
      o[27] = r.min_jacobian;
    });
    if (st != 0)
      for (int i = 0; i < 28; ++i) o[i] = 0.0;
    status[e] = st;
    if (st != 0) ++failures;
  }
  return failures;
}

// The standard Hex8 force at the same kinematics (modes held at zero).
int oracle_standard_force(const double* x0, const double* ue, const double* states_in, const double* mat,
                          int in_plane, int thickness, int axis, int kinematics, double* f, char* msg,
                          int msglen) {
  return guarded(msg, msglen, [&] {
    const Rule rule = make_rule(in_plane, thickness, axis);
    const MaterialData m = read_material(mat);
    IncompatibleHex8<double> el(m, rule, kinematics);
    el.set_up(x0, ue, 0);
    std::vector<PlasticState<double>> committed(static_cast<std::size_t>(rule.n));
    for (int q = 0; q < rule.n; ++q) committed[static_cast<std::size_t>(q)] = read_state<double>(states_in + q * kStateRecord);
    const Vec<double, 24> v = el.standard_force(committed);
    for (int i = 0; i < 24; ++i) f[i] = v(i);
  });
}

// Hex8IncompatibleElement::stiffness (Hex8Incompatible.cpp:94-127).
int oracle_linear_stiffness(const double* x0, double E, double nu, int in_plane, int thickness, int axis,
                            double* k, char* msg, int msglen) {
  return guarded(msg, msglen, [&] {
    const Rule rule = make_rule(in_plane, thickness, axis);
    bool ok = false;
    linear_incompatible_stiffness(x0, E, nu, rule, k, &ok);
    if (!ok) throw SolverError("Hex8 incompatible modes: the stiffness of the modes is not positive definite");
  });
}

// Assembler::build_pattern (Assembler.cpp:86-173). Call with outer == null
// to get nnz; then with outer (3 nn + 1), inner (nnz), block_offset (64 ne).
long long oracle_pattern(int nn, int ne, const int* conn, int* outer, int* inner, int* block_offset) {
  const Pattern p = build_pattern(nn, ne, conn);
  if (outer != nullptr) {
    std::memcpy(outer, p.outer.data(), p.outer.size() * sizeof(int));
    std::memcpy(inner, p.inner.data(), p.inner.size() * sizeof(int));
    std::memcpy(block_offset, p.block_offset.data(), p.block_offset.size() * sizeof(int));
  }
  return static_cast<long long>(p.inner.size());
}

// Assembler::scatter (Assembler.cpp:183-215): column-major values of K.
void oracle_scatter(int nn, int ne, const int* conn, const double* ke, double* values) {
  const Pattern p = build_pattern(nn, ne, conn);
  scatter(p, ne, conn, ke, values);
}

// NonlinearSystem.hpp:55-67 on the column-major values of the pattern of
// conn; product (3 nn) gets |K| |x|; returns the norm over the free DOFs.
double oracle_stiffness_gross(int nn, int ne, const int* conn, const double* values, const double* x,
                              const int* free_dofs, int num_free, double* product) {
  const Pattern p = build_pattern(nn, ne, conn);
  return stiffness_gross(p, values, x, free_dofs, num_free, product);
}

// PartitionedFactor::build + assemble (Forming.cpp:89-98, 208-236) on the
// pattern of conn for ascending free DOFs: block_outer (num_free + 1),
// block_inner and block_values (capacity entries each) and position (one
// per full nonzero); returns the block's nonzero count, or -1 when it
// exceeds capacity.
long long oracle_free_block(int nn, int ne, const int* conn, const double* values, const int* free_dofs,
                            int num_free, long long capacity, int* block_outer, int* block_inner,
                            double* block_values, long long* position) {
  const Pattern p = build_pattern(nn, ne, conn);
  const FreeBlock b = free_block(p, free_dofs, num_free);
  const long long nnz = static_cast<long long>(b.inner.size());
  if (nnz > capacity) return -1;
  std::copy(b.outer.begin(), b.outer.end(), block_outer);
  std::copy(b.inner.begin(), b.inner.end(), block_inner);
  std::copy(b.position.begin(), b.position.end(), position);
  free_block_values(b, values, block_values);
  return nnz;
}

// _interpolate_structured + raycast_top (results.py:262-283, raycast.py:27-100)
// without the nearest fill: surface (ny_s, nx_s, 3) -> z, hit (ny, nx);
// returns the folded triangles' count (nothing cast when positive).
long long oracle_surface_raycast(const double* surface, long long nxs, long long nys, double x0, double y0,
                                 long long nx, long long ny, double h, double* z, unsigned char* hit) {
  const HeightGrid g{x0, y0, h, nx, ny};
  return surface_raycast(surface, nxs, nys, g, z, hit);
}

// NonlinearSystem.hpp:240-262; returns the gross norm.
double oracle_gather_force(int nn, int ne, const int* conn, const double* fe, double* fint) {
  return gather_force(nn, ne, conn, fe, fint);
}

// Forming.cpp:280-297 on the column-major values of the cached pattern.
void oracle_add_node_blocks(int nn, int ne, const int* conn, double* values, int count, const int* nodes,
                            const double* blocks) {
  const Pattern p = build_pattern(nn, ne, conn);
  for (int i = 0; i < count; ++i) add_node_block(p, values, nodes[i], blocks + 9 * i);
}

// ToolContact::evaluate (RigidTool.cpp:398-555) for one tool.
// tool: shape, centre[3], centre0[3], radius, normal[3], axis[3], friction,
// tangential_ratio (16 doubles). scalars_out: gross, round_off, force_norm,
// resultant[3], active, slipping, max_penetration, symmetric (10).
// The tool record of the contact entries: shape, centre 3, centre0 3, radius,
// normal 3, axis 3, friction, tangential ratio.
static Tool tool_from_record(const double* tool) {
  Tool t;
  t.shape = static_cast<int>(tool[0]);
  for (int i = 0; i < 3; ++i) {
    t.centre[i] = tool[1 + i];
    t.centre0[i] = tool[4 + i];
    t.normal[i] = tool[8 + i];
    t.axis[i] = tool[11 + i];
  }
  t.radius = tool[7];
  t.friction = tool[14];
  t.tangential_ratio = tool[15];
  return t;
}

// ToolContact::begin_increment for one tool; candidates (ns capacity) gets
// the candidate slave indices in order; returns their count.
int oracle_contact_candidates(const double* tool, const double* coords, const double* u0, int ns, const int* node,
                              double margin, int* candidates) {
  const std::vector<int> c = contact_candidates(tool_from_record(tool), coords, u0, ns, node, margin);
  for (std::size_t i = 0; i < c.size(); ++i) candidates[i] = c[i];
  return static_cast<int>(c.size());
}

// ToolContact::anticipate for one tool over the listed slaves; returns the
// number added.
int oracle_contact_anticipate(const double* tool, const double* coords, const double* u, const double* du, int ns,
                              const int* node, const double* kappa_a, double* residual, double* blocks, int* active) {
  return contact_anticipate(tool_from_record(tool), coords, u, du, ns, node, kappa_a, residual, blocks, active);
}

int oracle_contact(const double* tool, int friction_symmetric, int nn, const double* coords, const double* u,
                   const double* start, int ns, const int* node, const double* kappa_a, const int* h_present,
                   const double* h_force, const double* h_relative, int want_tangent, double* residual,
                   double* blocks, int* active, int* t_present, double* t_force, double* t_relative,
                   double* scalars, char* msg, int msglen) {
  Tool t;
  t.shape = static_cast<int>(tool[0]);
  for (int i = 0; i < 3; ++i) {
    t.centre[i] = tool[1 + i];
    t.centre0[i] = tool[4 + i];
    t.normal[i] = tool[8 + i];
    t.axis[i] = tool[11 + i];
  }
  t.radius = tool[7];
  t.friction = tool[14];
  t.tangential_ratio = tool[15];
  return guarded(msg, msglen, [&] {
    const ContactOut o = contact_evaluate(t, friction_symmetric, nn, coords, u, start, ns, node, kappa_a,
                                          h_present, h_force, h_relative, want_tangent != 0, residual,
                                          blocks, active, t_present, t_force, t_relative);
    scalars[0] = o.gross;
    scalars[1] = o.round_off;
    scalars[2] = o.force_norm;
    for (int i = 0; i < 3; ++i) scalars[3 + i] = o.resultant[i];
    scalars[6] = o.active;
    scalars[7] = o.slipping;
    scalars[8] = o.max_penetration;
    scalars[9] = o.symmetric ? 1.0 : 0.0;
  });
}

}  // extern "C"
