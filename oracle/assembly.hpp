// Oracle (test infrastructure only): the reference's assembly and contact.
//
//   * src/fem/Assembler.cpp:86-173 (build_pattern: node incidence, sorted
//     neighbourhoods, column-major structure, per-element block offsets)
//   * src/fem/Assembler.cpp:183-215 (scatter, in element order)
//   * src/fem/NonlinearSystem.hpp:240-262 (gather of f_int and the gross norm)
//   * src/fem/RigidTool.cpp:237-270 (gap), :398-555 (ToolContact::evaluate)
//   * src/fem/Forming.cpp:280-297 (add_node_block)
#pragma once

#include "om.hpp"

#include <algorithm>
#include <limits>
#include <vector>

namespace oracle {

struct Pattern {
  std::vector<int> outer, inner, block_offset;
};

// Assembler.cpp:86-173 (three DOFs per node, eight nodes per element).
inline Pattern build_pattern(int nn, int ne, const int* conn) {
  const int npe = 8, dim = 3;
  std::vector<int> inc_ptr(static_cast<std::size_t>(nn) + 1, 0);
  for (int e = 0; e < ne; ++e)
    for (int a = 0; a < npe; ++a) ++inc_ptr[static_cast<std::size_t>(conn[e * npe + a]) + 1];
  for (int n = 0; n < nn; ++n) inc_ptr[n + 1] += inc_ptr[n];
  std::vector<int> inc(static_cast<std::size_t>(inc_ptr.back()));
  {
    std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int e = 0; e < ne; ++e)
      for (int a = 0; a < npe; ++a) inc[static_cast<std::size_t>(fill[conn[e * npe + a]]++)] = e;
  }
  std::vector<int> nbr_ptr(static_cast<std::size_t>(nn) + 1, 0), nbr;
  {
    std::vector<int> marker(static_cast<std::size_t>(nn), -1), list;
    for (int b = 0; b < nn; ++b) {
      list.clear();
      for (int k = inc_ptr[b]; k < inc_ptr[b + 1]; ++k) {
        const int* nodes = conn + inc[k] * npe;
        for (int a = 0; a < npe; ++a)
          if (marker[nodes[a]] != b) { marker[nodes[a]] = b; list.push_back(nodes[a]); }
      }
      std::sort(list.begin(), list.end());
      nbr.insert(nbr.end(), list.begin(), list.end());
      nbr_ptr[b + 1] = static_cast<int>(nbr.size());
    }
  }
  Pattern p;
  const int n = nn * dim;
  p.outer.assign(static_cast<std::size_t>(n) + 1, 0);
  for (int b = 0; b < nn; ++b) {
    const int deg = nbr_ptr[b + 1] - nbr_ptr[b];
    for (int kb = 0; kb < dim; ++kb) {
      const int c = b * dim + kb;
      p.outer[c + 1] = p.outer[c] + deg * dim;
    }
  }
  p.inner.resize(static_cast<std::size_t>(p.outer.back()));
  for (int b = 0; b < nn; ++b)
    for (int kb = 0; kb < dim; ++kb) {
      int q = p.outer[b * dim + kb];
      for (int k = nbr_ptr[b]; k < nbr_ptr[b + 1]; ++k)
        for (int ka = 0; ka < dim; ++ka) p.inner[static_cast<std::size_t>(q++)] = nbr[k] * dim + ka;
    }
  p.block_offset.resize(static_cast<std::size_t>(ne) * npe * npe);
  for (int e = 0; e < ne; ++e) {
    const int* nodes = conn + e * npe;
    for (int b = 0; b < npe; ++b) {
      const int* begin = nbr.data() + nbr_ptr[nodes[b]];
      const int* end = nbr.data() + nbr_ptr[nodes[b] + 1];
      for (int a = 0; a < npe; ++a) {
        const int* hit = std::lower_bound(begin, end, nodes[a]);
        p.block_offset[(static_cast<std::size_t>(e) * npe + a) * npe + b] = static_cast<int>(hit - begin) * dim;
      }
    }
  }
  return p;
}

// Assembler.cpp:183-215: the column-major values of K, in element order.
// ke: element-major 24 x 24, row-major (ke[e*576 + i*24 + j] = K_e(i, j)).
inline void scatter(const Pattern& p, int ne, const int* conn, const double* ke, double* values) {
  const int npe = 8, dim = 3;
  std::fill(values, values + p.inner.size(), 0.0);
  for (int e = 0; e < ne; ++e) {
    const double* k = ke + static_cast<std::size_t>(e) * 576;
    const int* nodes = conn + e * npe;
    const int* offsets = p.block_offset.data() + static_cast<std::size_t>(e) * npe * npe;
    for (int b = 0; b < npe; ++b)
      for (int kb = 0; kb < dim; ++kb) {
        const int column = dim * b + kb;
        const int base = p.outer[nodes[b] * dim + kb];
        for (int a = 0; a < npe; ++a) {
          double* target = values + base + offsets[a * npe + b];
          for (int ka = 0; ka < dim; ++ka) target[ka] += k[(dim * a + ka) * 24 + column];
        }
      }
  }
}

// NonlinearSystem.hpp:240-262: f_int and the gross norm sqrt(sum_dof (sum_e |f_e|)^2).
inline double gather_force(int nn, int ne, const int* conn, const double* fe, double* fint) {
  std::vector<double> gross(static_cast<std::size_t>(nn) * 3, 0.0);
  std::fill(fint, fint + static_cast<std::size_t>(nn) * 3, 0.0);
  for (int e = 0; e < ne; ++e)
    for (int a = 0; a < 8; ++a)
      for (int k = 0; k < 3; ++k) {
        const int dof = conn[e * 8 + a] * 3 + k;
        const double v = fe[static_cast<std::size_t>(e) * 24 + 3 * a + k];
        fint[dof] += v;
        gross[static_cast<std::size_t>(dof)] += std::fabs(v);
      }
  double s = 0.0;
  for (double g : gross) s += g * g;
  return std::sqrt(s);
}

// NonlinearSystem.hpp:55-67 (stiffness_gross) on the column-major values of
// the pattern: product = |K| |x| column by column, then the 2-norm over the
// listed free DOFs in their order.
inline double stiffness_gross(const Pattern& p, const double* values, const double* x, const int* free_dofs,
                              int num_free, double* product) {
  const std::size_t n = p.outer.size() - 1;
  for (std::size_t i = 0; i < n; ++i) product[i] = 0.0;
  for (std::size_t col = 0; col < n; ++col) {
    const double uc = std::fabs(x[col]);
    if (uc == 0.0) continue;
    for (int q = p.outer[col]; q < p.outer[col + 1]; ++q) product[p.inner[q]] += std::fabs(values[q]) * uc;
  }
  double sum = 0.0;
  for (int i = 0; i < num_free; ++i) sum += product[free_dofs[i]] * product[free_dofs[i]];
  return std::sqrt(sum);
}

// PartitionedFactor::reset and build (Forming.cpp:75-85, 208-236): the
// free-free block of a column-major pattern for ascending free DOFs. Its
// columns are the free columns in order, each keeping its free rows in the
// full matrix's (sorted) order, renumbered to free positions; position[q] is
// where entry q of the full values lands in the block's values, or -1.
struct FreeBlock {
  std::vector<int> outer, inner;
  std::vector<long long> position;
};

inline FreeBlock free_block(const Pattern& p, const int* free_dofs, int num_free) {
  const std::size_t n = p.outer.size() - 1;
  std::vector<int> map(n, -1);
  for (int i = 0; i < num_free; ++i) map[static_cast<std::size_t>(free_dofs[i])] = i;
  FreeBlock b;
  b.outer.assign(static_cast<std::size_t>(num_free) + 1, 0);
  b.position.assign(p.inner.size(), -1);
  for (std::size_t col = 0; col < n; ++col) {
    const int c = map[col];
    if (c < 0) continue;
    for (int q = p.outer[col]; q < p.outer[col + 1]; ++q) {
      const int r = map[static_cast<std::size_t>(p.inner[q])];
      if (r < 0) continue;
      b.position[static_cast<std::size_t>(q)] = static_cast<long long>(b.inner.size());
      b.inner.push_back(r);
    }
    b.outer[static_cast<std::size_t>(c) + 1] = static_cast<int>(b.inner.size());
  }
  return b;
}

// PartitionedFactor::assemble (Forming.cpp:89-98): the block's values out of
// the full values.
inline void free_block_values(const FreeBlock& b, const double* values, double* out) {
  for (std::size_t q = 0; q < b.position.size(); ++q)
    if (b.position[q] >= 0) out[b.position[q]] = values[q];
}

// ---------------------------------------------------------------------------
// Contact (RigidTool.cpp), three-dimensional meshes
// ---------------------------------------------------------------------------
struct Tool {
  int shape = 1;  // 0 plane, 1 sphere, 2 cylinder
  double centre[3] = {0, 0, 0};     // c(t)
  double centre0[3] = {0, 0, 0};    // c(t0), for nodes new to contact
  double radius = 0.0;
  double normal[3] = {0, 0, 1};
  double axis[3] = {0, 0, 1};
  double friction = 0.0;
  double tangential_ratio = 1.0;
};

// RigidTool.cpp:237-270
inline double gap(const Tool& t, const double x[3], double n[3], double dn[3][3]) {
  if (t.shape == 0) {
    const double nn = std::sqrt(t.normal[0] * t.normal[0] + t.normal[1] * t.normal[1] + t.normal[2] * t.normal[2]);
    for (int i = 0; i < 3; ++i) n[i] = t.normal[i] / nn;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dn[i][j] = 0.0;
    return n[0] * (x[0] - t.centre[0]) + n[1] * (x[1] - t.centre[1]) + n[2] * (x[2] - t.centre[2]);
  }
  double r[3] = {x[0] - t.centre[0], x[1] - t.centre[1], x[2] - t.centre[2]};
  double base[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (t.shape == 2) {
    const double an = std::sqrt(t.axis[0] * t.axis[0] + t.axis[1] * t.axis[1] + t.axis[2] * t.axis[2]);
    const double a[3] = {t.axis[0] / an, t.axis[1] / an, t.axis[2] / an};
    const double ar = a[0] * r[0] + a[1] * r[1] + a[2] * r[2];
    for (int i = 0; i < 3; ++i) r[i] -= ar * a[i];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) base[i][j] -= a[i] * a[j];
  }
  const double d = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (!(d > 0.5 * t.radius))
    throw SolverError("a node lies less than half the radius from the tool's centre or axis: the tool is through the surface");
  for (int i = 0; i < 3; ++i) n[i] = r[i] / d;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dn[i][j] = (base[i][j] - n[i] * n[j]) / d;
  return d - t.radius;
}

// RigidTool.cpp:44-50: the distance from p to the segment [a, b].
inline double segment_distance(const double p[3], const double a[3], const double b[3]) {
  const double s[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double len2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
  double t = 0.0;
  if (len2 > 0.0) t = std::clamp(((p[0] - a[0]) * s[0] + (p[1] - a[1]) * s[1] + (p[2] - a[2]) * s[2]) / len2, 0.0, 1.0);
  const double r[3] = {p[0] - a[0] - t * s[0], p[1] - a[1] - t * s[1], p[2] - a[2] - t * s[2]};
  return std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

// ToolContact::begin_increment (RigidTool.cpp:345-393) for one tool, three
// dimensions: the slaves (in slave order) whose node at u0 lies within
// radius + margin of the tool's path from c0 = centre0 to c1 = centre (a
// cylinder: in the plane normal to its axis; a plane: within margin of
// either position).
inline std::vector<int> contact_candidates(const Tool& t, const double* coords, const double* u0, int ns,
                                           const int* node, double margin) {
  std::vector<int> out;
  const double* c0 = t.centre0;
  const double* c1 = t.centre;
  for (int i = 0; i < ns; ++i) {
    const long n = node[i];
    double x[3];
    for (int k = 0; k < 3; ++k) x[k] = coords[3 * n + k] + u0[3 * n + k];
    bool candidate = false;
    if (t.shape == 1) {
      candidate = segment_distance(x, c0, c1) <= t.radius + margin;
    } else if (t.shape == 2) {
      const double an = std::sqrt(t.axis[0] * t.axis[0] + t.axis[1] * t.axis[1] + t.axis[2] * t.axis[2]);
      const double a[3] = {t.axis[0] / an, t.axis[1] / an, t.axis[2] / an};
      double px[3], pc[3];
      const double d0[3] = {x[0] - c0[0], x[1] - c0[1], x[2] - c0[2]};
      const double d1[3] = {c1[0] - c0[0], c1[1] - c0[1], c1[2] - c0[2]};
      for (int r = 0; r < 3; ++r) {
        px[r] = 0.0;
        pc[r] = 0.0;
        for (int k = 0; k < 3; ++k) {
          const double pk = (r == k ? 1.0 : 0.0) - a[r] * a[k];
          px[r] += pk * d0[k];
          pc[r] += pk * d1[k];
        }
      }
      const double zero[3] = {0.0, 0.0, 0.0};
      candidate = segment_distance(px, zero, pc) <= t.radius + margin;
    } else {
      const double nn = std::sqrt(t.normal[0] * t.normal[0] + t.normal[1] * t.normal[1] + t.normal[2] * t.normal[2]);
      const double nv[3] = {t.normal[0] / nn, t.normal[1] / nn, t.normal[2] / nn};
      const double g0 = nv[0] * (x[0] - c0[0]) + nv[1] * (x[1] - c0[1]) + nv[2] * (x[2] - c0[2]);
      const double g1 = nv[0] * (x[0] - c1[0]) + nv[1] * (x[1] - c1[1]) + nv[2] * (x[2] - c1[2]);
      candidate = std::min(g0, g1) <= margin;
    }
    if (candidate) out.push_back(i);
  }
  return out;
}

// ToolContact::anticipate (RigidTool.cpp:557-603) for one tool at its position
// centre, over the listed (candidate) slaves: a slave outside the tool at u
// whose node the correction du drives into it (or through it) gets
// residual += kappa A g n and the block kappa A n n^T; returns their count.
inline int contact_anticipate(const Tool& t, const double* coords, const double* u, const double* du, int ns,
                              const int* node, const double* kappa_a, double* residual, double* blocks,
                              int* active) {
  int added = 0;
  for (int i = 0; i < ns; ++i) {
    active[i] = 0;
    const long n = node[i];
    double x[3], xn[3];
    for (int k = 0; k < 3; ++k) {
      x[k] = coords[3 * n + k] + u[3 * n + k];
      xn[k] = x[k] + du[3 * n + k];
    }
    double nrm[3], dn[3][3], g = 0.0;
    try {
      g = gap(t, x, nrm, dn);
    } catch (const SolverError&) {
      continue;  // deep in the tool
    }
    if (!(g > 0.0)) continue;
    double g_next = 0.0;
    try {
      double n2[3];
      g_next = gap(t, xn, n2, dn);
    } catch (const SolverError&) {
      g_next = -std::numeric_limits<double>::infinity();
    }
    if (!(g_next < 0.0)) continue;
    const double ka = kappa_a[i];
    for (int k = 0; k < 3; ++k) residual[3 * n + k] += ka * g * nrm[k];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) blocks[9 * i + 3 * a + b] = ka * (nrm[a] * nrm[b]);
    active[i] = 1;
    ++added;
  }
  return added;
}

struct ContactOut {
  double gross = 0.0, round_off = 0.0, force_norm = 0.0;
  double resultant[3] = {0, 0, 0};
  int active = 0, slipping = 0;
  double max_penetration = 0.0;
  bool symmetric = true;
};

// RigidTool.cpp:398-555 for one tool over the listed slaves. Per slave i:
// node[i], kappa_a[i] = kappa * A, committed history (present, force, relative).
// residual: 3 nn, decremented by each contact force; blocks: 9 per slave
// (row-major), active[i]: the node is in contact; trial history per slave.
inline ContactOut contact_evaluate(const Tool& tool, int friction_symmetric, int nn, const double* coords,
                                   const double* u, const double* start, int ns, const int* node,
                                   const double* kappa_a, const int* h_present, const double* h_force,
                                   const double* h_relative, bool want_tangent, double* residual,
                                   double* blocks, int* active, int* t_present, double* t_force,
                                   double* t_relative) {
  const double kVanishing = 1.0e-12;
  ContactOut out;
  std::vector<double> gross(static_cast<std::size_t>(nn) * 3, 0.0), positions(static_cast<std::size_t>(nn) * 3, 0.0);
  const double* c = tool.centre;
  const double mu = tool.friction;
  for (int i = 0; i < ns; ++i) {
    active[i] = 0;
    t_present[i] = 0;
    const int nd = node[i];
    double x[3];
    for (int k = 0; k < 3; ++k) x[k] = coords[nd * 3 + k] + u[nd * 3 + k];
    double nrm[3], dn[3][3];
    const double g = gap(tool, x, nrm, dn);
    if (!(g < 0.0)) continue;
    const double ka = kappa_a[i];
    const double pn = g_plant == kPlantNormalForce ? -ka * g * (1.0 + 1.0e-6) : -ka * g;
    double force[3] = {pn * nrm[0], pn * nrm[1], pn * nrm[2]};
    double k_node[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) k_node[a][b] = ka * (nrm[a] * nrm[b]) - pn * dn[a][b];
    bool slipping = false;
    const double r[3] = {x[0] - c[0], x[1] - c[1], x[2] - c[2]};
    if (mu > 0.0) {
      const double kt = tool.tangential_ratio * ka;
      double f0[3] = {0, 0, 0}, q[3];
      if (h_present[i]) {
        for (int k = 0; k < 3; ++k) { f0[k] = h_force[i * 3 + k]; q[k] = h_relative[i * 3 + k]; }
      } else {
        for (int k = 0; k < 3; ++k) q[k] = (coords[nd * 3 + k] + start[nd * 3 + k]) - tool.centre0[k];
      }
      const double s0 = std::sqrt(f0[0] * f0[0] + f0[1] * f0[1] + f0[2] * f0[2]);
      const double fn = f0[0] * nrm[0] + f0[1] * nrm[1] + f0[2] * nrm[2];
      double f_tilde[3];
      for (int k = 0; k < 3; ++k) f_tilde[k] = f0[k] - fn * nrm[k];
      const double ftn = std::sqrt(f_tilde[0] * f_tilde[0] + f_tilde[1] * f_tilde[1] + f_tilde[2] * f_tilde[2]);
      double f_hat[3] = {0, 0, 0}, d_hat[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      if (s0 > 0.0 && ftn > kVanishing * s0) {
        double e[3];
        for (int k = 0; k < 3; ++k) { e[k] = f_tilde[k] / ftn; f_hat[k] = s0 * e[k]; }
        // d_tilde = -(n (f0^T dn) + (f0 . n) dn)
        double f0dn[3];
        for (int j = 0; j < 3; ++j) f0dn[j] = f0[0] * dn[0][j] + f0[1] * dn[1][j] + f0[2] * dn[2][j];
        double d_tilde[3][3];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) d_tilde[a][b] = -(nrm[a] * f0dn[b] + fn * dn[a][b]);
        const double sc = s0 / ftn;
        double pe[3][3];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) pe[a][b] = sc * ((a == b ? 1.0 : 0.0) - e[a] * e[b]);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) d_hat[a][b] = pe[a][0] * d_tilde[0][b] + pe[a][1] * d_tilde[1][b] + pe[a][2] * d_tilde[2][b];
      }
      double w[3] = {r[0] - q[0], r[1] - q[1], r[2] - q[2]};
      double p[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) p[a][b] = (a == b ? 1.0 : 0.0) - nrm[a] * nrm[b];
      double delta[3];
      for (int a = 0; a < 3; ++a) delta[a] = p[a][0] * w[0] + p[a][1] * w[1] + p[a][2] * w[2];
      const double nw = nrm[0] * w[0] + nrm[1] * w[1] + nrm[2] * w[2];
      double wdn[3];
      for (int j = 0; j < 3; ++j) wdn[j] = w[0] * dn[0][j] + w[1] * dn[1][j] + w[2] * dn[2][j];
      double d_delta[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) d_delta[a][b] = p[a][b] - nw * dn[a][b] - nrm[a] * wdn[b];
      double trial[3], d_trial[3][3];
      for (int a = 0; a < 3; ++a) trial[a] = f_hat[a] - kt * delta[a];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) d_trial[a][b] = d_hat[a][b] - kt * d_delta[a][b];
      const double tn = std::sqrt(trial[0] * trial[0] + trial[1] * trial[1] + trial[2] * trial[2]);
      const double limit = g_plant == kPlantSlipRadius ? mu * ka : mu * pn;
      double ft[3], d_ft[3][3], k_sym[3][3];
      if (tn <= limit) {
        for (int a = 0; a < 3; ++a) ft[a] = trial[a];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) { d_ft[a][b] = d_trial[a][b]; k_sym[a][b] = kt * p[a][b]; }
      } else {
        slipping = true;
        double e[3];
        for (int a = 0; a < 3; ++a) { e[a] = trial[a] / tn; ft[a] = limit * e[a]; }
        const double mk = mu * ka, lt = limit / tn;
        double pe[3][3];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) pe[a][b] = lt * ((a == b ? 1.0 : 0.0) - e[a] * e[b]);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            const double pd = pe[a][0] * d_trial[0][b] + pe[a][1] * d_trial[1][b] + pe[a][2] * d_trial[2][b];
            d_ft[a][b] = -mk * (e[a] * nrm[b]) + pd;
            k_sym[a][b] = lt * kt * (p[a][b] - e[a] * e[b]);
          }
      }
      for (int a = 0; a < 3; ++a) force[a] += ft[a];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) k_node[a][b] = friction_symmetric ? k_node[a][b] + k_sym[a][b] : k_node[a][b] - d_ft[a][b];
      t_present[i] = 1;
      for (int a = 0; a < 3; ++a) { t_force[i * 3 + a] = ft[a]; t_relative[i * 3 + a] = r[a]; }
    }
    const double xn = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const double cn = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const double position_scale = ka * (xn + cn);
    for (int k = 0; k < 3; ++k) {
      residual[nd * 3 + k] -= force[k];
      gross[static_cast<std::size_t>(nd) * 3 + k] += std::fabs(force[k]);
      positions[static_cast<std::size_t>(nd) * 3 + k] += position_scale;
    }
    ++out.active;
    if (slipping) ++out.slipping;
    for (int k = 0; k < 3; ++k) out.resultant[k] -= force[k];
    out.max_penetration = std::max(out.max_penetration, -g);
    active[i] = 1;
    if (want_tangent) {
      if (friction_symmetric) {
        double ks[3][3];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) ks[a][b] = 0.5 * (k_node[a][b] + k_node[b][a]);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) k_node[a][b] = ks[a][b];
      }
      double top = 0.0, asym = 0.0;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          top = std::max(top, std::fabs(k_node[a][b]));
          asym = std::max(asym, std::fabs(k_node[a][b] - k_node[b][a]));
        }
      if (asym > 1.0e-12 * top) out.symmetric = false;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) blocks[i * 9 + a * 3 + b] = k_node[a][b];
    }
  }
  double sg = 0.0, sp = 0.0, sr = 0.0;
  for (std::size_t d = 0; d < gross.size(); ++d) {
    sg += gross[d] * gross[d];
    sp += positions[d] * positions[d];
    sr += residual[d] * residual[d];
  }
  out.gross = std::sqrt(sg);
  out.round_off = std::sqrt(sp);
  out.force_norm = std::sqrt(sr);
  return out;
}

// Forming.cpp:280-297 on column-major values: block(a, b) into column 3 node + b.
inline void add_node_block(const Pattern& p, double* values, int node, const double* block) {
  for (int b = 0; b < 3; ++b) {
    const int col = node * 3 + b;
    const int begin = p.outer[col], end = p.outer[col + 1];
    for (int a = 0; a < 3; ++a) {
      const int row = node * 3 + a;
      const int* it = std::lower_bound(p.inner.data() + begin, p.inner.data() + end, row);
      values[it - p.inner.data()] += block[a * 3 + b];
    }
  }
}

}  // namespace oracle
