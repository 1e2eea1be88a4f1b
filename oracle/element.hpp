// Oracle (test infrastructure only): the incompatible-mode Hex8 of the reference.
//
//   * quadrature: src/elements/Quadrature.cpp:16-219 (line rules 1-4, Newton
//     rules 5-7, cube and box rules, xi fastest)
//   * shape gradients: src/elements/Hex8.cpp:17-19, 52-64
//   * mode gradients: src/elements/Hex8Incompatible.cpp:73-83
//   * the non-linear kernel: src/fem/IncompatibleModes.cpp:101-121 (B_NL, E),
//     :134-192 (set_up), :195-222 (kinematics_at), :227-247 (respond),
//     :252-275 (accumulate_local), :279-306 (accumulate_coupling),
//     :320-562 (solve_and_condense), :577-600 (the element)
//   * constants: src/fem/IncompatibleModes.hpp:98-127
//   * the linear condensed stiffness: src/elements/Hex8Incompatible.cpp:94-127
//     with src/elements/Hex8Kernels.hpp:60-120
#pragma once

#include "material.hpp"

#include <functional>
#include <limits>
#include <vector>

namespace oracle {

enum Kinematics { kSmallStrain = 0, kFinite = 1, kFiniteLogarithmic = 2 };

struct Rule {
  int n = 0;
  double xi[112], eta[112], zeta[112], w[112];
};

// Quadrature.cpp:16-48, 100-139: Gauss-Legendre on [-1, 1].
inline std::vector<std::pair<double, double>> line_rule(int n) {
  if (n == 1) return {{0.0, 2.0}};
  if (n == 2) {
    const double a2 = 1.0 / std::sqrt(3.0);
    return {{-a2, 1.0}, {a2, 1.0}};
  }
  if (n == 3) {
    const double a3 = std::sqrt(3.0 / 5.0);
    return {{-a3, 5.0 / 9.0}, {0.0, 8.0 / 9.0}, {a3, 5.0 / 9.0}};
  }
  if (n == 4) {
    const double r = std::sqrt(6.0 / 5.0);
    const double x1 = std::sqrt((3.0 - 2.0 * r) / 7.0);
    const double x2 = std::sqrt((3.0 + 2.0 * r) / 7.0);
    const double s = std::sqrt(30.0);
    const double w1 = (18.0 + s) / 36.0, w2 = (18.0 - s) / 36.0;
    return {{-x2, w2}, {-x1, w1}, {x1, w1}, {x2, w2}};
  }
  if (n < 1 || n > 7) throw ConfigError("Gauss-Legendre rule not available");
  std::vector<std::pair<double, double>> rule(static_cast<std::size_t>(n));
  const double pi = std::acos(-1.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double x = std::cos(pi * (static_cast<double>(i) + 0.75) / (static_cast<double>(n) + 0.5));
    double derivative = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double pn = n == 1 ? x : p1;
      const double pm = n == 1 ? 1.0 : p0;
      derivative = static_cast<double>(n) * (x * pn - pm) / (x * x - 1.0);
      const double step = pn / derivative;
      x -= step;
      if (std::abs(step) <= 1.0e-16) break;
    }
    if (2 * i + 1 == n) x = 0.0;
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    const double pn = n == 1 ? x : p1;
    const double pm = n == 1 ? 1.0 : p0;
    derivative = static_cast<double>(n) * (x * pn - pm) / (x * x - 1.0);
    const double weight = 2.0 / ((1.0 - x * x) * derivative * derivative);
    rule[static_cast<std::size_t>(i)] = {-x, weight};
    rule[static_cast<std::size_t>(n - 1 - i)] = {x, weight};
  }
  return rule;
}

// Hex8Kernels.hpp:125-130 with Quadrature.cpp:72-93, 165-219.
inline Rule make_rule(int in_plane, int thickness, int axis) {
  Rule r;
  if (in_plane < 1 || in_plane > 4) throw ConfigError("in-plane points must be 1..4");
  const int nt = (thickness <= 0 || thickness == in_plane) ? in_plane : thickness;
  if (nt < 1 || nt > 7) throw ConfigError("thickness points must be 1..7");
  if (axis < 0 || axis > 2) throw ConfigError("thickness axis must be 0, 1 or 2");
  const auto plane = line_rule(in_plane);
  const auto through = line_rule(nt);
  const auto& rx = (axis == 0) ? through : plane;
  const auto& ry = (axis == 1) ? through : plane;
  const auto& rz = (axis == 2) ? through : plane;
  for (const auto& gk : rz)
    for (const auto& gj : ry)
      for (const auto& gi : rx) {
        r.xi[r.n] = gi.first;
        r.eta[r.n] = gj.first;
        r.zeta[r.n] = gk.first;
        r.w[r.n] = gi.second * gj.second * gk.second;
        ++r.n;
      }
  return r;
}

constexpr double kCorner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                  {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// Hex8.cpp:52-64
inline void shape_gradients_natural(double xi, double eta, double zeta, double g[8][3]) {
  for (int a = 0; a < 8; ++a) {
    const double fx = 1.0 + xi * kCorner[a][0];
    const double fy = 1.0 + eta * kCorner[a][1];
    const double fz = 1.0 + zeta * kCorner[a][2];
    g[a][0] = 0.125 * kCorner[a][0] * fy * fz;
    g[a][1] = 0.125 * kCorner[a][1] * fx * fz;
    g[a][2] = 0.125 * kCorner[a][2] * fx * fy;
  }
}

// IncompatibleModes.hpp:98-127
constexpr int kMaxLocalIterations = 25;
constexpr double kFullStepReduction = 1.0 - 1.0e-4;
constexpr int kMaxLineSearch = 8;
constexpr double kLineSearchRatio = 0.5;
constexpr double kLocalShift = 1.0e-3;
constexpr double kLocalMarquardt = 1.0e-6;
constexpr double kLocalTolerance = 1.0e-8;
constexpr double kLocalRoundOff = 1.0e-13;

template <class T> struct PointWork {
  Mat<T, 3, 11> g;
  Mat<T, 3, 3> hu;
  T weight;
  Mat<T, 3, 3> h;
  Mat<T, 6, 33> mb;
  Vec<T, 6> strain;
  LogStrain<T> log;
  Vec<T, 6> s;
  Mat<T, 6, 6> material;
  Vec<T, 6> own;
  T energy;
  PlasticState<T> state;
  bool yielding = false;
  T strain_33;
};

template <class T> struct Sums {
  Vec<T, 33> r = Vec<T, 33>::zero();
  Vec<T, 9> scale = Vec<T, 9>::zero();
  Vec<T, 9> floor = Vec<T, 9>::zero();
  Mat<T, 9, 9> kaa = Mat<T, 9, 9>::zero();
  Mat<T, 24, 9> kua = Mat<T, 24, 9>::zero();
  Mat<T, 9, 24> kau = Mat<T, 9, 24>::zero();
  Mat<T, 24, 24> kuu = Mat<T, 24, 24>::zero();
  T energy = T(0.0);
};

// IncompatibleModes.cpp:101-106
template <class T> Vec<T, 6> green_lagrange_6(const Mat<T, 3, 3>& h) {
  Mat<T, 3, 3> e;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const T hth = h(0, i) * h(0, j) + h(1, i) * h(1, j) + h(2, i) * h(2, j);
      e(i, j) = T(0.5) * ((h(i, j) + h(j, i)) + hth);
    }
  Vec<T, 6> v;
  v(0) = e(0, 0); v(1) = e(1, 1); v(2) = e(2, 2);
  v(3) = T(2.0) * e(0, 1); v(4) = T(2.0) * e(1, 2); v(5) = T(2.0) * e(2, 0);
  return v;
}

// IncompatibleModes.cpp:110-121
template <class T> void green_lagrange_operator_11(const Mat<T, 3, 3>& f, const Mat<T, 3, 11>& g,
                                                   Mat<T, 6, 33>& b) {
  static constexpr int kPair[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {2, 0}};
  for (int c = 0; c < 6; ++c) {
    const int i = kPair[c][0], j = kPair[c][1];
    for (int a = 0; a < 11; ++a)
      for (int k = 0; k < 3; ++k)
        b(c, 3 * a + k) = i == j ? f(k, i) * g(i, a) : f(k, i) * g(j, a) + f(k, j) * g(i, a);
  }
}

template <class T> struct ElementResult {
  Vec<T, 24> force;
  Mat<T, 24, 24> tangent;
  T energy = T(0.0);
  std::vector<PlasticState<T>> states;
  int yielding_points = 0;
  bool symmetric = true;
  Vec<T, 9> internal;
  int internal_iterations = 0;
  int passes = 0;  // pass() calls, successful or not (diagnostic only)
  int path = 0;    // the local iteration's phases entered (kPath* bits; diagnostic only)
};

// Bits of ElementResult::path: the phases of the local iteration
// (IncompatibleModes.cpp:398-546) an evaluation went through.
constexpr int kPathRegulaFalsi = 1;   // the line search's regula falsi on d . r_a
constexpr int kPathNormSearch = 2;    // search_norm: halvings on |r_a|
constexpr int kPathIndefinite = 4;    // sym(K_aa) not positive definite
constexpr int kPathShifted = 8;       // the shifted direction (K_aa + mu I) d = -r_a
constexpr int kPathMarquardt = 16;    // Levenberg-Marquardt
constexpr int kPathHalved = 32;       // a trial that failed, halved in the line search

// include/sparlab/fem/Elastoplastic.hpp:206-221 (ElastoplasticStress).
template <class T> struct ElementStress {
  Vec<T, 6> cauchy = Vec<T, 6>::zero();
  Vec<T, 6> piola_kirchhoff = Vec<T, 6>::zero();
  Vec<T, 6> kirchhoff = Vec<T, 6>::zero();
  Vec<T, 6> logarithmic_strain = Vec<T, 6>::zero();
  T von_mises = T(0.0);
  T max_equivalent_plastic_strain = T(0.0);
  T max_strain = T(0.0);
  T min_jacobian = T(1.0);
};

// Plasticity.cpp:541-547
template <class T> T von_mises_stress(const Vec<T, 6>& s) {
  const T a = s(0) - s(1), b = s(1) - s(2), c = s(2) - s(0);
  return sqrt_(T(0.5) * (a * a + b * b + c * c) + T(3.0) * (s(3) * s(3) + s(4) * s(4) + s(5) * s(5)));
}

template <class T>
class IncompatibleHex8 {
 public:
  IncompatibleHex8(const MaterialData& m, const Rule& rule, int kinematics)
      : m_(m), rule_(rule), kin_(kinematics), symmetric_(m.symmetric_tangent()) {}

  // IncompatibleModes.cpp:134-192. x0: 8 nodes x 3 coordinates; ue: 24 node-major.
  void set_up(const double* x0, const T* ue, long e) {
    e_ = e;
    for (int a = 0; a < 8; ++a)
      for (int k = 0; k < 3; ++k) { u_(k, a) = ue[3 * a + k]; ue_(3 * a + k) = ue[3 * a + k]; }
    double dn0[8][3];
    shape_gradients_natural(0.0, 0.0, 0.0, dn0);
    Mat<T, 3, 3> jac0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        T s = T(x0[0 * 3 + i]) * T(dn0[0][j]);
        for (int a = 1; a < 8; ++a) s = s + T(x0[a * 3 + i]) * T(dn0[a][j]);
        jac0(i, j) = s;
      }
    const T det0 = det3(jac0);
    if (!(det0 > T(0.0))) {
      std::ostringstream os;
      os << "Hex8 element " << e << " has a Jacobian determinant of " << val(det0)
         << " m^3 at its centre; it is inverted or degenerate";
      throw MeshError(os.str());
    }
    const Mat<T, 3, 3> jinv0 = inverse3(jac0);
    points_.resize(static_cast<std::size_t>(rule_.n));
    for (int q = 0; q < rule_.n; ++q) {
      PointWork<T>& p = points_[static_cast<std::size_t>(q)];
      double dn[8][3];
      shape_gradients_natural(rule_.xi[q], rule_.eta[q], rule_.zeta[q], dn);
      Mat<T, 3, 3> jac;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          T s = T(x0[0 * 3 + i]) * T(dn[0][j]);
          for (int a = 1; a < 8; ++a) s = s + T(x0[a * 3 + i]) * T(dn[a][j]);
          jac(i, j) = s;
        }
      const T det = det3(jac);
      if (!(det > T(0.0))) {
        std::ostringstream os;
        os << "Hex8 element " << e << " has a Jacobian determinant of " << val(det)
           << " m^3 at an integration point; it is inverted or folded";
        throw MeshError(os.str());
      }
      const Mat<T, 3, 3> jinv = inverse3(jac);
      for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i)
          p.g(i, a) = T(dn[a][0]) * jinv(0, i) + T(dn[a][1]) * jinv(1, i) + T(dn[a][2]) * jinv(2, i);
      const T ratio = det0 / det;
      const double natural[3] = {rule_.xi[q], rule_.eta[q], rule_.zeta[q]};
      for (int k = 0; k < 3; ++k) {
        const T f = ratio * T(-2.0) * T(natural[k]);
        for (int i = 0; i < 3; ++i) p.g(i, 8 + k) = f * jinv0(k, i);
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          T s = u_(i, 0) * p.g(j, 0);
          for (int a = 1; a < 8; ++a) s = s + u_(i, a) * p.g(j, a);
          p.hu(i, j) = s;
        }
      p.weight = T(rule_.w[q]) * det;
      if (kin_ == kSmallStrain) green_lagrange_operator_11(Mat<T, 3, 3>::identity(), p.g, p.mb);
    }
  }

  // IncompatibleModes.cpp:195-222
  void kinematics_at(const Vec<T, 9>& alpha) {
    Mat<T, 3, 3> a;
    for (int m = 0; m < 3; ++m)
      for (int k = 0; k < 3; ++k) a(k, m) = alpha(3 * m + k);
    for (PointWork<T>& p : points_) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          p.h(i, j) = p.hu(i, j) + (a(i, 0) * p.g(j, 8) + a(i, 1) * p.g(j, 9) + a(i, 2) * p.g(j, 10));
      if (kin_ == kSmallStrain) {
        for (int c = 0; c < 6; ++c) {
          T s1 = p.mb(c, 0) * ue_(0);
          for (int k = 1; k < 24; ++k) s1 = s1 + p.mb(c, k) * ue_(k);
          T s2 = p.mb(c, 24) * alpha(0);
          for (int k = 1; k < 9; ++k) s2 = s2 + p.mb(c, 24 + k) * alpha(k);
          p.strain(c) = s1 + s2;
        }
      } else {
        Mat<T, 3, 3> f = p.h;
        for (int i = 0; i < 3; ++i) f(i, i) = T(1.0) + p.h(i, i);
        if (kin_ == kFiniteLogarithmic && !(det3(f) > T(0.0))) {
          std::ostringstream os;
          os << "an element is inverted (det F = " << val(det3(f))
             << "); the load step is too large or the mesh too coarse for this deformation";
          throw SolverError(os.str());
        }
        green_lagrange_operator_11(f, p.g, p.mb);
        if (kin_ == kFinite) {
          p.strain = green_lagrange_6(p.h);
        } else {
          p.log = logarithmic_strain(green_lagrange_6(p.h));
          p.strain = p.log.strain;
        }
      }
    }
  }

  // IncompatibleModes.cpp:227-247
  void respond(const std::vector<PlasticState<T>>& committed) {
    for (std::size_t q = 0; q < points_.size(); ++q) {
      PointWork<T>& p = points_[q];
      const PlasticResponse<T> r = plastic_return(m_, p.strain, committed[q], true);
      p.own = r.stress;
      p.energy = r.energy;
      p.state = r.state;
      p.yielding = r.yielding;
      p.strain_33 = r.strain_33;
      if (kin_ == kFiniteLogarithmic) {
        p.s = logarithmic_stress(p.log, r.stress);
        const Mat<T, 6, 6> pc = mul(mul_tn(p.log.projection, r.tangent), p.log.projection);
        const Mat<T, 6, 6> cv = logarithmic_curvature(p.log, r.stress);
        for (int i = 0; i < 36; ++i) p.material.a[i] = pc.a[i] + cv.a[i];
      } else {
        p.s = r.stress;
        p.material = r.tangent;
      }
    }
  }

  // IncompatibleModes.cpp:252-275
  void accumulate_local(Sums<T>& sums) const {
    const bool geometric = kin_ != kSmallStrain;
    for (const PointWork<T>& p : points_) {
      const T w = p.weight;
      for (int i = 0; i < 33; ++i) {
        T s = p.mb(0, i) * p.s(0);
        for (int c = 1; c < 6; ++c) s = s + p.mb(c, i) * p.s(c);
        sums.r(i) = sums.r(i) + w * s;
      }
      for (int m = 0; m < 9; ++m) {
        T s = abs_(p.mb(0, 24 + m)) * abs_(p.s(0));
        for (int c = 1; c < 6; ++c) s = s + abs_(p.mb(c, 24 + m)) * abs_(p.s(c));
        sums.scale(m) = sums.scale(m) + w * s;
      }
      const T h = max_abs(p.h);
      const T f = w * max_abs(p.material) * h * (T(1.0) + h);
      for (int m = 0; m < 9; ++m) {
        T s = abs_(p.mb(0, 24 + m));
        for (int c = 1; c < 6; ++c) s = s + abs_(p.mb(c, 24 + m));
        sums.floor(m) = sums.floor(m) + f * s;
      }
      sums.energy = sums.energy + w * p.energy;
      Mat<T, 6, 9> y;
      for (int c = 0; c < 6; ++c)
        for (int m = 0; m < 9; ++m) {
          T s = p.material(c, 0) * p.mb(0, 24 + m);
          for (int d = 1; d < 6; ++d) s = s + p.material(c, d) * p.mb(d, 24 + m);
          y(c, m) = s;
        }
      for (int m = 0; m < 9; ++m)
        for (int n = 0; n < 9; ++n) {
          T s = p.mb(0, 24 + m) * y(0, n);
          for (int c = 1; c < 6; ++c) s = s + p.mb(c, 24 + m) * y(c, n);
          sums.kaa(m, n) = sums.kaa(m, n) + w * s;
        }
      if (geometric) {
        const Mat<T, 3, 3> st = tensor_of(p.s, 1.0);
        Mat<T, 3, 3> gs, geo;
        for (int a = 0; a < 3; ++a)
          for (int j = 0; j < 3; ++j) gs(a, j) = p.g(0, 8 + a) * st(0, j) + p.g(1, 8 + a) * st(1, j) + p.g(2, 8 + a) * st(2, j);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) geo(a, b) = gs(a, 0) * p.g(0, 8 + b) + gs(a, 1) * p.g(1, 8 + b) + gs(a, 2) * p.g(2, 8 + b);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            for (int k = 0; k < 3; ++k) sums.kaa(3 * a + k, 3 * b + k) = sums.kaa(3 * a + k, 3 * b + k) + w * geo(a, b);
      }
    }
  }

  // IncompatibleModes.cpp:279-306
  void accumulate_coupling(bool full, Sums<T>& sums) const {
    const bool geometric = kin_ != kSmallStrain;
    for (const PointWork<T>& p : points_) {
      const T w = p.weight;
      Mat<T, 6, 9> ya;
      for (int c = 0; c < 6; ++c)
        for (int m = 0; m < 9; ++m) {
          T s = p.material(c, 0) * p.mb(0, 24 + m);
          for (int d = 1; d < 6; ++d) s = s + p.material(c, d) * p.mb(d, 24 + m);
          ya(c, m) = s;
        }
      for (int i = 0; i < 24; ++i)
        for (int m = 0; m < 9; ++m) {
          T s = p.mb(0, i) * ya(0, m);
          for (int c = 1; c < 6; ++c) s = s + p.mb(c, i) * ya(c, m);
          sums.kua(i, m) = sums.kua(i, m) + w * s;
        }
      // material * B_u, evaluated once per product as the reference writes them.
      auto mat_bu = [&]() {
        Mat<T, 6, 24> yu;
        for (int c = 0; c < 6; ++c)
          for (int j = 0; j < 24; ++j) {
            T s = p.material(c, 0) * p.mb(0, j);
            for (int d = 1; d < 6; ++d) s = s + p.material(c, d) * p.mb(d, j);
            yu(c, j) = s;
          }
        return yu;
      };
      {
        if (!symmetric_) {
          const Mat<T, 6, 24> yu = mat_bu();
          for (int m = 0; m < 9; ++m)
            for (int j = 0; j < 24; ++j) {
              T s = p.mb(0, 24 + m) * yu(0, j);
              for (int c = 1; c < 6; ++c) s = s + p.mb(c, 24 + m) * yu(c, j);
              sums.kau(m, j) = sums.kau(m, j) + w * s;
            }
        }
        if (full) {
          const Mat<T, 6, 24> yu = mat_bu();
          for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
              T s = p.mb(0, i) * yu(0, j);
              for (int c = 1; c < 6; ++c) s = s + p.mb(c, i) * yu(c, j);
              sums.kuu(i, j) = sums.kuu(i, j) + w * s;
            }
        }
      }
      if (geometric) {
        const Mat<T, 3, 3> st = tensor_of(p.s, 1.0);
        Mat<T, 11, 3> gs;
        for (int a = 0; a < 11; ++a)
          for (int j = 0; j < 3; ++j) gs(a, j) = p.g(0, a) * st(0, j) + p.g(1, a) * st(1, j) + p.g(2, a) * st(2, j);
        for (int a = 0; a < 8; ++a)
          for (int b = 0; b < 11; ++b) {
            const T geo = gs(a, 0) * p.g(0, b) + gs(a, 1) * p.g(1, b) + gs(a, 2) * p.g(2, b);
            const T v = w * geo;
            for (int k = 0; k < 3; ++k) {
              if (b < 8) {
                if (full) sums.kuu(3 * a + k, 3 * b + k) = sums.kuu(3 * a + k, 3 * b + k) + v;
              } else {
                sums.kua(3 * a + k, 3 * (b - 8) + k) = sums.kua(3 * a + k, 3 * (b - 8) + k) + v;
                if (!symmetric_) sums.kau(3 * (b - 8) + k, 3 * a + k) = sums.kau(3 * (b - 8) + k, 3 * a + k) + v;
              }
            }
          }
      }
    }
  }

  // IncompatibleModes.cpp:320-562 and :577-600.
  ElementResult<T> evaluate(const std::vector<PlasticState<T>>& committed, const T* start,
                            bool want_tangent) {
    if (committed.size() != points_.size())
      throw SolverError("elastoplastic element: the stored internal variables do not match the integration rule");
    ElementResult<T> out;
    Vec<T, 9> alpha = Vec<T, 9>::zero();
    if (start != nullptr)
      for (int i = 0; i < 9; ++i) alpha(i) = start[i];
    Sums<T> sums;
    FullPivLU<T, 9> lu;
    int iterations = 0;
    int cuts = 0;
    auto r_alpha = [&]() {
      Vec<T, 9> v;
      for (int i = 0; i < 9; ++i) v(i) = sums.r(24 + i);
      return v;
    };
    auto pass = [&](const Vec<T, 9>& at) {
      ++out.passes;
      kinematics_at(at);
      respond(committed);
      sums = Sums<T>();
      accumulate_local(sums);
    };
    auto try_pass = [&](const Vec<T, 9>& at) {
      try {
        pass(at);
        return true;
      } catch (const SolverError&) {
        return false;
      }
    };
    auto axpy = [](const Vec<T, 9>& a, T s, const Vec<T, 9>& d) {
      Vec<T, 9> o;
      for (int i = 0; i < 9; ++i) o(i) = a(i) + s * d(i);
      return o;
    };
    pass(alpha);
    T residual = norm2(r_alpha());
    auto search_norm = [&](const Vec<T, 9>& d, T reach) {
      out.path |= kPathNormSearch;
      T best_step = T(0.0);
      T best = residual;
      bool last_is_best = false;
      T trial_step = reach;
      for (int search = 0; search < kMaxLineSearch; ++search, trial_step = trial_step * T(0.5)) {
        ++cuts;
        if (!try_pass(axpy(alpha, trial_step, d))) { last_is_best = false; continue; }
        const T n = norm2(r_alpha());
        last_is_best = n < best;
        if (last_is_best) { best = n; best_step = trial_step; }
        if (n <= T(kFullStepReduction) * residual) break;
      }
      if (best_step == T(0.0)) return false;
      if (!last_is_best) pass(axpy(alpha, best_step, d));
      alpha = axpy(alpha, best_step, d);
      residual = best;
      return true;
    };
    auto line_search = [&](const Vec<T, 9>& d, const Vec<T, 9>& r0) {
      const T phi0 = dot(d, r0);
      T reach = T(1.0);
      bool evaluated = try_pass(axpy(alpha, T(1.0), d));
      for (int halving = 0; !evaluated && halving < kMaxLineSearch; ++halving) {
        out.path |= kPathHalved;
        reach = reach * T(0.5);
        ++cuts;
        evaluated = try_pass(axpy(alpha, reach, d));
      }
      if (!evaluated) return false;
      const T full = norm2(r_alpha());
      if (full <= T(kFullStepReduction) * residual ||
          max_abs(r_alpha()) <= T(kLocalTolerance) * max_coeff(sums.scale) + T(kLocalRoundOff) * max_coeff(sums.floor)) {
        alpha = axpy(alpha, reach, d);
        residual = full;
        return true;
      }
      const T phi1 = dot(d, r_alpha());
      if (!(phi0 < T(0.0)) || !(phi1 > T(0.0)) || !finite_(phi1)) return search_norm(d, T(0.5) * reach);
      T s0 = T(0.0), f0 = phi0, s1 = reach, f1 = phi1, step = reach;
      int side = 0;
      out.path |= kPathRegulaFalsi;
      for (int search = 0; search < kMaxLineSearch; ++search) {
        const T width = s1 - s0;
        step = clamp_((s0 * f1 - s1 * f0) / (f1 - f0), s0 + T(0.1) * width, s1 - T(0.1) * width);
        evaluated = try_pass(axpy(alpha, step, d));
        ++cuts;
        const T f = evaluated ? dot(d, r_alpha()) : T(std::numeric_limits<double>::infinity());
        if (evaluated && abs_(f) <= T(kLineSearchRatio) * abs_(phi0)) break;
        if (!evaluated || f > T(0.0)) {
          s1 = step;
          f1 = evaluated ? f : T(0.5) * f1;
          if (side == 1) f0 = f0 * T(0.5);
          side = 1;
        } else {
          s0 = step;
          f0 = f;
          if (side == -1) f1 = f1 * T(0.5);
          side = -1;
        }
      }
      if (!evaluated) {
        if (s0 == T(0.0)) return false;
        step = s0;
        pass(axpy(alpha, step, d));
      }
      alpha = axpy(alpha, step, d);
      residual = norm2(r_alpha());
      return true;
    };
    for (;;) {
      const T largest = max_abs(r_alpha());
      const T scale = max_coeff(sums.scale);
      const T floor = T(kLocalRoundOff) * max_coeff(sums.floor);
      lu.compute(sums.kaa);
      if (!lu.invertible()) {
        std::ostringstream os;
        os << "Hex8 element " << e_
           << ": the stiffness of the incompatible modes is singular (an unstable element "
              "state, e.g. under large compression); the step is cut";
        throw SolverError(os.str());
      }
      if (largest <= T(kLocalTolerance) * scale + floor || largest == T(0.0)) break;
      if (iterations == kMaxLocalIterations || !finite_(largest)) {
        std::ostringstream os;
        os << "Hex8 element " << e_ << ": the incompatible modes did not converge in "
           << kMaxLocalIterations << " local iterations (residual " << val(largest) << " N of "
           << val(scale) << " N); the step is cut";
        throw SolverError(os.str());
      }
      ++iterations;
      const Vec<T, 9> r0 = r_alpha();
      const Mat<T, 9, 9> k0 = sums.kaa;
      Vec<T, 9> newton = lu.solve(r0);
      for (int i = 0; i < 9; ++i) newton(i) = -newton(i);
      Mat<T, 9, 9> sym;
      for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) sym(i, j) = T(0.5) * (k0(i, j) + k0(j, i));
      LLT<T, 9> llt;
      llt.compute(sym);
      if (llt.ok) {
        if (line_search(newton, r0)) continue;
      } else {
        out.path |= kPathIndefinite;
        if (try_pass(axpy(alpha, T(1.0), newton))) {
          const T full = norm2(r_alpha());
          if (full <= T(kFullStepReduction) * residual) {
            alpha = axpy(alpha, T(1.0), newton);
            residual = full;
            continue;
          }
        }
        out.path |= kPathShifted;
        Vec<T, 9> ev;
        Mat<T, 9, 9> evec;
        jacobi_eigen<T, 9>(sym, ev, evec, false);
        T top = abs_(ev(0));
        for (int i = 1; i < 9; ++i) top = max_(top, abs_(ev(i)));
        const T mu = T(kLocalShift) * top - ev(0);
        Mat<T, 9, 9> shifted_k = k0;
        for (int i = 0; i < 9; ++i) shifted_k(i, i) = shifted_k(i, i) + mu;
        PartialPivLU<T, 9> plu;
        plu.compute(shifted_k);
        Vec<T, 9> shifted = plu.solve(r0);
        for (int i = 0; i < 9; ++i) shifted(i) = -shifted(i);
        if (line_search(shifted, r0) || search_norm(newton, T(0.5))) continue;
      }
      out.path |= kPathMarquardt;
      const Mat<T, 9, 9> normal = mul_tn(k0, k0);
      const Vec<T, 9> gradient = mul_tn(k0, r0);
      T dmax = normal(0, 0);
      for (int i = 1; i < 9; ++i) dmax = max_(dmax, normal(i, i));
      T mu = T(kLocalMarquardt) * dmax;
      bool moved = false;
      for (int search = 0; search < kMaxLineSearch && !moved; ++search, mu = mu * T(10.0)) {
        Mat<T, 9, 9> a = normal;
        for (int i = 0; i < 9; ++i) a(i, i) = a(i, i) + mu;
        LLT<T, 9> l2;
        l2.compute(a);
        Vec<T, 9> d = l2.solve(gradient);
        for (int i = 0; i < 9; ++i) d(i) = -d(i);
        ++cuts;
        if (!try_pass(axpy(alpha, T(1.0), d))) continue;
        const T n = norm2(r_alpha());
        if (n < residual) {
          alpha = axpy(alpha, T(1.0), d);
          residual = n;
          moved = true;
        }
      }
      if (!moved) {
        std::ostringstream os;
        os << "Hex8 element " << e_ << ": no step of the incompatible modes lowers their residual ("
           << val(largest) << " N of " << val(scale)
           << " N) without inverting a point (an unstable or non-smooth element state); the step is cut";
        throw SolverError(os.str());
      }
    }
    flop_mark() = flop_counter();
    accumulate_coupling(want_tangent, sums);
    const Vec<T, 9> correction = lu.solve(r_alpha());
    for (int i = 0; i < 24; ++i) {
      T s = sums.kua(i, 0) * correction(0);
      for (int m = 1; m < 9; ++m) s = s + sums.kua(i, m) * correction(m);
      out.force(i) = sums.r(i) - s;
    }
    out.energy = sums.energy;
    if (want_tangent) {
      Mat<T, 9, 24> coupling;
      for (int j = 0; j < 24; ++j) {
        Vec<T, 9> col;
        for (int m = 0; m < 9; ++m) col(m) = symmetric_ ? sums.kua(j, m) : sums.kau(m, j);
        const Vec<T, 9> x = lu.solve(col);
        for (int m = 0; m < 9; ++m) coupling(m, j) = x(m);
      }
      Mat<T, 24, 24> k;
      for (int i = 0; i < 24; ++i)
        for (int j = 0; j < 24; ++j) {
          T s = sums.kua(i, 0) * coupling(0, j);
          for (int m = 1; m < 9; ++m) s = s + sums.kua(i, m) * coupling(m, j);
          k(i, j) = sums.kuu(i, j) - s;
        }
      if (symmetric_) {
        Mat<T, 24, 24> ks;
        for (int i = 0; i < 24; ++i)
          for (int j = 0; j < 24; ++j) ks(i, j) = T(0.5) * (k(i, j) + k(j, i));
        k = ks;
      }
      out.tangent = k;
    } else {
      out.tangent = Mat<T, 24, 24>::zero();
    }
    out.internal = alpha;
    out.internal_iterations = iterations;
    out.symmetric = symmetric_;
    out.states.reserve(points_.size());
    for (const PointWork<T>& p : points_) {
      out.states.push_back(p.state);
      if (p.yielding) ++out.yielding_points;
    }
    return out;
  }

  // The standard (compatible) Hex8 force on the same kinematics: the modes
  // held at zero. Test helper for the homogeneous-deformation invariant.
  Vec<T, 24> standard_force(const std::vector<PlasticState<T>>& committed) {
    kinematics_at(Vec<T, 9>::zero());
    respond(committed);
    Sums<T> sums;
    accumulate_local(sums);
    Vec<T, 24> f;
    for (int i = 0; i < 24; ++i) f(i) = sums.r(i);
    return f;
  }

  // IncompatibleModes.cpp:602-664 (incompatible_elastoplastic_stress): the
  // element's stress measures at its committed history and modes, after
  // set_up. Each point's return runs again from the committed state (an
  // elastic check, without its tangent); the measures are those of F = I + H.
  ElementStress<T> stress(const std::vector<PlasticState<T>>& committed, const T* internal) {
    if (committed.size() != points_.size())
      throw SolverError("elastoplastic element: the stored internal variables do not match the integration rule");
    Vec<T, 9> alpha = Vec<T, 9>::zero();
    if (internal != nullptr)
      for (int i = 0; i < 9; ++i) alpha(i) = internal[i];
    kinematics_at(alpha);
    const bool finite = kin_ == kFinite, logarithmic = kin_ == kFiniteLogarithmic;
    ElementStress<T> out;
    out.min_jacobian = T(std::numeric_limits<double>::infinity());
    for (std::size_t q = 0; q < points_.size(); ++q) {
      const PointWork<T>& p = points_[q];
      const PlasticResponse<T> r = plastic_return(m_, p.strain, committed[q], false);
      Vec<T, 6> cauchy = r.stress, pk = r.stress;
      Mat<T, 3, 3> sym, f;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          sym(i, j) = T(0.5) * (p.h(i, j) + p.h(j, i));
          f(i, j) = (i == j ? T(1.0) : T(0.0)) + p.h(i, j);
        }
      if (logarithmic) {
        pk = logarithmic_stress(p.log, r.stress);
        const Vec<T, 6> tau = voigt_of(mul(mul(f, tensor_of(pk, 1.0)), transpose(f)), 1.0);
        Vec<T, 6> strain = p.strain;
        for (int i = 3; i < 6; ++i) strain(i) = strain(i) * T(0.5);  // tensor components
        const T j = exp_(strain(0) + strain(1) + strain(2));
        for (int i = 0; i < 6; ++i) {
          cauchy(i) = tau(i) / j;
          out.kirchhoff(i) = out.kirchhoff(i) + tau(i);
          out.logarithmic_strain(i) = out.logarithmic_strain(i) + strain(i);
        }
        T m = abs_(strain(0));
        for (int i = 1; i < 6; ++i) m = max_(m, abs_(strain(i)));
        out.max_strain = max_(out.max_strain, m);
        out.min_jacobian = min_(out.min_jacobian, det3(f));
      } else if (finite) {
        const T j = det3(f);
        const Mat<T, 3, 3> sf = mul(mul(f, tensor_of(r.stress, 1.0)), transpose(f));
        Mat<T, 3, 3> sj;
        for (int i = 0; i < 9; ++i) sj.a[i] = sf.a[i] / j;
        cauchy = voigt_of(sj, 1.0);
        T m = T(0.0);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            const T hth = p.h(0, a) * p.h(0, b) + p.h(1, a) * p.h(1, b) + p.h(2, a) * p.h(2, b);
            m = max_(m, abs_(sym(a, b) + T(0.5) * hth));
          }
        out.max_strain = max_(out.max_strain, m);
        out.min_jacobian = min_(out.min_jacobian, j);
      } else {
        T m = T(0.0);
        for (int i = 0; i < 9; ++i) m = max_(m, abs_(sym.a[i]));
        out.max_strain = max_(out.max_strain, m);
      }
      for (int i = 0; i < 6; ++i) {
        out.cauchy(i) = out.cauchy(i) + cauchy(i);
        out.piola_kirchhoff(i) = out.piola_kirchhoff(i) + pk(i);
      }
      out.max_equivalent_plastic_strain = max_(out.max_equivalent_plastic_strain, r.state.abar);
    }
    const T count = T(static_cast<double>(points_.size()));
    for (int i = 0; i < 6; ++i) {
      out.cauchy(i) = out.cauchy(i) / count;
      out.piola_kirchhoff(i) = out.piola_kirchhoff(i) / count;
      out.kirchhoff(i) = out.kirchhoff(i) / count;
      out.logarithmic_strain(i) = out.logarithmic_strain(i) / count;
    }
    out.von_mises = von_mises_stress(out.cauchy);
    if (!finite && !logarithmic) out.min_jacobian = T(1.0);
    return out;
  }

 private:
  const MaterialData& m_;
  const Rule& rule_;
  int kin_;
  bool symmetric_;
  long e_ = 0;
  Mat<T, 3, 8> u_;
  Vec<T, 24> ue_;
  std::vector<PointWork<T>> points_;
};

// Hex8Incompatible.cpp:94-127 (with Hex8Kernels.hpp:60-120): the linear
// condensed stiffness K* of an elastic isotropic material.
inline void linear_incompatible_stiffness(const double* x0, double E, double nu, const Rule& rule,
                                          double* k_out, bool* ok) {
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double g = E / (2.0 * (1.0 + nu));
  double d[6][6] = {{0}};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) d[i][j] = lam;
    d[i][i] = lam + 2.0 * g;
    d[3 + i][3 + i] = g;
  }
  auto mapping = [&](double xi, double eta, double zeta, double dn_dx[8][3], double& det,
                     double jinv[3][3]) {
    double dn[8][3];
    shape_gradients_natural(xi, eta, zeta, dn);
    double jac[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int a = 0; a < 8; ++a) s += x0[a * 3 + i] * dn[a][j];
        jac[i][j] = s;
      }
    const double c00 = jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1];
    const double c01 = jac[1][2] * jac[2][0] - jac[1][0] * jac[2][2];
    const double c02 = jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0];
    det = jac[0][0] * c00 + jac[0][1] * c01 + jac[0][2] * c02;
    if (!(det > 0.0)) throw MeshError("Hex8 Jacobian determinant is not positive");
    jinv[0][0] = c00 / det;
    jinv[1][0] = c01 / det;
    jinv[2][0] = c02 / det;
    jinv[0][1] = (jac[0][2] * jac[2][1] - jac[0][1] * jac[2][2]) / det;
    jinv[1][1] = (jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0]) / det;
    jinv[2][1] = (jac[0][1] * jac[2][0] - jac[0][0] * jac[2][1]) / det;
    jinv[0][2] = (jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1]) / det;
    jinv[1][2] = (jac[0][2] * jac[1][0] - jac[0][0] * jac[1][2]) / det;
    jinv[2][2] = (jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0]) / det;
    for (int a = 0; a < 8; ++a)
      for (int j = 0; j < 3; ++j) dn_dx[a][j] = dn[a][0] * jinv[0][j] + dn[a][1] * jinv[1][j] + dn[a][2] * jinv[2][j];
  };
  double dn0[8][3], det0, jinv0[3][3];
  mapping(0.0, 0.0, 0.0, dn0, det0, jinv0);
  double lkuu[24][24] = {{0}}, lkau[9][24] = {{0}}, lkaa[9][9] = {{0}};
  for (int q = 0; q < rule.n; ++q) {
    double dn_dx[8][3], det, jinv[3][3];
    mapping(rule.xi[q], rule.eta[q], rule.zeta[q], dn_dx, det, jinv);
    double b[6][24] = {{0}}, bt[6][9] = {{0}};
    for (int a = 0; a < 8; ++a) {
      const double dx = dn_dx[a][0], dy = dn_dx[a][1], dz = dn_dx[a][2];
      const int c = 3 * a;
      b[0][c] = dx; b[1][c + 1] = dy; b[2][c + 2] = dz;
      b[3][c] = dy; b[3][c + 1] = dx; b[4][c + 1] = dz; b[4][c + 2] = dy; b[5][c] = dz; b[5][c + 2] = dx;
    }
    const double ratio = det0 / det;
    const double natural[3] = {rule.xi[q], rule.eta[q], rule.zeta[q]};
    double gm[3][3];
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) gm[i][k] = (ratio * -2.0 * natural[k]) * jinv0[k][i];
    for (int k = 0; k < 3; ++k) {
      const double dx = gm[0][k], dy = gm[1][k], dz = gm[2][k];
      const int c = 3 * k;
      bt[0][c] = dx; bt[1][c + 1] = dy; bt[2][c + 2] = dz;
      bt[3][c] = dy; bt[3][c + 1] = dx; bt[4][c + 1] = dz; bt[4][c + 2] = dy; bt[5][c] = dz; bt[5][c + 2] = dx;
    }
    const double w = det * rule.w[q];
    double db[6][24], dbt[6][9];
    for (int i = 0; i < 6; ++i) {
      for (int j = 0; j < 24; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += d[i][k] * b[k][j];
        db[i][j] = s;
      }
      for (int j = 0; j < 9; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += d[i][k] * bt[k][j];
        dbt[i][j] = s;
      }
    }
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += b[k][i] * db[k][j];
        lkuu[i][j] += w * s;
      }
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 24; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += bt[k][i] * db[k][j];
        lkau[i][j] += w * s;
      }
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += bt[k][i] * dbt[k][j];
        lkaa[i][j] += w * s;
      }
  }
  Mat<double, 9, 9> sym;
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) sym(i, j) = 0.5 * (lkaa[i][j] + lkaa[j][i]);
  LLT<double, 9> llt;
  llt.compute(sym);
  *ok = llt.ok;
  if (!llt.ok) return;
  double coupling[9][24];
  for (int j = 0; j < 24; ++j) {
    Vec<double, 9> col;
    for (int i = 0; i < 9; ++i) col(i) = lkau[i][j];
    const Vec<double, 9> x = llt.solve(col);
    for (int i = 0; i < 9; ++i) coupling[i][j] = x(i);
  }
  double ke[24][24];
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      double s = 0.0;
      for (int m = 0; m < 9; ++m) s += lkau[m][i] * coupling[m][j];
      ke[i][j] = lkuu[i][j] - s;
    }
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) k_out[i * 24 + j] = 0.5 * (ke[i][j] + ke[j][i]);
}

}  // namespace oracle
