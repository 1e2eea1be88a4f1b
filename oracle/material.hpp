// Oracle (test infrastructure only): the material point of the reference.
//
//   * hardening laws and the Hill48 yield matrix:
//       src/material/Material.cpp:73-109 (yield, yield_slope, isotropic_energy,
//       von_mises_yield_matrix, symmetric_tangent), :113-217 (derive_hill)
//   * logarithmic strain and its derivatives:
//       src/material/LogarithmicStrain.cpp:29-197
//   * returns: src/material/Plasticity.cpp:18-171 (J2 radial return),
//       :185-471 (general closest-point return: Hill48 + Chaboche),
//       :513-539 (plastic_return dispatch, 3-D only here)
//
// Restated literally (same operation order and constants). Where the
// reference calls Eigen (3x3 symmetric eigensolver, 7x7 partial-pivot LU)
// the algorithms of om.hpp stand in; their results differ from Eigen's only
// by round-off.
#pragma once

#include "om.hpp"

#include <algorithm>
#include <sstream>

namespace oracle {

constexpr int kMaxBackstresses = 4;  // include/sparlab/material/Material.hpp:68

// PlasticityParameters + Material of include/sparlab/material/Material.hpp:148-257,
// in plain data.
struct MaterialData {
  double E = 0.0, nu = 0.0;
  double sy0 = 0.0, H = 0.0, Hkin = 0.0, Q = 0.0, delta = 0.0;
  int criterion = 0;  // 0 von Mises, 1 Hill48
  int nback = 0;
  double C[kMaxBackstresses] = {0, 0, 0, 0};
  double gamma[kMaxBackstresses] = {0, 0, 0, 0};
  int integration = 0;  // 0 exponential, 1 backward Euler
  double P[36];         // yield matrix, global axes, tensorial Voigt
  double shear() const { return E / (2.0 * (1.0 + nu)); }
  double lame() const { return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)); }
  bool enabled() const { return sy0 > 0.0; }
  bool general() const { return criterion == 1 || nback > 0; }
  bool symmetric_tangent() const {
    for (int i = 0; i < nback; ++i)
      if (gamma[i] > 0.0) return false;
    return true;
  }
  int kinematic_terms() const { return nback + (Hkin > 0.0 ? 1 : 0); }
  void term(int i, double& modulus, double& recovery) const {
    if (i < nback) { modulus = C[i]; recovery = gamma[i]; }
    else { modulus = Hkin; recovery = 0.0; }
  }
};

// Material.cpp:73-90
template <class T> T yield(const MaterialData& m, T a) {
  if (g_plant == kPlantVoceYield) return T(m.sy0) + T(m.H) * a + T(m.Q) * -expm1_(T(-1.01 * m.delta) * a);
  return T(m.sy0) + T(m.H) * a + T(m.Q) * -expm1_(T(-m.delta) * a);
}
template <class T> T yield_slope(const MaterialData& m, T a) {
  return T(m.H) + T(m.Q) * T(m.delta) * exp_(T(-m.delta) * a);
}
template <class T> T isotropic_energy(const MaterialData& m, T a) {
  T out = T(0.5) * T(m.H) * a * a;
  if (m.Q != 0.0) out = out + T(m.Q) * (a + expm1_(T(-m.delta) * a) / T(m.delta));
  if (g_plant == kPlantVoceEnergySign && m.Q != 0.0) out = out - T(2.0 * m.Q) * (a + expm1_(T(-m.delta) * a) / T(m.delta));
  return out;
}

// Material.cpp:92-102
inline void von_mises_yield_matrix(double* p) {
  for (int i = 0; i < 36; ++i) p[i] = 0.0;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) p[i * 6 + j] = -0.5;
    p[i * 6 + i] = 1.0;
    p[(3 + i) * 6 + 3 + i] = 3.0;
  }
}

struct HillInput {
  int calibration = 0;  // 0 r-values, 1 stress ratios, 2 coefficients
  double r0 = 1, r45 = 1, r90 = 1;
  double sigma45 = 1, sigma90 = 1, sigma_b = 1;
  double shear_l = 1.5, shear_m = 1.5;
  double F = 0.5, G = 0.5, Hh = 0.5, L = 1.5, M = 1.5, N = 1.5;
  double rd[3] = {1, 0, 0};
  double nd[3] = {0, 0, 1};
};

// Material.cpp:135-217: the coefficients, the material frame and the yield
// matrix P = T^T P_m T in global axes.
inline void derive_hill(HillInput& h, double* yield_matrix) {
  auto pos = [](double v) { return std::isfinite(v) && v > 0.0; };
  if (h.calibration == 0) {
    if (!pos(h.r0) || !pos(h.r45) || !pos(h.r90))
      throw ConfigError("the Hill48 r-values must be finite and positive");
    h.G = 1.0 / (1.0 + h.r0);
    h.Hh = h.r0 / (1.0 + h.r0);
    h.F = h.r0 / (h.r90 * (1.0 + h.r0));
    h.N = (h.r0 + h.r90) * (1.0 + 2.0 * h.r45) / (2.0 * h.r90 * (1.0 + h.r0));
    h.L = h.shear_l;
    h.M = h.shear_m;
  } else if (h.calibration == 1) {
    if (!pos(h.sigma45) || !pos(h.sigma90) || !pos(h.sigma_b))
      throw ConfigError("the Hill48 yield stress ratios must be finite and positive");
    const double a = 1.0 / (h.sigma90 * h.sigma90);
    const double b = 1.0 / (h.sigma_b * h.sigma_b);
    h.F = 0.5 * (a + b - 1.0);
    h.G = 0.5 * (1.0 - a + b);
    h.Hh = 0.5 * (1.0 + a - b);
    h.N = 0.5 * (4.0 / (h.sigma45 * h.sigma45) - h.F - h.G);
    h.L = h.shear_l;
    h.M = h.shear_m;
  }
  if (!(h.L > 0.0 && h.M > 0.0 && h.N > 0.0) || !(h.F + h.Hh > 0.0) ||
      !(h.F * h.G + h.G * h.Hh + h.Hh * h.F > 0.0))
    throw ConfigError("the Hill48 coefficients do not make a convex yield surface");
  double nd[3], rd[3];
  const double nn = std::sqrt(h.nd[0] * h.nd[0] + h.nd[1] * h.nd[1] + h.nd[2] * h.nd[2]);
  const double rn = std::sqrt(h.rd[0] * h.rd[0] + h.rd[1] * h.rd[1] + h.rd[2] * h.rd[2]);
  if (!(nn > 0.0) || !(rn > 0.0)) throw ConfigError("the rolling direction and the sheet normal must be non-zero");
  for (int i = 0; i < 3; ++i) { nd[i] = h.nd[i] / nn; rd[i] = h.rd[i] / rn; }
  const double c = rd[0] * nd[0] + rd[1] * nd[1] + rd[2] * nd[2];
  for (int i = 0; i < 3; ++i) rd[i] -= c * nd[i];
  const double r2 = std::sqrt(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
  if (!(r2 > 1.0e-6)) throw ConfigError("the rolling direction is parallel to the sheet normal");
  for (int i = 0; i < 3; ++i) rd[i] /= r2;
  const double td[3] = {nd[1] * rd[2] - nd[2] * rd[1], nd[2] * rd[0] - nd[0] * rd[2],
                        nd[0] * rd[1] - nd[1] * rd[0]};
  double q[3][3];
  for (int i = 0; i < 3; ++i) { q[0][i] = rd[i]; q[1][i] = td[i]; q[2][i] = nd[i]; }
  double pm[36] = {0};
  pm[0 * 6 + 0] = h.G + h.Hh;
  pm[1 * 6 + 1] = h.F + h.Hh;
  pm[2 * 6 + 2] = h.F + h.G;
  pm[0 * 6 + 1] = pm[1 * 6 + 0] = -h.Hh;
  pm[1 * 6 + 2] = pm[2 * 6 + 1] = -h.F;
  pm[0 * 6 + 2] = pm[2 * 6 + 0] = -h.G;
  pm[3 * 6 + 3] = 2.0 * h.N;
  pm[4 * 6 + 4] = 2.0 * h.L;
  pm[5 * 6 + 5] = 2.0 * h.M;
  // stress_rotation (Material.cpp:117-129)
  static const int ii[6] = {0, 1, 2, 0, 1, 2};
  static const int jj[6] = {0, 1, 2, 1, 2, 0};
  double t[36];
  for (int k = 0; k < 6; ++k) {
    double e[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    e[ii[k]][jj[k]] = 1.0;
    e[jj[k]][ii[k]] = 1.0;
    double qe[3][3], r[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double s = 0.0;
        for (int c2 = 0; c2 < 3; ++c2) s += q[a][c2] * e[c2][b];
        qe[a][b] = s;
      }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double s = 0.0;
        for (int c2 = 0; c2 < 3; ++c2) s += qe[a][c2] * q[b][c2];
        r[a][b] = s;
      }
    for (int row = 0; row < 6; ++row) t[row * 6 + k] = r[ii[row]][jj[row]];
  }
  double tp[36];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += t[k * 6 + i] * pm[k * 6 + j];
      tp[i * 6 + j] = s;
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += tp[i * 6 + k] * t[k * 6 + j];
      yield_matrix[i * 6 + j] = s;
    }
}

// ---------------------------------------------------------------------------
// PlasticState / PlasticResponse (include/sparlab/material/Plasticity.hpp:185-214)
// ---------------------------------------------------------------------------
template <class T>
struct PlasticState {
  Vec<T, 6> ep = Vec<T, 6>::zero();    // engineering shears
  Vec<T, 6> beta = Vec<T, 6>::zero();  // tensorial
  Vec<T, 6> alpha[kMaxBackstresses] = {Vec<T, 6>::zero(), Vec<T, 6>::zero(), Vec<T, 6>::zero(),
                                       Vec<T, 6>::zero()};
  T abar = T(0.0);
  T thickness_strain = T(0.0);
  bool loading = false;
};

template <class T>
struct PlasticResponse {
  Vec<T, 6> stress = Vec<T, 6>::zero();
  Mat<T, 6, 6> tangent = Mat<T, 6, 6>::zero();
  PlasticState<T> state;
  bool yielding = false;
  T energy = T(0.0);
  T strain_33 = T(0.0);
  bool symmetric = true;
};

// Plasticity.cpp:18-63
template <class T> T tensor_norm(const Vec<T, 6>& v) {
  return sqrt_(v(0) * v(0) + v(1) * v(1) + v(2) * v(2) +
               T(2.0) * (v(3) * v(3) + v(4) * v(4) + v(5) * v(5)));
}
template <class T> Vec<T, 6> deviatoric_stress(T g, const Vec<T, 6>& strain) {
  const T mean = (strain(0) + strain(1) + strain(2)) / T(3.0);
  Vec<T, 6> s;
  for (int i = 0; i < 3; ++i) s(i) = T(2.0) * g * (strain(i) - mean);
  for (int i = 3; i < 6; ++i) s(i) = g * strain(i);
  return s;
}
template <class T> T deviatoric_norm2(const Vec<T, 6>& strain) {
  const T mean = (strain(0) + strain(1) + strain(2)) / T(3.0);
  T out = T(0.0);
  for (int i = 0; i < 3; ++i) out = out + (strain(i) - mean) * (strain(i) - mean);
  for (int i = 3; i < 6; ++i) out = out + T(0.5) * strain(i) * strain(i);
  return out;
}
template <class T> T strain_norm2(const Vec<T, 6>& strain) {
  T out = T(0.0);
  for (int i = 0; i < 3; ++i) out = out + strain(i) * strain(i);
  for (int i = 3; i < 6; ++i) out = out + T(0.5) * strain(i) * strain(i);
  return out;
}
template <class T> Mat<T, 6, 6> elastic_tangent(T k, T g) {
  Mat<T, 6, 6> c = Mat<T, 6, 6>::zero();
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) c(i, j) = k - T(2.0) * g / T(3.0);
    c(i, i) = k + T(4.0) * g / T(3.0);
    c(3 + i, 3 + i) = g;
  }
  return c;
}

// Plasticity.cpp:66-171: the J2 radial return with linear/Voce isotropic and
// Prager kinematic hardening.
template <class T>
PlasticResponse<T> return_3d(const MaterialData& m, const Vec<T, 6>& strain,
                             const PlasticState<T>& committed) {
  const T g = T(m.shear());
  const T k = T(m.lame()) + T(2.0) * g / T(3.0);
  const T hk = T(m.Hkin);
  const T sqrt23 = T(std::sqrt(2.0 / 3.0));
  Vec<T, 6> elastic;
  for (int i = 0; i < 6; ++i) elastic(i) = strain(i) - committed.ep(i);
  const T volumetric = elastic(0) + elastic(1) + elastic(2);
  const Vec<T, 6> s_trial = deviatoric_stress(g, elastic);
  Vec<T, 6> xi_trial;
  for (int i = 0; i < 6; ++i) xi_trial(i) = s_trial(i) - committed.beta(i);
  const T xi_norm = tensor_norm(xi_trial);
  const T alpha_n = committed.abar;
  const T radius_n = sqrt23 * yield(m, alpha_n);
  const T f_trial = xi_norm - radius_n;

  PlasticResponse<T> out;
  out.state = committed;
  out.state.loading = false;
  out.strain_33 = strain(2);
  if (!m.enabled() || f_trial <= T(1.0e-12) * radius_n) {
    out.stress = s_trial;
    for (int i = 0; i < 3; ++i) out.stress(i) = out.stress(i) + k * volumetric;
    out.tangent = elastic_tangent(k, g);
    out.energy = T(0.5) * k * volumetric * volumetric + g * deviatoric_norm2(elastic) +
                 isotropic_energy(m, alpha_n) + hk / T(3.0) * strain_norm2(committed.ep);
    if (m.enabled() && committed.loading && f_trial > T(-1.0e-10) * radius_n && xi_norm > T(0.0)) {
      Vec<T, 6> n;
      for (int i = 0; i < 6; ++i) n(i) = xi_trial(i) / xi_norm;
      const T factor = T(2.0) * g / (T(1.0) + (yield_slope(m, alpha_n) + hk) / (T(3.0) * g));
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) out.tangent(i, j) = out.tangent(i, j) - factor * (n(i) * n(j));
      out.state.loading = true;
    }
    return out;
  }
  const T stiffness = T(2.0) * g + T(2.0) / T(3.0) * hk;
  T dgamma = f_trial / (stiffness + T(2.0) / T(3.0) * yield_slope(m, alpha_n));
  if (m.Q != 0.0) {
    dgamma = T(0.0);
    bool converged = false;
    for (int it = 0; it < 60; ++it) {
      const T a = alpha_n + sqrt23 * dgamma;
      const T residual = xi_norm - stiffness * dgamma - sqrt23 * yield(m, a);
      if (abs_(residual) <= T(1.0e-13) * radius_n) { converged = true; break; }
      dgamma = dgamma + residual / (stiffness + T(2.0) / T(3.0) * yield_slope(m, a));
    }
    if (!converged) {
      const T a = alpha_n + sqrt23 * dgamma;
      const T residual = xi_norm - stiffness * dgamma - sqrt23 * yield(m, a);
      converged = abs_(residual) <= T(8.0 * kEps) * xi_norm;
    }
    if (!converged) throw SolverError("the J2 return did not converge for the plastic multiplier");
  }
  const T alpha = alpha_n + sqrt23 * dgamma;
  Vec<T, 6> n;
  for (int i = 0; i < 6; ++i) n(i) = xi_trial(i) / xi_norm;
  out.yielding = true;
  for (int i = 0; i < 6; ++i) out.stress(i) = s_trial(i) - T(2.0) * g * dgamma * n(i);
  for (int i = 0; i < 3; ++i) out.stress(i) = out.stress(i) + k * volumetric;
  PlasticState<T>& st = out.state;
  st.loading = true;
  st.abar = alpha;
  for (int i = 0; i < 6; ++i) st.beta(i) = st.beta(i) + T(2.0) / T(3.0) * hk * dgamma * n(i);
  for (int i = 0; i < 3; ++i) st.ep(i) = st.ep(i) + dgamma * n(i);
  for (int i = 3; i < 6; ++i) st.ep(i) = st.ep(i) + T(2.0) * dgamma * n(i);
  const T theta = T(1.0) - T(2.0) * g * dgamma / xi_norm;
  const T theta_bar =
      T(1.0) / (T(1.0) + (yield_slope(m, alpha) + hk) / (T(3.0) * g)) - (T(1.0) - theta);
  Mat<T, 6, 6> c = Mat<T, 6, 6>::zero();
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) c(i, j) = k - T(2.0) * g * theta / T(3.0);
    c(i, i) = k + T(4.0) * g * theta / T(3.0);
    c(3 + i, 3 + i) = g * theta;
  }
  const T f2 = T(2.0) * g * theta_bar;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) c(i, j) = c(i, j) - f2 * (n(i) * n(j));
  out.tangent = c;
  Vec<T, 6> elastic_new = elastic;
  for (int i = 0; i < 3; ++i) elastic_new(i) = elastic_new(i) - dgamma * n(i);
  for (int i = 3; i < 6; ++i) elastic_new(i) = elastic_new(i) - T(2.0) * dgamma * n(i);
  out.energy = T(0.5) * k * volumetric * volumetric + g * deviatoric_norm2(elastic_new) +
               isotropic_energy(m, alpha) + hk / T(3.0) * strain_norm2(st.ep);
  return out;
}

// Plasticity.cpp:185-212
template <class T> struct Decay {
  T theta = T(1.0), chi = T(0.0), dtheta = T(0.0), dchi = T(1.0);
};
template <class T> Decay<T> decay(double gamma, T dl, int integration) {
  Decay<T> d;
  if (gamma == 0.0) { d.chi = dl; return d; }
  if (integration == 0) {
    d.theta = exp_(T(-gamma) * dl);
    d.chi = -expm1_(T(-gamma) * dl) / T(gamma);
    d.dtheta = T(-gamma) * d.theta;
    d.dchi = d.theta;
  } else {
    d.theta = T(1.0) / (T(1.0) + T(gamma) * dl);
    d.chi = d.theta * dl;
    d.dtheta = T(-gamma) * d.theta * d.theta;
    d.dchi = d.theta * d.theta;
    if (g_plant == kPlantBackwardEulerChi) { d.chi = dl; d.dchi = T(1.0); }
  }
  return d;
}

// Plasticity.cpp:235-244
template <class T> struct GeneralPoint {
  T equivalent = T(0.0);
  Vec<T, 6> m, mt;
  T b = T(0.0), db = T(0.0);
  Vec<T, 6> back_rate;
  Decay<T> decays[kMaxBackstresses];
  Vec<T, 7> residual;
};

// Plasticity.cpp:246-323
template <class T>
struct GeneralReturn {
  const MaterialData& p;
  T g;
  const PlasticState<T>& committed;
  const Vec<T, 6>& s_trial;
  int count;
  double mod[kMaxBackstresses], rec[kMaxBackstresses];
  GeneralReturn(const MaterialData& m, const PlasticState<T>& c, const Vec<T, 6>& st)
      : p(m), g(T(m.shear())), committed(c), s_trial(st), count(m.kinematic_terms()) {
    for (int i = 0; i < count; ++i) m.term(i, mod[i], rec[i]);
  }
  Vec<T, 6> pxi(const Vec<T, 6>& xi) const {
    Vec<T, 6> out;
    for (int i = 0; i < 6; ++i) {
      T s = T(p.P[i * 6]) * xi(0);
      for (int j = 1; j < 6; ++j) s = s + T(p.P[i * 6 + j]) * xi(j);
      out(i) = s;
    }
    return out;
  }
  bool evaluate(const Vec<T, 6>& xi, T dl, GeneralPoint<T>& e) const {
    const Vec<T, 6> px = pxi(xi);
    const T q = dot(xi, px);
    if (!(q > T(0.0)) || !finite_(q)) return false;
    e.equivalent = sqrt_(q);
    for (int i = 0; i < 6; ++i) e.m(i) = px(i) / e.equivalent;
    e.mt = e.m;
    for (int i = 3; i < 6; ++i) e.mt(i) = e.mt(i) * T(0.5);
    Vec<T, 6> xi_tilde = s_trial;
    e.b = T(2.0) * g * dl;
    e.db = T(2.0) * g;
    e.back_rate = Vec<T, 6>::zero();
    for (int i = 0; i < count; ++i) {
      const Decay<T> d = decay(rec[i], dl, p.integration);
      const Vec<T, 6>& an = committed.alpha[i];
      for (int j = 0; j < 6; ++j) xi_tilde(j) = xi_tilde(j) - d.theta * an(j);
      e.b = e.b + T(2.0) / T(3.0) * T(mod[i]) * d.chi;
      e.db = e.db + T(2.0) / T(3.0) * T(mod[i]) * d.dchi;
      for (int j = 0; j < 6; ++j) e.back_rate(j) = e.back_rate(j) + d.dtheta * an(j);
      e.decays[i] = d;
    }
    for (int j = 0; j < 6; ++j) e.residual(j) = xi(j) - xi_tilde(j) + e.b * e.mt(j);
    e.residual(6) = e.equivalent - yield(p, committed.abar + dl);
    return true;
  }
  Mat<T, 6, 6> flow_curvature(const GeneralPoint<T>& e) const {
    Mat<T, 6, 6> c;
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) c(i, j) = T(p.P[i * 6 + j]) - e.m(i) * e.m(j);
    for (int i = 3; i < 6; ++i)
      for (int j = 0; j < 6; ++j) c(i, j) = c(i, j) * T(0.5);
    return c;
  }
  Mat<T, 7, 7> jacobian(const GeneralPoint<T>& e, T dl) const {
    Mat<T, 7, 7> j;
    const Mat<T, 6, 6> fc = flow_curvature(e);
    const T r = e.b / e.equivalent;
    for (int i = 0; i < 6; ++i)
      for (int k = 0; k < 6; ++k) j(i, k) = (i == k ? T(1.0) : T(0.0)) + r * fc(i, k);
    for (int i = 0; i < 6; ++i) j(i, 6) = e.back_rate(i) + e.db * e.mt(i);
    for (int k = 0; k < 6; ++k) j(6, k) = e.m(k);
    j(6, 6) = -yield_slope(p, committed.abar + dl);
    return j;
  }
  Mat<T, 6, 6> tangent(const GeneralPoint<T>& e, T dl, const Mat<T, 6, 6>& elastic) const {
    Mat<T, 7, 6> rhs = Mat<T, 7, 6>::zero();
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) rhs(i, j) = T(-2.0) * g / T(3.0);
      rhs(i, i) = T(4.0) * g / T(3.0);
      rhs(3 + i, 3 + i) = g;
    }
    PartialPivLU<T, 7> lu;
    lu.compute(jacobian(e, dl));
    const Mat<T, 7, 6> da = lu.template solve<6>(rhs);
    Mat<T, 6, 6> c;
    const T g2 = T(2.0) * g;
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) c(i, j) = elastic(i, j) - g2 * (e.mt(i) * da(6, j));
    if (dl > T(0.0)) {
      const T s = T(2.0) * g * dl / e.equivalent;
      const Mat<T, 6, 6> fc = flow_curvature(e);
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
          T acc = fc(i, 0) * da(0, j);
          for (int k = 1; k < 6; ++k) acc = acc + fc(i, k) * da(k, j);
          c(i, j) = c(i, j) - s * acc;
        }
    }
    return c;
  }
};

// Plasticity.cpp:326-333
template <class T>
T kinematic_energy(const GeneralReturn<T>& r, const PlasticState<T>& st) {
  T out = T(0.0);
  for (int i = 0; i < r.count; ++i) {
    const T nrm = tensor_norm(st.alpha[i]);
    out = out + T(g_plant == kPlantKinematicEnergy ? 0.5 : 0.75) * nrm * nrm / T(r.mod[i]);
  }
  return out;
}

// Plasticity.cpp:337-471
template <class T>
PlasticResponse<T> return_general_3d(const MaterialData& m, const Vec<T, 6>& strain,
                                     const PlasticState<T>& committed, bool want_tangent) {
  const T g = T(m.shear());
  const T k = T(m.lame()) + T(2.0) * g / T(3.0);
  Vec<T, 6> elastic;
  for (int i = 0; i < 6; ++i) elastic(i) = strain(i) - committed.ep(i);
  const T volumetric = elastic(0) + elastic(1) + elastic(2);
  const Vec<T, 6> s_trial = deviatoric_stress(g, elastic);
  const GeneralReturn<T> gr(m, committed, s_trial);
  Vec<T, 6> xi_trial = s_trial;
  for (int i = 0; i < gr.count; ++i)
    for (int j = 0; j < 6; ++j) xi_trial(j) = xi_trial(j) - committed.alpha[i](j);
  const T alpha_n = committed.abar;
  const T sy_n = yield(m, alpha_n);
  const T eq_trial = sqrt_(max_(dot(xi_trial, gr.pxi(xi_trial)), T(0.0)));
  const T f_trial = eq_trial - sy_n;
  const Mat<T, 6, 6> ce = elastic_tangent(k, g);
  if (!finite_(f_trial)) throw SolverError("the closest-point return met a non-finite trial stress");

  PlasticResponse<T> out;
  out.state = committed;
  out.state.loading = false;
  out.strain_33 = strain(2);
  if (!m.enabled() || f_trial <= T(1.0e-12) * sy_n) {
    out.stress = s_trial;
    for (int i = 0; i < 3; ++i) out.stress(i) = out.stress(i) + k * volumetric;
    out.tangent = ce;
    out.energy = T(0.5) * k * volumetric * volumetric + g * deviatoric_norm2(elastic) +
                 isotropic_energy(m, alpha_n) + kinematic_energy(gr, committed);
    if (m.enabled() && committed.loading && f_trial > T(-1.0e-10) * sy_n && eq_trial > T(0.0)) {
      out.state.loading = true;
      GeneralPoint<T> e;
      if (want_tangent && gr.evaluate(xi_trial, T(0.0), e)) out.tangent = gr.tangent(e, T(0.0), ce);
    }
    return out;
  }
  Vec<T, 6> xi = xi_trial;
  T dl = T(0.0);
  GeneralPoint<T> e;
  gr.evaluate(xi, dl, e);
  const T scale = max_(max_(sy_n, max_abs(s_trial)), max_abs(xi_trial));
  const T eps_scale = T(kEps) * scale;
  const T tol = max_(T(1.0e-13) * sy_n, T(32.0) * eps_scale);
  bool converged = false;
  int it = 0;
  for (; it < 50; ++it) {
    const T r_inf = max_abs(e.residual);
    if (r_inf <= tol) { converged = true; break; }
    PartialPivLU<T, 7> lu;
    lu.compute(gr.jacobian(e, dl));
    Vec<T, 7> neg;
    for (int i = 0; i < 7; ++i) neg(i) = -e.residual(i);
    const Vec<T, 7> step = lu.solve(neg);
    const T r0 = norm2(e.residual);
    T t = T(1.0);
    bool accepted = false;
    GeneralPoint<T> trial;
    for (int ls = 0; ls < 40; ++ls) {
      Vec<T, 6> xi_t;
      for (int i = 0; i < 6; ++i) xi_t(i) = xi(i) + t * step(i);
      const T dl_t = max_(dl + t * step(6), T(0.0));
      if (gr.evaluate(xi_t, dl_t, trial) &&
          norm2(trial.residual) <= (T(1.0) - T(1.0e-4) * t) * r0) {
        xi = xi_t;
        dl = dl_t;
        e = trial;
        accepted = true;
        break;
      }
      t = t * T(0.5);
    }
    if (!accepted) { converged = r_inf <= T(4.0) * tol; break; }
  }
  if (!converged) {
    std::ostringstream os;
    os << "the " << (m.criterion == 1 ? "Hill48" : "von Mises")
       << " closest-point return did not converge (residual "
       << val(max_abs(e.residual)) / val(sy_n) << " of the yield stress after " << it
       << " iterations)";
    throw SolverError(os.str());
  }
  if (max_abs(e.residual) > T(4.0) * eps_scale) {
    PartialPivLU<T, 7> lu;
    lu.compute(gr.jacobian(e, dl));
    Vec<T, 7> neg;
    for (int i = 0; i < 7; ++i) neg(i) = -e.residual(i);
    const Vec<T, 7> step = lu.solve(neg);
    Vec<T, 6> xi_p;
    for (int i = 0; i < 6; ++i) xi_p(i) = xi(i) + step(i);
    const T dl_p = max_(dl + step(6), T(0.0));
    GeneralPoint<T> polished;
    if (gr.evaluate(xi_p, dl_p, polished) && max_abs(polished.residual) < max_abs(e.residual)) {
      xi = xi_p;
      dl = dl_p;
      e = polished;
    }
  }
  out.yielding = true;
  out.symmetric = m.symmetric_tangent();
  for (int i = 0; i < 6; ++i) out.stress(i) = s_trial(i) - T(2.0) * g * dl * e.mt(i);
  for (int i = 0; i < 3; ++i) out.stress(i) = out.stress(i) + k * volumetric;
  PlasticState<T>& st = out.state;
  st.loading = true;
  st.abar = alpha_n + dl;
  for (int i = 0; i < 6; ++i) st.ep(i) = st.ep(i) + dl * e.m(i);
  st.beta = Vec<T, 6>::zero();
  for (int i = 0; i < gr.count; ++i) {
    const Decay<T>& d = e.decays[i];
    const T f = T(2.0) / T(3.0) * T(gr.mod[i]) * d.chi;
    for (int j = 0; j < 6; ++j) st.alpha[i](j) = d.theta * committed.alpha[i](j) + f * e.mt(j);
    for (int j = 0; j < 6; ++j) st.beta(j) = st.beta(j) + st.alpha[i](j);
  }
  Vec<T, 6> elastic_new;
  for (int i = 0; i < 6; ++i) elastic_new(i) = elastic(i) - dl * e.m(i);
  out.energy = T(0.5) * k * volumetric * volumetric + g * deviatoric_norm2(elastic_new) +
               isotropic_energy(m, st.abar) + kinematic_energy(gr, st);
  out.tangent = want_tangent ? gr.tangent(e, dl, ce) : ce;
  return out;
}

// Plasticity.cpp:513-539 (three-dimensional stress state)
template <class T>
PlasticResponse<T> plastic_return(const MaterialData& m, const Vec<T, 6>& strain,
                                  const PlasticState<T>& committed, bool want_tangent) {
  if (!m.general()) return return_3d(m, strain, committed);
  return return_general_3d(m, strain, committed, want_tangent);
}

// ---------------------------------------------------------------------------
// Logarithmic strain (LogarithmicStrain.cpp)
// ---------------------------------------------------------------------------
constexpr int kRow[6] = {0, 1, 2, 0, 1, 2};
constexpr int kCol[6] = {0, 1, 2, 1, 2, 0};
constexpr double kSpread = 5.0e-3;

template <class T> struct LogStrain {
  Mat<T, 3, 3> axes = Mat<T, 3, 3>::identity();
  Vec<T, 3> green = Vec<T, 3>::zero();
  Vec<T, 6> strain = Vec<T, 6>::zero();
  Mat<T, 6, 6> projection = Mat<T, 6, 6>::identity();
  Mat<T, 3, 3> first = Mat<T, 3, 3>::identity();
};

// LogarithmicStrain.cpp:29-35
template <class T> T log_divided(T mx, T my) {
  if (mx > my) { T t = mx; mx = my; my = t; }
  const T y = T(1.0) + T(2.0) * my;
  const T d = T(2.0) * (mx - my);
  if (d == T(0.0)) return T(1.0) / y;
  return log1p_(d / y) / d;
}

// LogarithmicStrain.cpp:40-64
template <class T>
T log_divided2(const Vec<T, 3>& mu, const Mat<T, 3, 3>& first, int a, int b, int c) {
  int i = a, j = b, k = c;
  if (mu(i) < mu(j)) std::swap(i, j);
  if (mu(j) < mu(k)) std::swap(j, k);
  if (mu(i) < mu(j)) std::swap(i, j);
  const T y = T(1.0) + T(2.0) * mu(j);
  const T spread = T(2.0) * (mu(i) - mu(k));
  if (spread > T(kSpread) * y) return (first(i, j) - first(j, k)) / spread;
  const T u = T(2.0) * (mu(i) - mu(j)) / y;
  const T w = T(2.0) * (mu(k) - mu(j)) / y;
  T h = T(1.0), wn = T(1.0), sum = T(-0.5);
  for (int n = 1; n <= 5; ++n) {
    wn = wn * w;
    h = u * h + wn;
    sum = sum + (n % 2 == 1 ? T(1.0) : T(-1.0)) * h / T(double(n + 2));
  }
  return sum / (y * y);
}

template <class T> Mat<T, 3, 3> tensor_of(const Vec<T, 6>& v, double shear) {
  Mat<T, 3, 3> t;
  t(0, 0) = v(0); t(1, 1) = v(1); t(2, 2) = v(2);
  t(0, 1) = t(1, 0) = T(shear) * v(3);
  t(1, 2) = t(2, 1) = T(shear) * v(4);
  t(2, 0) = t(0, 2) = T(shear) * v(5);
  return t;
}
template <class T> Vec<T, 6> voigt_of(const Mat<T, 3, 3>& t, double shear) {
  Vec<T, 6> v;
  v(0) = t(0, 0); v(1) = t(1, 1); v(2) = t(2, 2);
  v(3) = T(shear) * t(0, 1); v(4) = T(shear) * t(1, 2); v(5) = T(shear) * t(2, 0);
  return v;
}

// LogarithmicStrain.cpp:86-98
template <class T> Mat<T, 6, 6> principal_rotation(const Mat<T, 3, 3>& q) {
  Mat<T, 6, 6> m;
  for (int r = 0; r < 6; ++r) {
    const int a = kRow[r], b = kCol[r];
    for (int c = 0; c < 6; ++c) {
      const int i = kRow[c], j = kCol[c];
      m(r, c) = i == j ? q(i, a) * q(i, b) : q(i, a) * q(j, b) + q(j, a) * q(i, b);
    }
  }
  return m;
}

// LogarithmicStrain.cpp:109-142
template <class T> LogStrain<T> logarithmic_strain(const Vec<T, 6>& green) {
  if (!all_finite(green)) throw SolverError("the logarithmic strain met a non-finite Green-Lagrange strain");
  LogStrain<T> out;
  jacobi_eigen<T, 3>(tensor_of(green, 0.5), out.green, out.axes, true);
  Vec<T, 3> principal;
  for (int a = 0; a < 3; ++a) {
    const T stretch2 = T(1.0) + T(2.0) * out.green(a);
    if (!(stretch2 > T(0.0)) || !finite_(stretch2)) {
      std::ostringstream os;
      os << "a point is degenerate: a principal stretch squared of C is " << val(stretch2)
         << "; the load step is too large or the mesh too coarse";
      throw SolverError(os.str());
    }
    principal(a) = T(0.5) * log1p_(T(2.0) * out.green(a));
  }
  const Mat<T, 3, 3>& q = out.axes;
  Mat<T, 3, 3> qd, e;
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 3; ++a) qd(i, a) = q(i, a) * principal(a);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) e(i, j) = qd(i, 0) * q(j, 0) + qd(i, 1) * q(j, 1) + qd(i, 2) * q(j, 2);
  out.strain = voigt_of(e, 2.0);
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) out.first(a, b) = out.first(b, a) = log_divided(out.green(a), out.green(b));
  const Mat<T, 6, 6> m = principal_rotation(q);
  Vec<T, 6> sd;
  for (int k = 0; k < 6; ++k) sd(k) = T(k < 3 ? 1.0 : 2.0) * out.first(kRow[k], kCol[k]);
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) {
      T s = m(0, r) * sd(0) * m(0, c);
      for (int k = 1; k < 6; ++k) s = s + m(k, r) * sd(k) * m(k, c);
      out.projection(r, c) = c < 3 ? s : s * T(0.5);
    }
  return out;
}

// LogarithmicStrain.cpp:144-149
template <class T> Vec<T, 6> logarithmic_stress(const LogStrain<T>& log, const Vec<T, 6>& t) {
  const Mat<T, 3, 3>& q = log.axes;
  const Mat<T, 3, 3> tp = mul(mul_tn(q, tensor_of(t, 1.0)), q);
  Mat<T, 3, 3> pr;
  for (int i = 0; i < 9; ++i) pr.a[i] = tp.a[i] * log.first.a[i];
  return voigt_of(mul(mul(q, pr), transpose(q)), 1.0);
}

// LogarithmicStrain.cpp:151-197
template <class T> Mat<T, 6, 6> logarithmic_curvature(const LogStrain<T>& log, const Vec<T, 6>& t) {
  const Mat<T, 3, 3>& q = log.axes;
  const Mat<T, 3, 3> tp = mul(mul_tn(q, tensor_of(t, 1.0)), q);
  T second[27];
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b)
      for (int c = b; c < 3; ++c) {
        const T v = log_divided2(log.green, log.first, a, b, c);
        const int perm[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
        for (const auto& i : perm) second[i[0] + 3 * i[1] + 9 * i[2]] = v;
      }
  auto entries = [](int k, int (&rows)[2], int (&cols)[2]) {
    rows[0] = kRow[k]; cols[0] = kCol[k]; rows[1] = kCol[k]; cols[1] = kRow[k];
    return k < 3 ? 1 : 2;
  };
  auto half = [&](int kx, int ky) {
    int xr[2], xc[2], yr[2], yc[2];
    const int nx = entries(kx, xr, xc);
    const int ny = entries(ky, yr, yc);
    T sum = T(0.0);
    for (int s = 0; s < nx; ++s) {
      const int a = xr[s], c = xc[s];
      for (int r = 0; r < ny; ++r) {
        if (yr[r] != c) continue;
        const int b = yc[r];
        sum = sum + tp(a, b) * second[a + 3 * b + 9 * c];
      }
    }
    return sum;
  };
  Mat<T, 6, 6> hp;
  for (int k = 0; k < 6; ++k)
    for (int l = k; l < 6; ++l) hp(k, l) = hp(l, k) = T(2.0) * (half(k, l) + half(l, k));
  Mat<T, 6, 6> m = principal_rotation(q);
  for (int r = 0; r < 6; ++r)
    for (int c = 3; c < 6; ++c) m(r, c) = m(r, c) * T(0.5);
  return mul(mul_tn(m, hp), m);
}

}  // namespace oracle
