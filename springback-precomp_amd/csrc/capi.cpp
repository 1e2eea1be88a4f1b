// The C ABI of include/sparlab_gpu.h: host-side validation, the material's
// derived data, the integration rule, and the element-kernel launches.
#include "sparlab_gpu.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "contact_kernels.hpp"
#include "element_kernels.hpp"
#include "surface_kernels.hpp"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& message) {
  g_error = message;
  return code;
}

bool finite_positive(double v) { return std::isfinite(v) && v > 0.0; }

// The kernels' copy of a tool record.
sbk::ContactTool contact_tool(const sparlab_tool& tool) {
  sbk::ContactTool t{};
  t.shape = tool.shape;
  for (int k = 0; k < 3; ++k) {
    t.centre[k] = tool.centre[k];
    t.centre0[k] = tool.centre0[k];
    t.normal[k] = tool.normal[k];
    t.axis[k] = tool.axis[k];
  }
  t.radius = tool.radius;
  t.friction = tool.friction;
  t.tangential_ratio = tool.tangential_ratio;
  return t;
}

// What is wrong with a tool record, or null.
const char* tool_problem(const sparlab_tool& tool) {
  if (tool.shape < 0 || tool.shape > 2) return "tool shape must be 0, 1 or 2";
  if (tool.shape != 0 && !(tool.radius > 0.0)) return "a sphere or cylinder needs a radius";
  return nullptr;
}

// Gauss-Legendre abscissae and weights on [-1, 1]: the closed forms for 1..4
// points and Newton's iteration on P_n for 5..7
// (src/elements/Quadrature.cpp:16-48, 100-139).
std::vector<std::pair<double, double>> gauss_line(int n) {
  switch (n) {
    case 1: return {{0.0, 2.0}};
    case 2: {
      const double a = 1.0 / std::sqrt(3.0);
      return {{-a, 1.0}, {a, 1.0}};
    }
    case 3: {
      const double a = std::sqrt(3.0 / 5.0);
      return {{-a, 5.0 / 9.0}, {0.0, 8.0 / 9.0}, {a, 5.0 / 9.0}};
    }
    case 4: {
      const double r = std::sqrt(6.0 / 5.0);
      const double inner = std::sqrt((3.0 - 2.0 * r) / 7.0);
      const double outer = std::sqrt((3.0 + 2.0 * r) / 7.0);
      const double s30 = std::sqrt(30.0);
      return {{-outer, (18.0 - s30) / 36.0},
              {-inner, (18.0 + s30) / 36.0},
              {inner, (18.0 + s30) / 36.0},
              {outer, (18.0 - s30) / 36.0}};
    }
    default: break;
  }
  std::vector<std::pair<double, double>> out(static_cast<std::size_t>(n));
  const double pi = std::acos(-1.0);
  auto legendre = [n](double x, double& pn, double& pm) {
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    pn = p1;
    pm = p0;
  };
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double x = std::cos(pi * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double pn, pm;
      legendre(x, pn, pm);
      const double dp = n * (x * pn - pm) / (x * x - 1.0);
      const double step = pn / dp;
      x -= step;
      if (std::abs(step) <= 1.0e-16) break;
    }
    if (2 * i + 1 == n) x = 0.0;
    double pn, pm;
    legendre(x, pn, pm);
    const double dp = n * (x * pn - pm) / (x * x - 1.0);
    const double w = 2.0 / ((1.0 - x * x) * dp * dp);
    out[static_cast<std::size_t>(i)] = {-x, w};
    out[static_cast<std::size_t>(n - 1 - i)] = {x, w};
  }
  return out;
}

// The Hex8 stiffness rule with xi fastest, then eta, then zeta
// (src/elements/Quadrature.cpp:72-93, 165-219).
int build_rule(const sparlab_rule& r, sbk::RuleArg& out, std::string& why) {
  const int n = r.stiffness_points;
  if (n < 1 || n > 4) { why = "stiffness_points must be 1..4"; return -1; }
  const int m = (r.thickness_points <= 0 || r.thickness_points == n) ? n : r.thickness_points;
  if (m < 1 || m > 7) { why = "thickness_points must be 0 or 1..7"; return -1; }
  if (r.thickness_axis < 0 || r.thickness_axis > 2) { why = "thickness_axis must be 0, 1 or 2"; return -1; }
  const auto plane = gauss_line(n);
  const auto through = gauss_line(m);
  const auto& rx = r.thickness_axis == 0 ? through : plane;
  const auto& ry = r.thickness_axis == 1 ? through : plane;
  const auto& rz = r.thickness_axis == 2 ? through : plane;
  const std::size_t total = rx.size() * ry.size() * rz.size();
  if (total > static_cast<std::size_t>(sbk::kMaxPoints)) {
    why = "the element kernels take integration rules of at most " + std::to_string(sbk::kMaxPoints) + " points (got " +
          std::to_string(total) + ")";
    return -2;
  }
  int q = 0;
  for (const auto& gk : rz)
    for (const auto& gj : ry)
      for (const auto& gi : rx) {
        out.xi[q] = gi.first;
        out.eta[q] = gj.first;
        out.zeta[q] = gk.first;
        out.w[q] = gi.second * gj.second * gk.second;
        ++q;
      }
  out.n = q;
  return 0;
}

int kinematic_terms(const sparlab_material& m) {
  return m.num_backstresses + (m.kinematic_hardening_modulus > 0.0 ? 1 : 0);
}
bool general_law(const sparlab_material& m) { return m.criterion == 1 || m.num_backstresses > 0; }
int kin_fields(const sparlab_material& m) { return general_law(m) ? kinematic_terms(m) : 1; }

// The material frame's rows (Material.cpp:190-204): the rolling direction
// made orthogonal to the sheet normal, the transverse direction, the normal.
bool material_frame(const sparlab_material& m, double q[3][3], std::string& why) {
  const double* rdi = m.rolling_direction;
  const double* ndi = m.sheet_normal;
  const double nn = std::sqrt(ndi[0] * ndi[0] + ndi[1] * ndi[1] + ndi[2] * ndi[2]);
  const double rn = std::sqrt(rdi[0] * rdi[0] + rdi[1] * rdi[1] + rdi[2] * rdi[2]);
  if (!(nn > 0.0) || !(rn > 0.0) || !std::isfinite(nn) || !std::isfinite(rn)) {
    why = "the rolling direction and the sheet normal must be finite non-zero vectors";
    return false;
  }
  double nd[3], rd[3];
  for (int i = 0; i < 3; ++i) { nd[i] = ndi[i] / nn; rd[i] = rdi[i] / rn; }
  const double c = rd[0] * nd[0] + rd[1] * nd[1] + rd[2] * nd[2];
  for (int i = 0; i < 3; ++i) rd[i] -= c * nd[i];
  const double r2 = std::sqrt(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
  if (!(r2 > 1.0e-6)) {
    why = "the rolling direction is parallel to the sheet normal";
    return false;
  }
  for (int i = 0; i < 3; ++i) rd[i] /= r2;
  const double td[3] = {nd[1] * rd[2] - nd[2] * rd[1], nd[2] * rd[0] - nd[0] * rd[2], nd[0] * rd[1] - nd[1] * rd[0]};
  for (int j = 0; j < 3; ++j) {
    q[0][j] = rd[j];
    q[1][j] = td[j];
    q[2][j] = nd[j];
  }
  return true;
}

// P = T^T P_m T: the Hill48 matrix of (F, G, H, L, M, N) in the material
// frame, mapped to global axes (Material.cpp:205-216). Linear in the
// coefficients, so it maps coefficient derivatives as well.
void hill_global(const double q[3][3], double F, double G, double H, double L, double M, double N, double* out) {
  double pm[6][6] = {};
  pm[0][0] = G + H;
  pm[1][1] = F + H;
  pm[2][2] = F + G;
  pm[0][1] = pm[1][0] = -H;
  pm[1][2] = pm[2][1] = -F;
  pm[0][2] = pm[2][0] = -G;
  pm[3][3] = 2.0 * N;
  pm[4][4] = 2.0 * L;
  pm[5][5] = 2.0 * M;
  // Voigt map of sigma' = q sigma q^T, column k = the image of basis tensor k
  static const int vi[6] = {0, 1, 2, 0, 1, 2};
  static const int vj[6] = {0, 1, 2, 1, 2, 0};
  double t[6][6];
  for (int k = 0; k < 6; ++k) {
    double e[3][3] = {};
    e[vi[k]][vj[k]] = 1.0;
    e[vj[k]][vi[k]] = 1.0;
    double qe[3][3], r[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) qe[i][j] = q[i][0] * e[0][j] + q[i][1] * e[1][j] + q[i][2] * e[2][j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r[i][j] = qe[i][0] * q[j][0] + qe[i][1] * q[j][1] + qe[i][2] * q[j][2];
    for (int row = 0; row < 6; ++row) t[row][k] = r[vi[row]][vj[row]];
  }
  double tp[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += t[k][i] * pm[k][j];
      tp[i][j] = s;
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += tp[i][k] * t[k][j];
      out[i * 6 + j] = s;
    }
}

}  // namespace

namespace sbk {

// dP/d(r0, r45, r90) of a finalized Hill48 material calibrated by r-values
// (the derivatives of Material.cpp:140-151 mapped to global axes); false
// (and zeros) for any other criterion or calibration.
bool hill_r_derivatives(const sparlab_material& m, double dP[3][36]) {
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 36; ++i) dP[k][i] = 0.0;
  if (m.criterion != 1 || m.hill_calibration != 0) return false;
  double q[3][3];
  std::string why;
  if (!material_frame(m, q, why)) return false;
  const double r0 = m.hill_r[0], r45 = m.hill_r[1], r90 = m.hill_r[2];
  const double a = 1.0 + r0;
  // G = 1/(1+r0), H = r0/(1+r0), F = r0/(r90(1+r0)), N = (r0+r90)(1+2r45)/(2 r90 (1+r0))
  const double dG[3] = {-1.0 / (a * a), 0.0, 0.0};
  const double dH[3] = {1.0 / (a * a), 0.0, 0.0};
  const double dF[3] = {1.0 / (r90 * a * a), 0.0, -r0 / (r90 * r90 * a)};
  const double dN[3] = {(1.0 + 2.0 * r45) * (1.0 - r90) / (2.0 * r90 * a * a), (r0 + r90) / (r90 * a),
                        -(1.0 + 2.0 * r45) * r0 / (2.0 * r90 * r90 * a)};
  for (int k = 0; k < 3; ++k) hill_global(q, dF[k], dG[k], dH[k], 0.0, 0.0, dN[k], dP[k]);
  return true;
}

}  // namespace sbk

namespace {

// Eigen-decomposition of the symmetric 6 x 6 matrix a by cyclic Jacobi
// rotations (a is destroyed): a = V diag(l) V^T, V row-major, column k the
// eigenvector of l[k].
void symmetric_eigen6(double a[6][6], double l[6], double V[36]) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) V[i * 6 + j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 6; ++p)
      for (int q = p + 1; q < 6; ++q) off += a[p][q] * a[p][q];
    if (off == 0.0) break;
    for (int p = 0; p < 5; ++p)
      for (int q = p + 1; q < 6; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = 0.5 * (a[q][q] - a[p][p]) / a[p][q];
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < 6; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - sn * akq;
          a[k][q] = sn * akp + c * akq;
        }
        for (int k = 0; k < 6; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - sn * aqk;
          a[q][k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < 6; ++k) {
          const double vkp = V[k * 6 + p], vkq = V[k * 6 + q];
          V[k * 6 + p] = c * vkp - sn * vkq;
          V[k * 6 + q] = sn * vkp + c * vkq;
        }
      }
  }
  for (int k = 0; k < 6; ++k) l[k] = a[k][k];
}

sbk::DevMaterial device_material(const sparlab_material& m) {
  sbk::DevMaterial d{};
  d.shear = m.youngs_modulus / (2.0 * (1.0 + m.poisson_ratio));
  const double lame = m.youngs_modulus * m.poisson_ratio / ((1.0 + m.poisson_ratio) * (1.0 - 2.0 * m.poisson_ratio));
  d.bulk = lame + 2.0 * d.shear / 3.0;
  d.sy0 = m.yield_stress;
  d.hard = m.hardening_modulus;
  d.hkin = m.kinematic_hardening_modulus;
  d.qsat = m.saturation_stress;
  d.delta = m.saturation_rate;
  std::memcpy(d.P, m.yield_matrix, sizeof(d.P));
  {
    // D P D with D = diag(1, 1, 1, 2^-1/2 (x3)), for the block solve of the
    // general return's Jacobian
    double dpd[6][6];
    const double dd[6] = {1.0, 1.0, 1.0, std::sqrt(0.5), std::sqrt(0.5), std::sqrt(0.5)};
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) dpd[i][j] = 0.5 * (dd[i] * m.yield_matrix[i * 6 + j] * dd[j] +
                                                     dd[j] * m.yield_matrix[j * 6 + i] * dd[i]);
    symmetric_eigen6(dpd, d.L, d.V);
  }
  d.general = general_law(m) ? 1 : 0;
  d.terms = kinematic_terms(m);
  for (int i = 0; i < d.terms; ++i) {
    d.mod[i] = i < m.num_backstresses ? m.backstress_modulus[i] : m.kinematic_hardening_modulus;
    d.rec[i] = i < m.num_backstresses ? m.backstress_recovery[i] : 0.0;
  }
  d.exponential = m.kinematic_integration == 0 ? 1 : 0;
  d.enabled = m.yield_stress > 0.0 ? 1 : 0;
  d.symmetric = 1;
  for (int i = 0; i < m.num_backstresses; ++i)
    if (m.backstress_recovery[i] > 0.0) d.symmetric = 0;
  d.kin_fields = kin_fields(m);
  return d;
}

}  // namespace

namespace sbk {

int capi_fail(int code, const std::string& message) { return fail(code, message); }

// Validates the arguments the element kernels share and fills the launch
// record's rule and material.
int capi_prepare_elements(const sparlab_material* material, const sparlab_rule* rule, int kinematics,
                          ElementArgs& a) {
  if (material == nullptr || rule == nullptr) return fail(SPARLAB_E_ARGUMENT, "null argument");
  if (kinematics < 0 || kinematics > 2) return fail(SPARLAB_E_ARGUMENT, "kinematics must be 0, 1 or 2");
  if (kinematic_terms(*material) > 4) return fail(SPARLAB_E_ARGUMENT, "at most 4 kinematic terms");
  std::string why;
  const int rs = build_rule(*rule, a.rule, why);
  if (rs == -1) return fail(SPARLAB_E_ARGUMENT, why);
  if (rs == -2) return fail(SPARLAB_E_UNSUPPORTED, why);
  a.kinematics = kinematics;
  a.mat = device_material(*material);
  a.ordered = 1;
  return SPARLAB_OK;
}

}  // namespace sbk

extern "C" {

const char* sparlab_last_error(void) { return g_error.c_str(); }

const char* sparlab_version(void) { return "sparlab-gpu 0.1 (gfx950)"; }

int sparlab_material_finalize(sparlab_material* m) {
  if (m == nullptr) return fail(SPARLAB_E_ARGUMENT, "null material");
  if (!(m->youngs_modulus > 0.0)) return fail(SPARLAB_E_ARGUMENT, "the material needs a positive Young's modulus");
  if (!(m->poisson_ratio > -1.0 && m->poisson_ratio < 0.5))
    return fail(SPARLAB_E_ARGUMENT, "the material needs a Poisson ratio in (-1, 0.5)");
  const double values[] = {m->yield_stress, m->hardening_modulus, m->kinematic_hardening_modulus,
                           m->saturation_stress, m->saturation_rate};
  for (double v : values)
    if (!std::isfinite(v) || v < 0.0)
      return fail(SPARLAB_E_ARGUMENT, "the material needs finite, non-negative plasticity parameters");
  if (m->saturation_stress > 0.0 && !(m->saturation_rate > 0.0))
    return fail(SPARLAB_E_ARGUMENT, "a Voce saturation stress needs a positive saturation rate");
  if (m->num_backstresses < 0 || m->num_backstresses > 4)
    return fail(SPARLAB_E_ARGUMENT, "at most 4 backstresses are supported");
  for (int i = 0; i < m->num_backstresses; ++i)
    if (!finite_positive(m->backstress_modulus[i]) || !std::isfinite(m->backstress_recovery[i]) ||
        m->backstress_recovery[i] < 0.0)
      return fail(SPARLAB_E_ARGUMENT, "a backstress needs a finite positive modulus and a non-negative recovery");
  const bool enabled = m->yield_stress > 0.0;
  if (!enabled && (m->hardening_modulus > 0.0 || m->kinematic_hardening_modulus > 0.0 ||
                   m->saturation_stress > 0.0 || m->num_backstresses > 0))
    return fail(SPARLAB_E_ARGUMENT, "hardening parameters need a yield stress");
  if (!enabled && m->criterion != 0) return fail(SPARLAB_E_ARGUMENT, "a yield criterion needs a yield stress");
  if (general_law(*m) && kinematic_terms(*m) > 4)
    return fail(SPARLAB_E_ARGUMENT, "backstresses and Prager's modulus: at most 4 kinematic terms in all");
  if (m->criterion != 1) {
    // von Mises: (3/2) |dev xi|^2 (Material.cpp:92-102)
    std::memset(m->yield_matrix, 0, sizeof(m->yield_matrix));
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) m->yield_matrix[i * 6 + j] = -0.5;
      m->yield_matrix[i * 6 + i] = 1.0;
      m->yield_matrix[(3 + i) * 6 + 3 + i] = 3.0;
    }
    return SPARLAB_OK;
  }
  // Hill48 (Material.cpp:135-217)
  double F, G, H, L, M, N;
  if (m->hill_calibration == 0) {
    const double r0 = m->hill_r[0], r45 = m->hill_r[1], r90 = m->hill_r[2];
    if (!finite_positive(r0) || !finite_positive(r45) || !finite_positive(r90))
      return fail(SPARLAB_E_ARGUMENT, "the Hill48 r-values must be finite and positive");
    G = 1.0 / (1.0 + r0);
    H = r0 / (1.0 + r0);
    F = r0 / (r90 * (1.0 + r0));
    N = (r0 + r90) * (1.0 + 2.0 * r45) / (2.0 * r90 * (1.0 + r0));
    L = m->hill_shear_lm[0];
    M = m->hill_shear_lm[1];
  } else if (m->hill_calibration == 1) {
    const double s45 = m->hill_stress_ratio[0], s90 = m->hill_stress_ratio[1], sb = m->hill_stress_ratio[2];
    if (!finite_positive(s45) || !finite_positive(s90) || !finite_positive(sb))
      return fail(SPARLAB_E_ARGUMENT, "the Hill48 yield stress ratios must be finite and positive");
    const double a = 1.0 / (s90 * s90), b = 1.0 / (sb * sb);
    F = 0.5 * (a + b - 1.0);
    G = 0.5 * (1.0 - a + b);
    H = 0.5 * (1.0 + a - b);
    N = 0.5 * (4.0 / (s45 * s45) - F - G);
    L = m->hill_shear_lm[0];
    M = m->hill_shear_lm[1];
  } else {
    F = m->hill_coefficients[0];
    G = m->hill_coefficients[1];
    H = m->hill_coefficients[2];
    L = m->hill_coefficients[3];
    M = m->hill_coefficients[4];
    N = m->hill_coefficients[5];
  }
  for (double v : {F, G, H, L, M, N})
    if (!std::isfinite(v)) return fail(SPARLAB_E_ARGUMENT, "the Hill48 coefficients must be finite");
  if (!(L > 0.0 && M > 0.0 && N > 0.0) || !(F + H > 0.0) || !(F * G + G * H + H * F > 0.0))
    return fail(SPARLAB_E_ARGUMENT, "the Hill48 coefficients do not make a convex yield surface");
  double q[3][3];
  std::string why;
  if (!material_frame(*m, q, why)) return fail(SPARLAB_E_ARGUMENT, why);
  hill_global(q, F, G, H, L, M, N, m->yield_matrix);
  const double coeffs[6] = {F, G, H, L, M, N};
  std::memcpy(m->hill_coefficients, coeffs, sizeof(coeffs));
  return SPARLAB_OK;
}

int sparlab_rule_points(const sparlab_rule* rule) {
  if (rule == nullptr) return fail(SPARLAB_E_ARGUMENT, "null rule");
  sbk::RuleArg r;
  std::string why;
  const int st = build_rule(*rule, r, why);
  if (st == -1) return fail(SPARLAB_E_ARGUMENT, why);
  if (st == -2) return fail(SPARLAB_E_UNSUPPORTED, why);
  return r.n;
}

int sparlab_state_fields(const sparlab_material* m) {
  if (m == nullptr) return fail(SPARLAB_E_ARGUMENT, "null material");
  return sbk::state_fields(kin_fields(*m));
}

int64_t sparlab_element_workspace_doubles(int64_t num_elements, const sparlab_rule* rule) {
  if (rule == nullptr || num_elements < 0) return fail(SPARLAB_E_ARGUMENT, "null rule or negative size");
  sbk::RuleArg r{};
  std::string why;
  const int rs = build_rule(*rule, r, why);
  if (rs == -1) return fail(SPARLAB_E_ARGUMENT, why);
  if (rs == -2) return fail(SPARLAB_E_UNSUPPORTED, why);
  return static_cast<int64_t>(sbk::workspace_doubles(num_elements, r.n));
}

}  // extern "C"

namespace sbk {

// The checks of an element batch that need no device: null pointers, sizes,
// kinematics, stages, the rule and the material. Fills a's rule and material
// and returns SPARLAB_OK, or an error code with the message set. The batch's
// workspace pointer is not looked at (the caller supplies the workspace).
int capi_check_element_batch(const sparlab_material* material, const sparlab_rule* rule, int kinematics,
                             int want_tangent, const sparlab_element_batch* b, int stages, ElementArgs& a) {
  if (material == nullptr || rule == nullptr || b == nullptr) return fail(SPARLAB_E_ARGUMENT, "null argument");
  if (kinematics < 0 || kinematics > 2) return fail(SPARLAB_E_ARGUMENT, "kinematics must be 0, 1 or 2");
  if ((stages & ~(3 | SPARLAB_STAGE_INDEX_ORDER)) != 0 || (stages & 3) == 0)
    return fail(SPARLAB_E_ARGUMENT, "stages must be 1, 2 or 3, optionally with SPARLAB_STAGE_INDEX_ORDER");
  if (b->num_elements < 0 || b->num_nodes < 0) return fail(SPARLAB_E_ARGUMENT, "negative sizes");
  if (b->num_elements == 0) return SPARLAB_OK;
  if (b->coords == nullptr || b->conn == nullptr || b->u == nullptr || b->state_in == nullptr ||
      b->state_out == nullptr || b->alpha_out == nullptr || b->force == nullptr || b->energy == nullptr ||
      b->yielding == nullptr || b->iterations == nullptr || b->status == nullptr || b->fail_value == nullptr)
    return fail(SPARLAB_E_ARGUMENT, "a required buffer of the element batch is null");
  if (want_tangent && b->tangent == nullptr) return fail(SPARLAB_E_ARGUMENT, "want_tangent needs a tangent buffer");
  return capi_prepare_elements(material, rule, kinematics, a);
}

// Queues the stages of an element batch already checked by
// capi_check_element_batch (a holds its rule and material) with `workspace`
// (workspace_doubles(num_elements, points) doubles of device memory).
int capi_launch_element_batch(int want_tangent, const sparlab_element_batch* b, double* workspace, int stages,
                              void* stream, ElementArgs& a) {
  if (b->num_elements == 0) return SPARLAB_OK;
  // the condensation copies the workspace's records in 16-byte chunks
  if (reinterpret_cast<std::uintptr_t>(workspace) % 16 != 0)
    return fail(SPARLAB_E_ARGUMENT, "the workspace must be 16-byte aligned");
  a.coords = b->coords;
  a.conn = b->conn;
  a.u = b->u;
  a.ne = b->num_elements;
  a.state_in = b->state_in;
  a.state_stride = b->num_elements * a.rule.n;
  a.alpha_in = b->alpha_in;
  a.alpha_out = b->alpha_out;
  a.state_out = b->state_out;
  a.force = b->force;
  a.tangent = want_tangent ? b->tangent : nullptr;
  a.energy = b->energy;
  a.yielding = b->yielding;
  a.iterations = b->iterations;
  a.status = b->status;
  a.fail_value = b->fail_value;
  a.local_path = b->local_path;
  set_workspace(a, workspace);
  if (stages & SPARLAB_STAGE_INDEX_ORDER) a.ordered = 0;
  // The local solve writes every element's status; the condensation reads it.
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (stages & 1) {
    const int err = launch_local_solve(a, s);
    if (err != 0)
      return fail(SPARLAB_E_DEVICE, std::string("the local-solve kernel did not launch: ") +
                                        hipGetErrorString(static_cast<hipError_t>(err)));
  }
  if (stages & 2) {
    const int err = launch_condense(a, want_tangent != 0, s);
    if (err != 0)
      return fail(SPARLAB_E_DEVICE, std::string("the condensation kernel did not launch: ") +
                                        hipGetErrorString(static_cast<hipError_t>(err)));
  }
  return SPARLAB_OK;
}

}  // namespace sbk

extern "C" {

int sparlab_elements_evaluate_stage_device(const sparlab_material* material, const sparlab_rule* rule,
                                           int kinematics, int want_tangent, const sparlab_element_batch* b,
                                           int stages, void* stream) {
  sbk::ElementArgs a{};
  const int st = sbk::capi_check_element_batch(material, rule, kinematics, want_tangent, b, stages, a);
  if (st != SPARLAB_OK) return st;
  if (b->num_elements > 0 && b->workspace == nullptr)
    return fail(SPARLAB_E_ARGUMENT, "a required buffer of the element batch is null");
  return sbk::capi_launch_element_batch(want_tangent, b, b->workspace, stages, stream, a);
}

int sparlab_elements_evaluate_device(const sparlab_material* material, const sparlab_rule* rule, int kinematics,
                                     int want_tangent, const sparlab_element_batch* b, void* stream) {
  return sparlab_elements_evaluate_stage_device(material, rule, kinematics, want_tangent, b, 3, stream);
}

int sparlab_contact_candidates_device(const sparlab_tool* tool, const double* coords, const double* u0,
                                      int64_t num_slaves, const int32_t* nodes, double margin, int32_t* candidates,
                                      int32_t* count, int32_t* work, void* stream) {
  if (tool == nullptr || coords == nullptr || u0 == nullptr || count == nullptr ||
      (num_slaves > 0 && (nodes == nullptr || candidates == nullptr || work == nullptr)))
    return fail(SPARLAB_E_ARGUMENT, "null argument");
  if (num_slaves < 0 || num_slaves > (int64_t{1} << 31) - 257) return fail(SPARLAB_E_ARGUMENT, "bad slave count");
  if (const char* why = tool_problem(*tool)) return fail(SPARLAB_E_ARGUMENT, why);
  if (sbk::launch_contact_candidates(contact_tool(*tool), coords, u0, num_slaves, nodes, margin, candidates, count, work,
                                     static_cast<hipStream_t>(stream)) != 0)
    return fail(SPARLAB_E_DEVICE, "candidate launch failed");
  return SPARLAB_OK;
}

int sparlab_contact_anticipate_device(const sparlab_tool* tool, const double* coords, const double* u,
                                      const double* du, int64_t num_slaves, const int32_t* nodes,
                                      const double* kappa_area, double* residual, double* blocks, int32_t* active,
                                      int32_t* added, void* stream) {
  if (tool == nullptr || coords == nullptr || u == nullptr || du == nullptr || residual == nullptr ||
      added == nullptr || (num_slaves > 0 && (nodes == nullptr || kappa_area == nullptr || blocks == nullptr ||
                                              active == nullptr)))
    return fail(SPARLAB_E_ARGUMENT, "null argument");
  if (num_slaves < 0) return fail(SPARLAB_E_ARGUMENT, "bad slave count");
  if (const char* why = tool_problem(*tool)) return fail(SPARLAB_E_ARGUMENT, why);
  if (sbk::launch_contact_anticipate(contact_tool(*tool), coords, u, du, num_slaves, nodes, kappa_area, residual,
                                     blocks, active, added, static_cast<hipStream_t>(stream)) != 0)
    return fail(SPARLAB_E_DEVICE, "anticipation launch failed");
  return SPARLAB_OK;
}

int sparlab_contact_evaluate_device(const sparlab_tool* tool, int friction_symmetric, int want_tangent,
                                    const double* coords, const double* u, const double* start,
                                    int64_t num_slaves, const int32_t* nodes, const double* kappa_area,
                                    const int32_t* hist_present, const double* hist_force,
                                    const double* hist_relative, double* residual, double* blocks,
                                    int32_t* active, int32_t* trial_present, double* trial_force,
                                    double* trial_relative, double* work, int32_t* work_status,
                                    double* scalars, void* stream) {
  if (tool == nullptr || coords == nullptr || u == nullptr || start == nullptr || residual == nullptr ||
      active == nullptr || trial_present == nullptr || trial_force == nullptr || trial_relative == nullptr ||
      work == nullptr || work_status == nullptr || scalars == nullptr || (num_slaves > 0 && (nodes == nullptr || kappa_area == nullptr)))
    return fail(SPARLAB_E_ARGUMENT, "null argument");
  if (want_tangent && blocks == nullptr) return fail(SPARLAB_E_ARGUMENT, "want_tangent needs a blocks buffer");
  if (const char* why = tool_problem(*tool)) return fail(SPARLAB_E_ARGUMENT, why);
  if (hist_present != nullptr && (hist_force == nullptr || hist_relative == nullptr))
    return fail(SPARLAB_E_ARGUMENT, "a friction history needs its forces and positions");
  sbk::ContactArgs a{};
  a.tool = contact_tool(*tool);
  a.friction_symmetric = friction_symmetric;
  a.want_tangent = want_tangent;
  a.coords = coords;
  a.u = u;
  a.start = start;
  a.num_slaves = num_slaves;
  a.nodes = nodes;
  a.kappa_area = kappa_area;
  a.hist_present = hist_present;
  a.hist_force = hist_force;
  a.hist_relative = hist_relative;
  a.residual = residual;
  a.blocks = blocks;
  a.active = active;
  a.trial_present = trial_present;
  a.trial_force = trial_force;
  a.trial_relative = trial_relative;
  a.node_terms = work;
  a.status = work_status;
  if (sbk::launch_contact(a, scalars, static_cast<hipStream_t>(stream)) != 0)
    return fail(SPARLAB_E_DEVICE, "the contact kernel did not launch");
  return SPARLAB_OK;
}

int sparlab_add_node_blocks_device(const double* blocks, const int32_t* active, const int32_t* first_value,
                                   const int32_t* row_len, int64_t num_slaves, int column_major, double* values,
                                   void* stream) {
  if (num_slaves > 0 && (blocks == nullptr || active == nullptr || first_value == nullptr || row_len == nullptr ||
                         values == nullptr))
    return fail(SPARLAB_E_ARGUMENT, "null argument");
  if (sbk::launch_add_node_blocks(blocks, active, first_value, row_len, num_slaves, column_major, values,
                                  static_cast<hipStream_t>(stream)) != 0)
    return fail(SPARLAB_E_DEVICE, "the node-block kernel did not launch");
  return SPARLAB_OK;
}

}  // extern "C"

int sparlab_surface_raycast_device(const double* surface, int64_t nx_s, int64_t ny_s, double x0, double y0,
                                   int64_t nx, int64_t ny, double h, double* z, uint8_t* hit, int64_t* folded,
                                   void* stream) {
  if (!surface || !z || !hit || !folded) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "null argument");
  if (nx_s < 2 || ny_s < 2 || nx < 1 || ny < 1) return sbk::capi_fail(SPARLAB_E_ARGUMENT, "bad surface or grid size");
  if (!(h > 0.0) || !std::isfinite(h) || !std::isfinite(x0) || !std::isfinite(y0))
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, "the grid spacing must be finite and > 0, its origin finite");
  hipStream_t st = static_cast<hipStream_t>(stream);
  unsigned long long* d_folded = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d_folded), sizeof(unsigned long long)) != hipSuccess)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "device allocation failed");
  unsigned long long count = 0;
  const bool ok = hipMemsetAsync(d_folded, 0, sizeof(count), st) == hipSuccess &&
                  sbk::launch_surface_folds(surface, nx_s, ny_s, d_folded, st) == 0 &&
                  hipMemcpyAsync(&count, d_folded, sizeof(count), hipMemcpyDeviceToHost, st) == hipSuccess &&
                  hipStreamSynchronize(st) == hipSuccess;
  (void)hipFree(d_folded);
  if (!ok) return sbk::capi_fail(SPARLAB_E_DEVICE, "the surface's fold check failed");
  *folded = static_cast<int64_t>(count);
  if (count > 0)
    return sbk::capi_fail(SPARLAB_E_ARGUMENT, std::to_string(count) +
                                                  " surface triangles fold over in plan view; the deformed surface is "
                                                  "not a height field");
  const sbk::HeightGrid g{x0, y0, h, nx, ny};
  if (sbk::launch_surface_raycast(surface, nx_s, ny_s, g, z, hit, st) != 0)
    return sbk::capi_fail(SPARLAB_E_DEVICE, "the surface ray-cast failed");
  return SPARLAB_OK;
}
