// Device material-point returns (gfx950, FP64): the J2 radial return and the
// closest-point return of Hill48 with Chaboche/Prager backstresses, each with
// its consistent tangent, for one integration point per lane.
//
// Algorithm: src/material/Plasticity.cpp:66-171 (return_3d) and :185-471
// (GeneralReturn, return_general_3d); the same tolerances and iteration
// limits (Voce Newton: 60 iterations, 1e-13 radius, fallback 8 eps |xi_tr|;
// general: 50 Newton iterations, 40 halvings, tol max(1e-13 sy, 32 eps scale),
// stagnation within 4 tol, one polishing step).
#pragma once

#include "device_math.hpp"

namespace sbk {

// History of one point, in registers; NT 6-vectors of kinematic history
// (kernels are instantiated for the smallest NT that holds the material's).
template <int NT>
struct PointStateT {
  double ep[6];
  double kin[NT][6];  // radial return: kin[0] = beta; general: alpha_i
  double abar;
  bool loading;
};
using PointState = PointStateT<kMaxTerms>;

struct ReturnOut {
  double stress[6];
  double C[6][6];
  double energy;
  bool yielding;
};

enum ReturnFail { kReturnOk = 0, kReturnNotConverged = 5, kReturnNonFinite = 9 };

__device__ __forceinline__ double tensor_norm(const double v[6]) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 2.0 * (v[3] * v[3] + v[4] * v[4] + v[5] * v[5]));
}
__device__ __forceinline__ double dev_norm2(const double e[6]) {
  const double mean = (e[0] + e[1] + e[2]) / 3.0;
  double out = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) out = out + (e[i] - mean) * (e[i] - mean);
#pragma unroll
  for (int i = 3; i < 6; ++i) out = out + 0.5 * e[i] * e[i];
  return out;
}
__device__ __forceinline__ double strain_norm2(const double e[6]) {
  double out = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) out = out + e[i] * e[i];
#pragma unroll
  for (int i = 3; i < 6; ++i) out = out + 0.5 * e[i] * e[i];
  return out;
}
__device__ __forceinline__ void elastic_C(double k, double g, double C[6][6]) {
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) C[i][j] = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) C[i][j] = k - 2.0 * g / 3.0;
    C[i][i] = k + 4.0 * g / 3.0;
    C[3 + i][3 + i] = g;
  }
}

// ---------------------------------------------------------------------------
// J2 radial return (Plasticity.cpp:66-171); the state is updated in place.
// ---------------------------------------------------------------------------
template <int NT>
__device__ int return_radial(const DevMaterial& m, const double strain[6], PointStateT<NT>& st, ReturnOut& out) {
  const double g = m.shear, k = m.bulk, hk = m.hkin;
  const double sqrt23 = 0.816496580927726;  // sqrt(2/3), rounded as std::sqrt(2.0 / 3.0)
  double el[6], s_tr[6], xi[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) el[i] = strain[i] - st.ep[i];
  const double vol = el[0] + el[1] + el[2];
  const double mean = vol / 3.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) s_tr[i] = 2.0 * g * (el[i] - mean);
#pragma unroll
  for (int i = 3; i < 6; ++i) s_tr[i] = g * el[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) xi[i] = s_tr[i] - st.kin[0][i];
  const double xn = tensor_norm(xi);
  const double an = st.abar;
  const double rn = sqrt23 * yield_stress(m, an);
  const double f_tr = xn - rn;
  const bool was_loading = st.loading;
  st.loading = false;
  if (!m.enabled || f_tr <= 1.0e-12 * rn) {
#pragma unroll
    for (int i = 0; i < 6; ++i) out.stress[i] = s_tr[i] + (i < 3 ? k * vol : 0.0);
    elastic_C(k, g, out.C);
    out.energy = 0.5 * k * vol * vol + g * dev_norm2(el) + isotropic_energy(m, an) + hk / 3.0 * strain_norm2(st.ep);
    out.yielding = false;
    if (m.enabled && was_loading && f_tr > -1.0e-10 * rn && xn > 0.0) {
      double n[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) n[i] = xi[i] / xn;
      const double fac = 2.0 * g / (1.0 + (yield_slope(m, an) + hk) / (3.0 * g));
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) out.C[i][j] = out.C[i][j] - fac * (n[i] * n[j]);
      st.loading = true;
    }
    return kReturnOk;
  }
  const double stiff = 2.0 * g + 2.0 / 3.0 * hk;
  double dg = f_tr / (stiff + 2.0 / 3.0 * yield_slope(m, an));
  if (m.qsat != 0.0) {
    dg = 0.0;
    bool conv = false;
    for (int it = 0; it < 60; ++it) {
      const double a = an + sqrt23 * dg;
      const double res = xn - stiff * dg - sqrt23 * yield_stress(m, a);
      if (fabs(res) <= 1.0e-13 * rn) { conv = true; break; }
      dg = dg + res / (stiff + 2.0 / 3.0 * yield_slope(m, a));
    }
    if (!conv) {
      const double a = an + sqrt23 * dg;
      const double res = xn - stiff * dg - sqrt23 * yield_stress(m, a);
      conv = fabs(res) <= 8.0 * kEps * xn;
    }
    if (!conv) return kReturnNotConverged;
  }
  const double alpha = an + sqrt23 * dg;
  double n[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) n[i] = xi[i] / xn;
  out.yielding = true;
#pragma unroll
  for (int i = 0; i < 6; ++i) out.stress[i] = s_tr[i] - 2.0 * g * dg * n[i] + (i < 3 ? k * vol : 0.0);
  st.loading = true;
  st.abar = alpha;
#pragma unroll
  for (int i = 0; i < 6; ++i) st.kin[0][i] = st.kin[0][i] + 2.0 / 3.0 * hk * dg * n[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) st.ep[i] = st.ep[i] + (i < 3 ? 1.0 : 2.0) * dg * n[i];
  const double theta = 1.0 - 2.0 * g * dg / xn;
  const double theta_bar = 1.0 / (1.0 + (yield_slope(m, alpha) + hk) / (3.0 * g)) - (1.0 - theta);
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) out.C[i][j] = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) out.C[i][j] = k - 2.0 * g * theta / 3.0;
    out.C[i][i] = k + 4.0 * g * theta / 3.0;
    out.C[3 + i][3 + i] = g * theta;
  }
  const double f2 = 2.0 * g * theta_bar;
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) out.C[i][j] = out.C[i][j] - f2 * (n[i] * n[j]);
  double eln[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) eln[i] = el[i] - (i < 3 ? 1.0 : 2.0) * dg * n[i];
  out.energy = 0.5 * k * vol * vol + g * dev_norm2(eln) + isotropic_energy(m, alpha) + hk / 3.0 * strain_norm2(st.ep);
  return kReturnOk;
}

// ---------------------------------------------------------------------------
// The general closest-point return (Plasticity.cpp:246-471)
// ---------------------------------------------------------------------------
struct GenPoint {
  double eq;        // sigma_bar(xi)
  double inv_eq;    // 1 / sigma_bar
  double m[6];      // P xi / sigma_bar (engineering)
  double b, db;
  double slope;     // the yield stress's slope at abar + dl
  double back_rate[6];
  double res[7];
};

__device__ __forceinline__ void decay(const DevMaterial& mat, int i, double dl, double& theta, double& chi,
                                      double& dtheta, double& dchi) {
  const double gm = mat.rec[i];
  if (gm == 0.0) {
    theta = 1.0; chi = dl; dtheta = 0.0; dchi = 1.0;
    return;
  }
  if (mat.exponential) {
    // exp(-gm dl) as 1 + expm1(-gm dl): one transcendental for both, theta
    // within an absolute 1e-16 (it scales backstresses in sums of stresses)
    const double em1 = expm1(-gm * dl);
    theta = 1.0 + em1;
    chi = -em1 * (1.0 / gm);  // 1 / gm: the same on every iteration
    dtheta = -gm * theta;
    dchi = theta;
  } else {
    theta = 1.0 / (1.0 + gm * dl);
    chi = theta * dl;
    dtheta = -gm * theta * theta;
    dchi = theta * theta;
  }
}

__device__ __forceinline__ void p_times(const DevMaterial& m, const double x[6], double out[6]) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = m.P[i * 6] * x[0];
#pragma unroll
    for (int j = 1; j < 6; ++j) s = s + m.P[i * 6 + j] * x[j];
    out[i] = s;
  }
}

template <int NT>
__device__ __forceinline__ bool gen_evaluate(const DevMaterial& mat, const PointStateT<NT>& c, const double s_tr[6],
                                             const double xi[6], double dl, GenPoint& e) {
  double px[6];
  p_times(mat, xi, px);
  double q = xi[0] * px[0];
#pragma unroll
  for (int i = 1; i < 6; ++i) q = q + xi[i] * px[i];
  if (!(q > 0.0) || !isfinite(q)) return false;
  // sigma_bar and its reciprocal from one inverse square root (each within an
  // ulp); the reciprocal serves the seven quotients by sigma_bar (m and, in
  // jac_setup, r). A q below 1e-290 (no stress state reaches it) takes the
  // IEEE operations.
  if (q >= 1.0e-290) {
    e.inv_eq = rsqrt_positive(q);
    e.eq = q * e.inv_eq;
  } else {
    e.eq = sqrt(q);
    e.inv_eq = 1.0 / e.eq;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) e.m[i] = px[i] * e.inv_eq;
  double xt[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) { xt[i] = s_tr[i]; e.back_rate[i] = 0.0; }
  e.b = 2.0 * mat.shear * dl;
  e.db = 2.0 * mat.shear;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t >= mat.terms) break;
    double th, chi, dth, dchi;
    decay(mat, t, dl, th, chi, dth, dchi);
#pragma unroll
    for (int i = 0; i < 6; ++i) xt[i] = xt[i] - th * c.kin[t][i];
    e.b = e.b + 2.0 / 3.0 * mat.mod[t] * chi;
    e.db = e.db + 2.0 / 3.0 * mat.mod[t] * dchi;
#pragma unroll
    for (int i = 0; i < 6; ++i) e.back_rate[i] = e.back_rate[i] + dth * c.kin[t][i];
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) e.res[i] = xi[i] - xt[i] + e.b * (i < 3 ? e.m[i] : e.m[i] * 0.5);
  double sy;
  hardening_at(mat, c.abar + dl, sy, e.slope);
  e.res[6] = e.eq - sy;
  return true;
}

// flow_curvature: S^-1 (P - m m^T)
__device__ __forceinline__ double flow_curv(const DevMaterial& mat, const GenPoint& e, int i, int j) {
  const double v = mat.P[i * 6 + j] - e.m[i] * e.m[j];
  return i < 3 ? v : v * 0.5;
}

// The closest-point Jacobian (Plasticity.cpp:278-300)
//   J = [[A, c], [m^T, -h]],  A = I + r S^-1 (P - m m^T),  r = b / sigma_bar,
//   c = back_rate + db S^-1 m,  h = the yield slope,  S^-1 = diag(1, 1, 1, 1/2, 1/2, 1/2),
// solved by blocks instead of a pivoted 7 x 7 LU: B = I + r S^-1 P =
// D V (I + r L) V^T D^-1 with D P D = V L V^T (the material's, from the host),
// A^-1 = B^-1 + B^-1 u m^T B^-1 / (1 - m^T B^-1 u) with u = r S^-1 m
// (Sherman-Morrison), and the Schur complement m^T A^-1 c + h for the plastic
// multiplier. A is similar to an SPD matrix with eigenvalues >= 1 (P - m m^T is
// the Hessian of the convex sigma_bar times sigma_bar), so no pivoting is needed.
struct JacSolver {
  double m[6];   // flow direction
  double d[6];   // 1 / (1 + r L_k)
  double z[6];   // B^-1 u
  double sm;     // 1 / (1 - m^T z)
  double wc[6];  // A^-1 c
  double schur;  // 1 / (m^T A^-1 c + h)
};

__device__ __forceinline__ double dot6(const double a[6], const double b[6]) {
  double s = a[0] * b[0];
#pragma unroll
  for (int i = 1; i < 6; ++i) s = s + a[i] * b[i];
  return s;
}

__device__ __forceinline__ void b_inverse(const DevMaterial& mat, const double d[6], const double v[6], double out[6]) {
  constexpr double kRoot2 = 1.4142135623730951;
  constexpr double kInvRoot2 = 0.7071067811865476;
  const double dv[6] = {v[0], v[1], v[2], kRoot2 * v[3], kRoot2 * v[4], kRoot2 * v[5]};
  double t[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double s = mat.V[k] * dv[0];
#pragma unroll
    for (int i = 1; i < 6; ++i) s = s + mat.V[i * 6 + k] * dv[i];
    t[k] = d[k] * s;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = mat.V[i * 6] * t[0];
#pragma unroll
    for (int k = 1; k < 6; ++k) s = s + mat.V[i * 6 + k] * t[k];
    out[i] = i < 3 ? s : kInvRoot2 * s;
  }
}

__device__ __forceinline__ void a_inverse(const DevMaterial& mat, const JacSolver& J, const double v[6], double out[6]) {
  b_inverse(mat, J.d, v, out);
  const double f = dot6(J.m, out) * J.sm;
#pragma unroll
  for (int i = 0; i < 6; ++i) out[i] = out[i] + J.z[i] * f;
}

// e: gen_evaluate's at the point's (xi, dl)
__device__ __forceinline__ void jac_setup(const DevMaterial& mat, const GenPoint& e, JacSolver& J) {
  const double r = e.b * e.inv_eq;
#pragma unroll
  for (int i = 0; i < 6; ++i) J.m[i] = e.m[i];
#pragma unroll
  for (int k = 0; k < 6; ++k) J.d[k] = recip_positive(1.0 + r * mat.L[k]);  // r >= 0, L_k >= 0
  double u[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) u[i] = r * (i < 3 ? e.m[i] : e.m[i] * 0.5);
  b_inverse(mat, J.d, u, J.z);
  J.sm = recip_positive(1.0 - dot6(J.m, J.z));  // 1 - m^T B^-1 u > 0 (A's eigenvalues >= 1)
  double c[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) c[i] = e.back_rate[i] + e.db * (i < 3 ? e.m[i] : e.m[i] * 0.5);
  a_inverse(mat, J, c, J.wc);
  J.schur = 1.0 / (dot6(J.m, J.wc) + e.slope);
}

// [x; y] = J^-1 [b1; b2]
__device__ __forceinline__ void jac_solve(const DevMaterial& mat, const JacSolver& J, const double b1[6], double b2,
                                          double x[6], double& y) {
  double w1[6];
  a_inverse(mat, J, b1, w1);
  y = (dot6(J.m, w1) - b2) * J.schur;
#pragma unroll
  for (int i = 0; i < 6; ++i) x[i] = w1[i] - J.wc[i] * y;
}

__device__ __forceinline__ double norm7(const double r[7]) {
  double s = r[0] * r[0];
#pragma unroll
  for (int i = 1; i < 7; ++i) s = s + r[i] * r[i];
  return sqrt(s);
}
__device__ __forceinline__ double maxabs7(const double r[7]) {
  double s = fabs(r[0]);
#pragma unroll
  for (int i = 1; i < 7; ++i) s = fmax(s, fabs(r[i]));
  return s;
}

// C_alg = C^e - 2G [m_t a^T + (dl / sigma_bar) S^-1 (P - m m^T) A],
// [A; a^T] = J^-1 [D_dev; 0] (Plasticity.cpp:302-311), one column at a time.
template <int NT>
__device__ __forceinline__ void gen_tangent(const DevMaterial& mat, const PointStateT<NT>& c, const GenPoint& e,
                                            double dl, double C[6][6]) {
  const double g = mat.shear;
  JacSolver J;
  jac_setup(mat, e, J);
  elastic_C(mat.bulk, g, C);
  const double g2 = 2.0 * g;
  const double s = 2.0 * g * dl * e.inv_eq;
#pragma unroll 1
  for (int j = 0; j < 6; ++j) {
    double b1[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      if (i < 3 && j < 3) b1[i] = (i == j) ? 4.0 * g / 3.0 : -2.0 * g / 3.0;
      else b1[i] = (i == j) ? g : 0.0;
    }
    double x[6], y;
    jac_solve(mat, J, b1, 0.0, x, y);
    double col[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double cij = 0.0;
#pragma unroll
      for (int jj = 0; jj < 6; ++jj) if (jj == j) cij = C[i][jj];
      const double mt = i < 3 ? e.m[i] : e.m[i] * 0.5;
      col[i] = cij - g2 * (mt * y);
    }
    if (dl > 0.0) {
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double acc = flow_curv(mat, e, i, 0) * x[0];
#pragma unroll
        for (int k = 1; k < 6; ++k) acc = acc + flow_curv(mat, e, i, k) * x[k];
        col[i] = col[i] - s * acc;
      }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int jj = 0; jj < 6; ++jj) if (jj == j) C[i][jj] = col[i];
  }
}

template <int NT>
__device__ __forceinline__ double kinematic_energy(const DevMaterial& mat, const PointStateT<NT>& st) {
  double out = 0.0;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t >= mat.terms) break;
    const double nrm = tensor_norm(st.kin[t]);
    out = out + 0.75 * nrm * nrm / mat.mod[t];
  }
  return out;
}

// The elastic trial of the general return (Plasticity.cpp:337-364).
struct GenTrial {
  double el[6], s_tr[6], xi_tr[6];
  double vol, sy_n, eq_tr, f_tr;
};

template <int NT>
__device__ __forceinline__ void gen_trial(const DevMaterial& mat, const double strain[6], const PointStateT<NT>& st,
                                          GenTrial& tr) {
  const double g = mat.shear;
#pragma unroll
  for (int i = 0; i < 6; ++i) tr.el[i] = strain[i] - st.ep[i];
  tr.vol = tr.el[0] + tr.el[1] + tr.el[2];
  const double mean = tr.vol / 3.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) tr.s_tr[i] = 2.0 * g * (tr.el[i] - mean);
#pragma unroll
  for (int i = 3; i < 6; ++i) tr.s_tr[i] = g * tr.el[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) tr.xi_tr[i] = tr.s_tr[i];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t >= mat.terms) break;
#pragma unroll
    for (int i = 0; i < 6; ++i) tr.xi_tr[i] = tr.xi_tr[i] - st.kin[t][i];
  }
  tr.sy_n = yield_stress(mat, st.abar);
  double px[6];
  p_times(mat, tr.xi_tr, px);
  double q = tr.xi_tr[0] * px[0];
#pragma unroll
  for (int i = 1; i < 6; ++i) q = q + tr.xi_tr[i] * px[i];
  tr.eq_tr = sqrt(fmax(q, 0.0));
  tr.f_tr = tr.eq_tr - tr.sy_n;
}

// The closest-point Newton from the trial (Plasticity.cpp:365-448): xi and
// dl converge with e the evaluation at the solution; false when the
// iteration does not converge.
template <int NT>
__device__ __forceinline__ bool gen_newton(const DevMaterial& mat, const PointStateT<NT>& st, const GenTrial& trial,
                                           double xi[6], double& dl, GenPoint& e) {
#pragma unroll
  for (int i = 0; i < 6; ++i) xi[i] = trial.xi_tr[i];
  dl = 0.0;
  gen_evaluate(mat, st, trial.s_tr, xi, dl, e);
  double scale = trial.sy_n;
#pragma unroll
  for (int i = 0; i < 6; ++i) scale = fmax(scale, fabs(trial.s_tr[i]));
#pragma unroll
  for (int i = 0; i < 6; ++i) scale = fmax(scale, fabs(trial.xi_tr[i]));
  const double eps_scale = kEps * scale;
  const double tol = fmax(1.0e-13 * trial.sy_n, 32.0 * eps_scale);
  bool conv = false;
  for (int it = 0; it < 50; ++it) {
    const double r_inf = maxabs7(e.res);
    if (r_inf <= tol) { conv = true; break; }
    double step[7][1];
    {
      JacSolver J;
      jac_setup(mat, e, J);
      double b1[6], x[6], y;
#pragma unroll
      for (int i = 0; i < 6; ++i) b1[i] = -e.res[i];
      jac_solve(mat, J, b1, -e.res[6], x, y);
#pragma unroll
      for (int i = 0; i < 6; ++i) step[i][0] = x[i];
      step[6][0] = y;
    }
    const double r0 = norm7(e.res);
    double t = 1.0;
    bool accepted = false;
    for (int ls = 0; ls < 40; ++ls) {
      double xt[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) xt[i] = xi[i] + t * step[i][0];
      const double dlt = fmax(dl + t * step[6][0], 0.0);
      GenPoint tr;
      if (gen_evaluate(mat, st, trial.s_tr, xt, dlt, tr) && norm7(tr.res) <= (1.0 - 1.0e-4 * t) * r0) {
#pragma unroll
        for (int i = 0; i < 6; ++i) xi[i] = xt[i];
        dl = dlt;
        e = tr;
        accepted = true;
        break;
      }
      t = t * 0.5;
    }
    if (!accepted) { conv = r_inf <= 4.0 * tol; break; }
  }
  if (!conv) return false;
  if (maxabs7(e.res) > 4.0 * eps_scale) {
    double step[7][1];
    {
      JacSolver J;
      jac_setup(mat, e, J);
      double b1[6], x[6], y;
#pragma unroll
      for (int i = 0; i < 6; ++i) b1[i] = -e.res[i];
      jac_solve(mat, J, b1, -e.res[6], x, y);
#pragma unroll
      for (int i = 0; i < 6; ++i) step[i][0] = x[i];
      step[6][0] = y;
    }
    double xp[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) xp[i] = xi[i] + step[i][0];
    const double dlp = fmax(dl + step[6][0], 0.0);
    GenPoint pol;
    if (gen_evaluate(mat, st, trial.s_tr, xp, dlp, pol) && maxabs7(pol.res) < maxabs7(e.res)) {
#pragma unroll
      for (int i = 0; i < 6; ++i) xi[i] = xp[i];
      dl = dlp;
      e = pol;
    }
  }
  return true;
}

// What the return of a point will cost at this strain, from its elastic
// trial alone (the tests that open return_3d and return_general_3d,
// Plasticity.cpp:87-108, 352-364): 2 for a plastic trial (the Newton
// iteration and the consistent tangent), 1 for an elastic trial that takes the
// continuum tangent of a loading point, 0 otherwise. Used only to order the
// elements of a launch by their expected work.
template <int NT>
__device__ __forceinline__ int return_cost(const DevMaterial& mat, const double strain[6], const PointStateT<NT>& st) {
  if (!mat.enabled) return 0;
  double f_tr, sy, eq;
  if (mat.general) {
    GenTrial trial;
    gen_trial(mat, strain, st, trial);
    f_tr = trial.f_tr;
    sy = trial.sy_n;
    eq = trial.eq_tr;
  } else {
    const double sqrt23 = 0.816496580927726;
    double el[6], xi[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) el[i] = strain[i] - st.ep[i];
    const double mean = (el[0] + el[1] + el[2]) / 3.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) xi[i] = (i < 3 ? 2.0 * mat.shear * (el[i] - mean) : mat.shear * el[i]) - st.kin[0][i];
    eq = tensor_norm(xi);
    sy = sqrt23 * yield_stress(mat, st.abar);
    f_tr = eq - sy;
  }
  if (!(f_tr <= 1.0e-12 * sy)) return 2;
  return (st.loading && f_tr > -1.0e-10 * sy && eq > 0.0) ? 1 : 0;
}

// return_general_3d; the state is updated in place (want_tangent = true, as
// IncompatibleModes.cpp:232 asks for it on every pass).
template <int NT>
__device__ int return_general(const DevMaterial& mat, const double strain[6], PointStateT<NT>& st, ReturnOut& out) {
  const double g = mat.shear, k = mat.bulk;
  GenTrial trial;
  gen_trial(mat, strain, st, trial);
  const double an = st.abar;
  if (!isfinite(trial.f_tr)) return kReturnNonFinite;
  const bool was_loading = st.loading;
  st.loading = false;
  if (!mat.enabled || trial.f_tr <= 1.0e-12 * trial.sy_n) {
#pragma unroll
    for (int i = 0; i < 6; ++i) out.stress[i] = trial.s_tr[i] + (i < 3 ? k * trial.vol : 0.0);
    elastic_C(k, g, out.C);
    out.energy = 0.5 * k * trial.vol * trial.vol + g * dev_norm2(trial.el) + isotropic_energy(mat, an) +
                 kinematic_energy(mat, st);
    out.yielding = false;
    if (mat.enabled && was_loading && trial.f_tr > -1.0e-10 * trial.sy_n && trial.eq_tr > 0.0) {
      st.loading = true;
      GenPoint e;
      if (gen_evaluate(mat, st, trial.s_tr, trial.xi_tr, 0.0, e)) gen_tangent(mat, st, e, 0.0, out.C);
    }
    return kReturnOk;
  }
  double xi[6], dl;
  GenPoint e;
  if (!gen_newton(mat, st, trial, xi, dl, e)) return kReturnNotConverged;
  out.yielding = true;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const double mt = i < 3 ? e.m[i] : e.m[i] * 0.5;
    out.stress[i] = trial.s_tr[i] - 2.0 * g * dl * mt + (i < 3 ? k * trial.vol : 0.0);
  }
  // tangent from the committed backstresses (the Jacobian's back_rate uses alpha_n)
  gen_tangent(mat, st, e, dl, out.C);
  st.loading = true;
  st.abar = an + dl;
#pragma unroll
  for (int i = 0; i < 6; ++i) st.ep[i] = st.ep[i] + dl * e.m[i];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t >= mat.terms) break;
    double th, chi, dth, dchi;
    decay(mat, t, dl, th, chi, dth, dchi);
    const double f = 2.0 / 3.0 * mat.mod[t] * chi;
#pragma unroll
    for (int i = 0; i < 6; ++i) st.kin[t][i] = th * st.kin[t][i] + f * (i < 3 ? e.m[i] : e.m[i] * 0.5);
  }
  double eln[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) eln[i] = trial.el[i] - dl * e.m[i];
  out.energy = 0.5 * k * trial.vol * trial.vol + g * dev_norm2(eln) + isotropic_energy(mat, st.abar) +
               kinematic_energy(mat, st);
  return kReturnOk;
}

}  // namespace sbk
