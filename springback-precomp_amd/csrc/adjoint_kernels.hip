// The adjoint of the element evaluation: vector-Jacobian products of the
// condensed internal force with respect to the nodal reference coordinates,
// the nodal displacements and the Hill48/Chaboche parameters, at the
// converged modes and the fixed committed history of the last evaluation.
//
// At the converged modes r_a(X, u, alpha, theta) = 0 defines alpha, and
// f* = r_u, so (implicit function theorem)
//   v^T df*/dy = v^T dr_u/dy - lambda^T dr_a/dy = w^T dr/dy,
//   lambda = K_aa^-T K_ua^T v,  w = [v; -lambda]  (33 entries),
// with r = sum_q w_q B_NL^T S the element's full residual
// (src/fem/IncompatibleModes.cpp:252-275) and K_aa, K_ua its blocks
// (:279-306). At fixed u, alpha and committed history, S depends on y only
// through the Green-Lagrange strain E, with dS/dE = the point's tangent (the
// "material" of :227-247), so
//   w^T dr/dy = sum_q [ d(w_q B_q w)/dy . S_q + w_q (B_q w) . (mat_q dE_q/dy) ].
// The coordinate and displacement derivatives of w_q, B_q and E_q are the
// tangent-linear forms of the set-up and the kinematics
// (IncompatibleModes.cpp:134-222) for each element DOF, formed on the point's
// values (point_products). The
// parameters enter through the stress alone: dS = P_log^T dT, and the general
// return's stress (src/material/Plasticity.cpp:337-471) moves with its
// converged (xi, dl) by d(xi, dl) = -J^-1 dR/dtheta, J the closest-point
// Jacobian, solved by the return's own block solver.
//
// One team of 8 lanes per element (lane = integration point for the point
// work, lane = node for the node rows), 8 elements per wavefront.
#include "adjoint_kernels.hpp"
#include "element_pass.hpp"
#include "team_linalg.hpp"
#include "workspace_records.hpp"

namespace sbk {

namespace {


// The point's share of w^T dr/dy for all 24 directions y of the element's
// coordinates (WRT_U false) or displacements (true), out[3 b + l] for node b,
// axis l. With the point's values formed once (the set-up and kinematics of
// IncompatibleModes.cpp:134-222: Jacobians, gradients g = [G | G~], h, F, E,
// w = w_q det J), each direction is linear algebra on them:
// * coordinates: the Jacobians change in row l only (dJ = e_l dn_b^T), so
//   d(ln det J) = t = g[l][b], d(ln det J0) = t0 = g0[l][b] (g0: the centre's
//   gradients), dJ^-1 = -J^-1 e_l g_b^T, and
//     dg[i][a] = -g[l][a] g[i][b]                       (a < 8),
//     dG~[i][k] = (t0 - t) G~[i][k] - G~[l][k] g0[i][b]  (G~ ~ det J0 / det J J0^-T),
//     dh = -hu[:, l] g_b^T + (t0 - t) ha - ha[:, l] g0_b^T, dw = w t,
//   with h = hu + ha (displacement and mode parts), hu[i][m] = sum_a u_a,i g[m][a];
// * displacements: only h moves, dh = e_l g_b^T.
// The residual side uses y = B_NL(F, g) w33 = voigt(F^T M), M[k][m] =
// sum_a w33[3 a + k] g[m][a] (so dM has the same form as dg), and
//   d(w_q B w33 . S) + w_q (B w33) . (mat dE) = (dw y + w dy) . S + (mat^T z) . dE,
// z = w y, dE from dh (Green-Lagrange: dE = sym(dh) + sym(h^T dh)).
template <int KIN, bool WRT_U>
__device__ void point_products(const double x0[8][3], const double ue[24], const double alpha[9], double xi,
                               double eta, double zeta, double wq, const double S[6], const double mat[6][6],
                               const double w33[33], double out[24]) {
  double dn0[8][3], dn[8][3];
  shape_gradients(0.0, 0.0, 0.0, dn0);
  shape_gradients(xi, eta, zeta, dn);
  double j0[3][3], jq[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double s0 = x0[0][i] * dn0[0][k], s = x0[0][i] * dn[0][k];
#pragma unroll
      for (int a = 1; a < 8; ++a) {
        s0 = s0 + x0[a][i] * dn0[a][k];
        s = s + x0[a][i] * dn[a][k];
      }
      j0[i][k] = s0;
      jq[i][k] = s;
    }
  const double det0 = det3(j0), det = det3(jq);
  double ji0[3][3], ji[3][3];
  inverse3(j0, ji0);
  inverse3(jq, ji);
  double g[3][8], g0[3][8], gm[3][3];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      g[i][a] = dn[a][0] * ji[0][i] + dn[a][1] * ji[1][i] + dn[a][2] * ji[2][i];
      g0[i][a] = dn0[a][0] * ji0[0][i] + dn0[a][1] * ji0[1][i] + dn0[a][2] * ji0[2][i];
    }
  const double ratio = det0 / det;
  const double nat[3] = {xi, eta, zeta};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double f = ratio * (-2.0 * nat[k]);
#pragma unroll
    for (int i = 0; i < 3; ++i) gm[i][k] = f * ji0[k][i];
  }
  // h = hu + ha; M = Mu + Mm (the w33-weighted gradients, u and mode columns)
  double hu[3][3], ha[3][3], h[3][3], mu[3][3], mm[3][3], M[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      double su = ue[i] * g[m][0], sw = w33[i] * g[m][0];
#pragma unroll
      for (int a = 1; a < 8; ++a) {
        su = su + ue[3 * a + i] * g[m][a];
        sw = sw + w33[3 * a + i] * g[m][a];
      }
      hu[i][m] = su;
      ha[i][m] = alpha[i] * gm[m][0] + alpha[3 + i] * gm[m][1] + alpha[6 + i] * gm[m][2];
      h[i][m] = su + ha[i][m];
      mu[i][m] = sw;
      mm[i][m] = w33[24 + i] * gm[m][0] + w33[27 + i] * gm[m][1] + w33[30 + i] * gm[m][2];
      M[i][m] = sw + mm[i][m];
    }
  const double w = wq * det;
  double F[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) F[i][k] = (KIN == kSmall ? 0.0 : h[i][k]) + (i == k ? 1.0 : 0.0);
  // y = voigt(F^T M), z = w y, q = mat^T z
  double P[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P[i][j] = F[0][i] * M[0][j] + F[1][i] * M[1][j] + F[2][i] * M[2][j];
  const double y[6] = {P[0][0], P[1][1], P[2][2], P[0][1] + P[1][0], P[1][2] + P[2][1], P[2][0] + P[0][2]};
  double q[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    double s = w * y[0] * mat[0][d];
#pragma unroll
    for (int c = 1; c < 6; ++c) s = s + w * y[c] * mat[c][d];
    q[d] = s;
  }
#pragma unroll
  for (int b = 0; b < 8; ++b)
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      double dh[3][3], dM[3][3], dw;
      if constexpr (WRT_U) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int m = 0; m < 3; ++m) { dh[i][m] = i == l ? g[m][b] : 0.0; dM[i][m] = 0.0; }
        dw = 0.0;
      } else {
        const double t = g[l][b], dlt = g0[l][b] - t;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            dh[i][m] = -g[m][b] * hu[i][l] + dlt * ha[i][m] - g0[m][b] * ha[i][l];
            dM[i][m] = -g[m][b] * mu[i][l] + dlt * mm[i][m] - g0[m][b] * mm[i][l];
          }
        dw = w * t;
      }
      double dE[6];
      if constexpr (KIN == kSmall) {
        dE[0] = dh[0][0];
        dE[1] = dh[1][1];
        dE[2] = dh[2][2];
        dE[3] = dh[0][1] + dh[1][0];
        dE[4] = dh[1][2] + dh[2][1];
        dE[5] = dh[2][0] + dh[0][2];
      } else {
        double de[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double s = dh[0][i] * h[0][k] + h[0][i] * dh[0][k];
#pragma unroll
            for (int m = 1; m < 3; ++m) s = s + dh[m][i] * h[m][k] + h[m][i] * dh[m][k];
            de[i][k] = 0.5 * ((dh[i][k] + dh[k][i]) + s);
          }
        dE[0] = de[0][0];
        dE[1] = de[1][1];
        dE[2] = de[2][2];
        dE[3] = 2.0 * de[0][1];
        dE[4] = 2.0 * de[1][2];
        dE[5] = 2.0 * de[2][0];
      }
      // dP = dF^T M + F^T dM (dF = dh; 0 for small strain)
      double dP[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = F[0][i] * dM[0][j] + F[1][i] * dM[1][j] + F[2][i] * dM[2][j];
          if constexpr (KIN != kSmall) s = s + dh[0][i] * M[0][j] + dh[1][i] * M[1][j] + dh[2][i] * M[2][j];
          dP[i][j] = s;
        }
      const double dy[6] = {dP[0][0], dP[1][1], dP[2][2], dP[0][1] + dP[1][0], dP[1][2] + dP[2][1],
                            dP[2][0] + dP[0][2]};
      double o = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) o = o + (dw * y[c] + w * dy[c]) * S[c] + q[c] * dE[c];
      out[3 * b + l] = o;
    }
}

// d(theta_t, chi_t)/d(gamma_t) of a backstress's decay (Plasticity.cpp:192-212).
__device__ __forceinline__ void decay_rate_derivative(const DevMaterial& mat, int t, double dl, double& dth,
                                                      double& dchi) {
  const double gm = mat.rec[t];
  if (mat.exponential) {
    if (gm == 0.0) {
      dth = -dl;
      dchi = -0.5 * dl * dl;
    } else {
      const double th = exp(-gm * dl);
      const double chi = -expm1(-gm * dl) / gm;
      dth = -dl * th;
      dchi = (dl * th - chi) / gm;
    }
  } else {
    const double th = 1.0 / (1.0 + gm * dl);
    dth = -dl * th * th;
    dchi = -dl * dl * th * th;
  }
}

// The point's share of w^T dr/dtheta for every parameter (out[kAdjointParams]).
template <int KIN>
__device__ void point_parameters(const AdjointArgs& Q, const double x0[8][3], const double ue[24],
                                 const double alpha[9], int p, const PointState& st, const double w33[33],
                                 double out[kAdjointParams]) {
  const DevMaterial& mat = Q.e.mat;
#pragma unroll
  for (int k = 0; k < kAdjointParams; ++k) out[k] = 0.0;
  if (!mat.general || !mat.enabled) return;
  PointGeom pg;
  double fv = 0.0;
  if (setup_point(x0, ue, Q.e.rule.xi[p], Q.e.rule.eta[p], Q.e.rule.zeta[p], Q.e.rule.w[p], pg, fv) != kOk) return;
  ModeGeom mg;
  mode_geom<KIN>(pg, ue, mg);
  double h[3][3], F[3][3];
  mode_kinematics<KIN>(mg.hu, mg.gm, alpha, h, F);
  double strain[6];
  LogData L;
  if constexpr (KIN == kSmall) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double s2 = bnl_mode(F, mg.gm, c, 0, 0) * alpha[0];
#pragma unroll
      for (int kk = 1; kk < 9; ++kk) s2 = s2 + bnl_mode(F, mg.gm, c, kk / 3, kk % 3) * alpha[kk];
      strain[c] = mg.eps_u[c] + s2;
    }
  } else {
    double e3[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double hth = h[0][i] * h[0][j] + h[1][i] * h[1][j] + h[2][i] * h[2][j];
        e3[i][j] = 0.5 * ((h[i][j] + h[j][i]) + hth);
      }
    double green[6] = {e3[0][0], e3[1][1], e3[2][2], 2.0 * e3[0][1], 2.0 * e3[1][2], 2.0 * e3[2][0]};
    if constexpr (KIN == kLog) {
      if (logarithmic_strain(green, L) != 0) return;
#pragma unroll
      for (int i = 0; i < 6; ++i) strain[i] = L.strain[i];
    } else {
#pragma unroll
      for (int i = 0; i < 6; ++i) strain[i] = green[i];
    }
  }
  GenTrial trial;
  gen_trial(mat, strain, st, trial);
  if (!(trial.f_tr > 1.0e-12 * trial.sy_n)) return;  // elastic: the stress does not depend on theta
  double xi[6], dl;
  GenPoint e;
  if (!gen_newton(mat, st, trial, xi, dl, e)) return;
  // z = w (B w) at the converged alpha
  double zc[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double y = bnl(F, pg.g, c, 0, 0) * w33[0];
#pragma unroll
    for (int col = 1; col < 33; ++col) y = y + bnl(F, pg.g, c, col / 3, col % 3) * w33[col];
    zc[c] = pg.w * y;
  }
  double lp[6][6];
  if constexpr (KIN == kLog) log_projection(L, lp);
  JacSolver J;
  jac_setup(mat, e, J);
  const double g2 = 2.0 * mat.shear;
  for (int k = 0; k < kAdjointParams; ++k) {
    const bool hill = k < 3;
    const int t = (k - 3) / 2;
    if (hill && !Q.hill) continue;
    if (!hill && t >= Q.backstresses) continue;
    double dR[7] = {0, 0, 0, 0, 0, 0, 0}, dpx[6] = {0, 0, 0, 0, 0, 0};
    double xdx = 0.0;
    if (hill) {
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double s = Q.dP[k][i * 6] * xi[0];
#pragma unroll
        for (int jj = 1; jj < 6; ++jj) s = s + Q.dP[k][i * 6 + jj] * xi[jj];
        dpx[i] = s;
      }
      xdx = dot6(xi, dpx);
      const double deq = xdx / (2.0 * e.eq);
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const double dm = dpx[i] / e.eq - e.m[i] * deq / e.eq;
        dR[i] = e.b * (i < 3 ? dm : 0.5 * dm);
      }
      dR[6] = deq;
    } else if ((k - 3) % 2 == 0) {  // C_t
      double th, chi, dth, dchi;
      decay(mat, t, dl, th, chi, dth, dchi);
#pragma unroll
      for (int i = 0; i < 6; ++i) dR[i] = 2.0 / 3.0 * chi * (i < 3 ? e.m[i] : 0.5 * e.m[i]);
    } else {  // gamma_t
      double dthg, dchig;
      decay_rate_derivative(mat, t, dl, dthg, dchig);
#pragma unroll
      for (int i = 0; i < 6; ++i)
        dR[i] = dthg * st.kin[t][i] + 2.0 / 3.0 * mat.mod[t] * dchig * (i < 3 ? e.m[i] : 0.5 * e.m[i]);
    }
    double b1[6], dxi[6], ddl;
#pragma unroll
    for (int i = 0; i < 6; ++i) b1[i] = -dR[i];
    jac_solve(mat, J, b1, -dR[6], dxi, ddl);
    // dm = (dP xi + P dxi) / eq - m deq / eq,  deq = xi.dP.xi / (2 eq) + m.dxi
    double pdx[6];
    p_times(mat, dxi, pdx);
    const double deq = xdx / (2.0 * e.eq) + dot6(e.m, dxi);
    double dT[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const double dm = (dpx[i] + pdx[i]) / e.eq - e.m[i] * deq / e.eq;
      const double mt = i < 3 ? e.m[i] : 0.5 * e.m[i];
      const double dmt = i < 3 ? dm : 0.5 * dm;
      dT[i] = -g2 * (ddl * mt + dl * dmt);
    }
    double dS[6];
    if constexpr (KIN == kLog) {
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double s = lp[0][i] * dT[0];
#pragma unroll
        for (int r = 1; r < 6; ++r) s = s + lp[r][i] * dT[r];
        dS[i] = s;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 6; ++i) dS[i] = dT[i];
    }
    out[k] = dot6(zc, dS);
  }
}

}  // namespace

// A team of TW lanes per element (8 for rules of up to 8 points, 32 for the
// 9..32-point box rules): lane p owns point p; lanes 0..7 also own node p's
// rows (K_ua's rows for lambda, the node's gradient entries).
template <int KIN, int TW>
__global__ __launch_bounds__(64) void adjoint_kernel(AdjointArgs Q) {
  constexpr int kTeams = 64 / TW;
  __shared__ LuSlot slots[kTeams];
  const ElementArgs& A = Q.e;
  const int lane = threadIdx.x;
  const int t = lane / TW;
  const int p = lane % TW;
  LuSlot& T = slots[t];
  const long long e = static_cast<long long>(blockIdx.x) * kTeams + t;
  if (e >= A.ne) return;
  const int n = A.rule.n;
  const long long q0 = e * n;
  if (A.status[e] != kOk) {
    // a failed element contributes nothing
    if (p < 8)
      for (int k = 0; k < 3; ++k) {
        if (Q.grad_x) Q.grad_x[e * 24 + 3 * p + k] = 0.0;
        if (Q.grad_u) Q.grad_u[e * 24 + 3 * p + k] = 0.0;
      }
    if (Q.grad_p)
      for (int k = p; k < kAdjointParams; k += TW) Q.grad_p[k * A.ne + e] = 0.0;
    return;
  }
  // K_aa's factors of the final local iteration
  const double* er = A.elem_ws + e * kElemRecord;
  for (int i = p; i < 81; i += TW) T.a[i] = er[kErFactors + i];
  for (int i = p; i < 9; i += TW) {
    T.rt[i] = static_cast<int>(er[kErRows + i]);
    T.ct[i] = static_cast<int>(er[kErCols + i]);
  }
  team_sync();
  double x0[8][3], ue[24];
  load_element(A.coords, A.conn, A.u, e, x0, ue);
  // the geometry records of the element's points at the converged modes (F,
  // [G | G~], w), which the coupling rows read
  if (p < n) {
    PointGeom pg;
    double fv = 0.0;
    setup_point(x0, ue, A.rule.xi[p], A.rule.eta[p], A.rule.zeta[p], A.rule.w[p], pg, fv);
    double alpha[9], gm[3][3], h[3][3], F[3][3];
#pragma unroll
    for (int i = 0; i < 9; ++i) alpha[i] = A.alpha_out[e * 9 + i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) gm[i][j] = pg.g[i][8 + j];
    mode_kinematics<KIN>(pg.hu, gm, alpha, h, F);
    double* gr = A.geom_ws + (q0 + p) * kGeomRecord;
#pragma unroll
    for (int i = 0; i < 9; ++i) gr[kGrF + i] = F[i / 3][i % 3];
#pragma unroll
    for (int i = 0; i < 33; ++i) gr[kGrG + i] = pg.g[i / 11][i % 11];
    gr[kGrW] = pg.w;
  }
  team_sync_global();
  double w33[33];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const long long node = A.conn[e * 8 + a];
#pragma unroll
    for (int k = 0; k < 3; ++k) w33[3 * a + k] = Q.v[node * 3 + k];
  }
  {
    // lambda = K_aa^-T K_ua^T v_e (lane a < 8 holds K_ua's rows 3a..3a+2)
    double ru[3], kua[3][9];
    const int a = p < 8 ? p : 0;
    coupling_rows<KIN>(A, q0, n, a, ru, kua);
    double g9[9], lam[9];
#pragma unroll
    for (int nn = 0; nn < 9; ++nn) {
      double s = kua[0][nn] * w33[3 * a];
      s = s + kua[1][nn] * w33[3 * a + 1];
      s = s + kua[2][nn] * w33[3 * a + 2];
      g9[nn] = team_sum<TW>(p < 8 ? s : 0.0);
    }
    lu_solve_transposed(T, g9, lam);
#pragma unroll
    for (int nn = 0; nn < 9; ++nn) w33[24 + nn] = -lam[nn];
  }
  double alpha[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) alpha[i] = A.alpha_out[e * 9 + i];
  const bool has_point = p < n;
  double S[6] = {0, 0, 0, 0, 0, 0}, mat[6][6];
  {
    const double* pr = A.point_ws + (q0 + (has_point ? p : 0)) * kPointRecord;
#pragma unroll
    for (int c = 0; c < 6; ++c) S[c] = has_point ? pr[kWsStress + c] : 0.0;
#pragma unroll
    for (int i = 0; i < 36; ++i) mat[i / 6][i % 6] = has_point ? pr[kWsTangent + i] : 0.0;
  }
  const double xi = has_point ? A.rule.xi[p] : 0.0, eta = has_point ? A.rule.eta[p] : 0.0;
  const double zeta = has_point ? A.rule.zeta[p] : 0.0, wq = has_point ? A.rule.w[p] : 0.0;
  // the team's sums of the 24 directions: lane b keeps node b's three
  if (Q.wrt & kWrtCoords) {
    double c[24];
    if (has_point) {
      point_products<KIN, false>(x0, ue, alpha, xi, eta, zeta, wq, S, mat, w33, c);
    } else {
#pragma unroll
      for (int j = 0; j < 24; ++j) c[j] = 0.0;
    }
    double mine[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 24; ++j) {
      const double tot = team_sum<TW>(c[j]);
      if (j / 3 == p) mine[j % 3] = tot;
    }
    if (p < 8)
#pragma unroll
      for (int k = 0; k < 3; ++k) Q.grad_x[e * 24 + 3 * p + k] = mine[k];
  }
  if (Q.wrt & kWrtDisplacement) {
    double c[24];
    if (has_point) {
      point_products<KIN, true>(x0, ue, alpha, xi, eta, zeta, wq, S, mat, w33, c);
    } else {
#pragma unroll
      for (int j = 0; j < 24; ++j) c[j] = 0.0;
    }
    double mine[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 24; ++j) {
      const double tot = team_sum<TW>(c[j]);
      if (j / 3 == p) mine[j % 3] = tot;
    }
    if (p < 8)
#pragma unroll
      for (int k = 0; k < 3; ++k) Q.grad_u[e * 24 + 3 * p + k] = mine[k];
  }
  if (Q.wrt & kWrtParams) {
    double c[kAdjointParams];
    if (has_point) {
      PointState st;
      load_state(A.mat, A.state_in, A.state_stride, q0 + p, st);
      point_parameters<KIN>(Q, x0, ue, alpha, p, st, w33, c);
    } else {
#pragma unroll
      for (int k = 0; k < kAdjointParams; ++k) c[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < kAdjointParams; ++k) {
      const double tot = team_sum<TW>(c[k]);
      if (p == k % TW) Q.grad_p[k * A.ne + e] = tot;
    }
  }
}

template <int TW>
void launch_adjoint_tw(const AdjointArgs& a, hipStream_t stream) {
  constexpr int teams = 64 / TW;
  const unsigned int blocks = static_cast<unsigned int>((a.e.ne + teams - 1) / teams);
  switch (a.e.kinematics) {
    case kSmall: hipLaunchKernelGGL((adjoint_kernel<kSmall, TW>), dim3(blocks), dim3(64), 0, stream, a); break;
    case kFinite: hipLaunchKernelGGL((adjoint_kernel<kFinite, TW>), dim3(blocks), dim3(64), 0, stream, a); break;
    default: hipLaunchKernelGGL((adjoint_kernel<kLog, TW>), dim3(blocks), dim3(64), 0, stream, a); break;
  }
}

int launch_adjoint(const AdjointArgs& a, hipStream_t stream) {
  (void)hipGetLastError();
  if (a.e.ne == 0) return 0;
  if (a.e.rule.n < 1 || a.e.rule.n > kMaxPoints) return static_cast<int>(hipErrorInvalidValue);
  if (a.e.rule.n <= kTeam)
    launch_adjoint_tw<kTeam>(a, stream);
  else
    launch_adjoint_tw<kWideTeam>(a, stream);
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
