// Device building blocks of the incompatible-mode Hex8 kernels (gfx950, FP64).
//
// Execution model: a team of 8 consecutive lanes of a 64-wide wavefront owns
// one element; lane p of the team owns integration point p of the element's
// rule (rules of up to 8 points; rules of 9 to 32 points, the reference's
// 2 x 2 x m box rules, take teams of 32 lanes). Team-wide sums are DPP
// butterflies inside the team (quad_perm xor 1, quad_perm xor 2,
// row_half_mirror; for 32 lanes also row_mirror and a swizzle between the
// two rows of 16), so every lane of a team ends with the bitwise same total,
// in a fixed order.
//
// The constitutive algorithms follow the reference's material point:
//   src/material/Material.cpp:73-90 (hardening), src/material/Plasticity.cpp
//   (J2 radial return :66-171, general closest-point return :185-471),
//   src/material/LogarithmicStrain.cpp:29-197 (Hencky strain, P, T:L),
//   src/fem/IncompatibleModes.cpp:101-121, 195-275 (kinematics, sums).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sbk {

constexpr int kTeam = 8;           // lanes per element (rules of up to 8 points)
constexpr int kWideTeam = 32;      // lanes per element for rules of 9..32 points
constexpr int kMaxPoints = 32;     // integration points per element
constexpr int kMaxTerms = 4;       // backstresses (include/sparlab/material/Material.hpp:68)
constexpr double kEps = 2.220446049250313e-16;

// 1 / x for a positive normal x far from the overflow threshold (the
// arguments below are >= 1 or bounded sums of positive terms): the hardware's
// reciprocal estimate (v_rcp_f64) refined by two Newton-Raphson steps,
// r <- r + r (1 - x r), each squaring the relative error; within an ulp of
// the correctly rounded quotient, in 5 dependent instructions instead of the
// IEEE division's 11 (which scales for denormal and huge operands).
__device__ __forceinline__ double recip_positive(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  r = fma(fma(-x, r, 1.0), r, r);
  return r;
}

// 1 / sqrt(x) for a positive normal x (not below 1e-290): the hardware's
// estimate (v_rsq_f64) refined by two Newton-Raphson steps for x r^2 = 1,
// r <- r + r (1/2 - (x/2) r^2); within an ulp of the rounded 1 / sqrt(x).
__device__ __forceinline__ double rsqrt_positive(double x) {
  double r = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  r = fma(r, fma(-h * r, r, 0.5), r);
  r = fma(r, fma(-h * r, r, 0.5), r);
  return r;
}

// ---------------------------------------------------------------------------
// Material, broadcast to every lane as a kernel argument
// ---------------------------------------------------------------------------
struct DevMaterial {
  double shear, bulk;            // G, K = lambda + 2G/3
  double sy0, hard, hkin, qsat, delta;
  double P[36];                  // yield matrix, tensorial Voigt, global axes
  // D P D = V diag(L) V^T with D = diag(1, 1, 1, 2^-1/2, 2^-1/2, 2^-1/2):
  // V row-major (column k is eigenvector k), L the eigenvalues
  double V[36], L[6];
  double mod[kMaxTerms], rec[kMaxTerms];  // kinematic terms: Chaboche, then Prager
  int general;                   // Hill48 or backstresses: closest-point return
  int terms;                     // kinematic terms of the general return
  int exponential;               // 1 exponential, 0 backward-Euler integration
  int enabled;                   // yield stress > 0
  int symmetric;                 // no recovering backstress
  int kin_fields;                // 6-vectors of kinematic history per point
};

// Per-point state record in structure-of-arrays order (field f of point q at
// state[f * stride + q]): plastic strain (6, engineering shears), then
// kin_fields 6-vectors (the radial return's beta, or the general return's
// backstresses), the equivalent plastic strain and the loading flag (0/1).
__host__ __device__ constexpr int state_fields(int kin_fields) { return 8 + 6 * kin_fields; }

__device__ __forceinline__ double yield_stress(const DevMaterial& m, double a) {
  return m.sy0 + m.hard * a + m.qsat * -expm1(-m.delta * a);
}
__device__ __forceinline__ double yield_slope(const DevMaterial& m, double a) {
  return m.hard + m.qsat * m.delta * exp(-m.delta * a);
}
// The yield stress and its slope at a from one exponential: the stress as
// yield_stress, the slope with exp(-delta a) = 1 + expm1(-delta a) (within an
// absolute 1e-16).
__device__ __forceinline__ void hardening_at(const DevMaterial& m, double a, double& sy, double& slope) {
  const double em1 = expm1(-m.delta * a);
  sy = m.sy0 + m.hard * a + m.qsat * -em1;
  slope = m.hard + m.qsat * m.delta * (1.0 + em1);
}
__device__ __forceinline__ double isotropic_energy(const DevMaterial& m, double a) {
  double out = 0.5 * m.hard * a * a;
  if (m.qsat != 0.0) out = out + m.qsat * (a + expm1(-m.delta * a) / m.delta);
  return out;
}

// ---------------------------------------------------------------------------
// Team (8-lane) reductions
// ---------------------------------------------------------------------------
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  int lo = static_cast<int>(b);
  int hi = static_cast<int>(b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, kCtrl, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, kCtrl, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <int kCtrl>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xF, 0xF, false);
}
constexpr int kXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kHalfMirror = 0x141; // row_half_mirror: lane i <-> 7 - i in each 8

constexpr int kRowMirror = 0x140;  // row_mirror: lane i <-> 15 - i in each 16
// ds_swizzle in bit mode over 32 lanes: lane i reads lane i ^ 16
constexpr int kSwap16 = 0x1F | (0x10 << 10);

__device__ __forceinline__ double swap16_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_swizzle(static_cast<int>(b), kSwap16);
  const int hi = __builtin_amdgcn_ds_swizzle(static_cast<int>(b >> 32), kSwap16);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ int swap16_i32(int v) { return __builtin_amdgcn_ds_swizzle(v, kSwap16); }

// A team is TW consecutive lanes (8 or 32) of the wavefront; every lane of
// the team receives the same result, in a fixed order (commutativity makes
// each butterfly step symmetric). The 8-lane steps are quad_perm xor 1, xor
// 2 and row_half_mirror; 32 lanes add row_mirror and a swap of the two rows
// of 16.
template <int TW = kTeam, class Op>
__device__ __forceinline__ double team_fold(double v, Op op) {
  v = op(v, dpp_f64<kXor1>(v));
  v = op(v, dpp_f64<kXor2>(v));
  v = op(v, dpp_f64<kHalfMirror>(v));
  if constexpr (TW == kWideTeam) {
    v = op(v, dpp_f64<kRowMirror>(v));
    v = op(v, swap16_f64(v));
  }
  return v;
}
template <int TW = kTeam, class Op>
__device__ __forceinline__ int team_ifold(int v, Op op) {
  v = op(v, dpp_i32<kXor1>(v));
  v = op(v, dpp_i32<kXor2>(v));
  v = op(v, dpp_i32<kHalfMirror>(v));
  if constexpr (TW == kWideTeam) {
    v = op(v, dpp_i32<kRowMirror>(v));
    v = op(v, swap16_i32(v));
  }
  return v;
}
template <int TW = kTeam>
__device__ __forceinline__ double team_sum(double v) {
  return team_fold<TW>(v, [](double a, double b) { return a + b; });
}
template <int TW = kTeam>
__device__ __forceinline__ int team_or(int v) {
  return team_ifold<TW>(v, [](int a, int b) { return a | b; });
}
template <int TW = kTeam>
__device__ __forceinline__ double team_max(double v) {
  return team_fold<TW>(v, [](double a, double b) { return fmax(a, b); });
}
template <int TW = kTeam>
__device__ __forceinline__ double team_min(double v) {
  return team_fold<TW>(v, [](double a, double b) { return fmin(a, b); });
}
template <int TW = kTeam>
__device__ __forceinline__ int team_isum(int v) {
  return team_ifold<TW>(v, [](int a, int b) { return a + b; });
}
template <int TW = kTeam>
__device__ __forceinline__ int team_imin(int v) {
  return team_ifold<TW>(v, [](int a, int b) { return min(a, b); });
}
// Orders the team's lanes' LDS traffic. A team lives in one wavefront, whose
// LDS instructions execute in issue order, so wavefront-scope fences suffice:
// they keep the compiler from moving LDS accesses across and cost no waits.
__device__ __forceinline__ void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Orders global-memory traffic between the lanes of a team (the stores are
// waited for, so other lanes' loads see them).
__device__ __forceinline__ void team_sync_global() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---------------------------------------------------------------------------
// Hex8 geometry
// ---------------------------------------------------------------------------
__device__ __forceinline__ double corner(int a, int k) {
  // VTK hexahedron corners (src/elements/Hex8.cpp:17-19)
  const int bits = (k == 0) ? ((a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0)
                 : (k == 1) ? ((a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0)
                            : (a >= 4 ? 1 : 0);
  return bits ? 1.0 : -1.0;
}

// dN_a/dxi_k at (xi, eta, zeta) (src/elements/Hex8.cpp:52-64)
__device__ __forceinline__ void shape_gradients(double xi, double eta, double zeta, double dn[8][3]) {
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const double fx = 1.0 + xi * corner(a, 0);
    const double fy = 1.0 + eta * corner(a, 1);
    const double fz = 1.0 + zeta * corner(a, 2);
    dn[a][0] = 0.125 * corner(a, 0) * fy * fz;
    dn[a][1] = 0.125 * corner(a, 1) * fx * fz;
    dn[a][2] = 0.125 * corner(a, 2) * fx * fy;
  }
}

__device__ __forceinline__ double det3(const double m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
         m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}
// Inverse by cofactors times 1/det, det from the first column's cofactors.
__device__ __forceinline__ void inverse3(const double m[3][3], double r[3][3]) {
  double c[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      c[i][j] = m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
    }
  const double det = c[0][0] * m[0][0] + c[1][0] * m[1][0] + c[2][0] * m[2][0];
  const double inv = 1.0 / det;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r[i][j] = c[j][i] * inv;
}

// ---------------------------------------------------------------------------
// Symmetric 3x3 eigen-decomposition: cyclic Jacobi, eigenvalues ascending,
// eigenvectors in the columns of v. A rotation is skipped once its
// off-diagonal entry no longer changes either diagonal entry.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void jacobi_rotate(double a[3][3], double v[3][3], int p, int q, int r,
                                              bool& any) {
  const double apq = a[p][q];
  if (apq == 0.0) return;
  const double g = 100.0 * fabs(apq);
  const double app = fabs(a[p][p]), aqq = fabs(a[q][q]);
  if (app + g == app && aqq + g == aqq) {
    a[p][q] = 0.0;
    a[q][p] = 0.0;
    return;
  }
  any = true;
  const double h = a[q][q] - a[p][p];
  double t;
  if (fabs(h) + g == fabs(h)) {
    t = apq / h;
  } else {
    const double theta = 0.5 * h / apq;
    const double q = 1.0 + theta * theta;  // >= 1
    t = recip_positive(fabs(theta) + q * rsqrt_positive(q));
    if (theta < 0.0) t = -t;
  }
  const double c = rsqrt_positive(1.0 + t * t);  // |t| <= 1: argument in [1, 2]
  const double s = t * c;
  const double tau = s * recip_positive(1.0 + c);
  a[p][p] = a[p][p] - t * apq;
  a[q][q] = a[q][q] + t * apq;
  a[p][q] = 0.0;
  a[q][p] = 0.0;
  const double gr = a[r][p], hr = a[r][q];
  a[r][p] = gr - s * (hr + gr * tau);
  a[p][r] = a[r][p];
  a[r][q] = hr + s * (gr - hr * tau);
  a[q][r] = a[r][q];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double gv = v[k][p], hv = v[k][q];
    v[k][p] = gv - s * (hv + gv * tau);
    v[k][q] = hv + s * (gv - hv * tau);
  }
}

__device__ __forceinline__ void eigen_sym3(const double in[3][3], double w[3], double v[3][3]) {
  double a[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) { a[i][j] = in[i][j]; v[i][j] = (i == j) ? 1.0 : 0.0; }
  for (int sweep = 0; sweep < 64; ++sweep) {
    bool any = false;
    jacobi_rotate(a, v, 0, 1, 2, any);
    jacobi_rotate(a, v, 0, 2, 1, any);
    jacobi_rotate(a, v, 1, 2, 0, any);
    if (!any) break;
  }
  w[0] = a[0][0];
  w[1] = a[1][1];
  w[2] = a[2][2];
  // ascending, columns follow (insertion sort on three)
  auto swap_cols = [&](int i, int j) {
    double t = w[i]; w[i] = w[j]; w[j] = t;
#pragma unroll
    for (int k = 0; k < 3; ++k) { double u = v[k][i]; v[k][i] = v[k][j]; v[k][j] = u; }
  };
  if (w[1] < w[0]) swap_cols(1, 0);
  if (w[2] < w[1]) {
    swap_cols(2, 1);
    if (w[1] < w[0]) swap_cols(1, 0);
  }
}

// ---------------------------------------------------------------------------
// Logarithmic strain of the Green-Lagrange strain E (engineering Voigt)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int vrow(int k) { return k < 3 ? k : k - 3; }
__device__ __forceinline__ int vcol(int k) { return k < 3 ? k : (k == 3 ? 1 : (k == 4 ? 2 : 0)); }

struct LogData {
  double q[3][3];    // principal axes (columns)
  double mu[3];      // principal Green-Lagrange strains
  double first[3][3];
  double strain[6];  // E_log, engineering
};

__device__ __forceinline__ double log_divided(double mx, double my) {
  if (mx > my) { double t = mx; mx = my; my = t; }
  const double y = 1.0 + 2.0 * my;  // a principal value of C: positive
  const double d = 2.0 * (mx - my);
  const double iy = recip_positive(y);
  if (d == 0.0) return iy;
  return log1p(d * iy) / d;
}

// The principal-axes map of tensorial Voigt components, t(Q^T A Q) = M t(A).
__device__ __forceinline__ void principal_rotation(const double q[3][3], double m[6][6]) {
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int a = vrow(r), b = vcol(r), i = vrow(c), j = vcol(c);
      m[r][c] = (i == j) ? q[i][a] * q[i][b] : q[i][a] * q[j][b] + q[j][a] * q[i][b];
    }
}

// Returns 0, or a failure code (non-finite E / non-positive principal stretch).
__device__ __forceinline__ int logarithmic_strain(const double e[6], LogData& L) {
#pragma unroll
  for (int k = 0; k < 6; ++k)
    if (!isfinite(e[k])) return 1;
  double t[3][3];
  t[0][0] = e[0]; t[1][1] = e[1]; t[2][2] = e[2];
  t[0][1] = t[1][0] = 0.5 * e[3];
  t[1][2] = t[2][1] = 0.5 * e[4];
  t[2][0] = t[0][2] = 0.5 * e[5];
  eigen_sym3(t, L.mu, L.q);
  double pr[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double s2 = 1.0 + 2.0 * L.mu[a];
    if (!(s2 > 0.0) || !isfinite(s2)) return 2;
    pr[a] = 0.5 * log1p(2.0 * L.mu[a]);
  }
  double qd[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) qd[i][a] = L.q[i][a] * pr[a];
  double el[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) el[i][j] = qd[i][0] * L.q[j][0] + qd[i][1] * L.q[j][1] + qd[i][2] * L.q[j][2];
  L.strain[0] = el[0][0]; L.strain[1] = el[1][1]; L.strain[2] = el[2][2];
  L.strain[3] = 2.0 * el[0][1]; L.strain[4] = 2.0 * el[1][2]; L.strain[5] = 2.0 * el[2][0];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b) L.first[a][b] = L.first[b][a] = log_divided(L.mu[a], L.mu[b]);
  return 0;
}

// The projection dE_log = P dE (LogarithmicStrain.cpp:128-141), formed from the
// axes and first divided differences when it is needed (after the return, so
// it is not held across it).
__device__ __forceinline__ void log_projection(const LogData& L, double P[6][6]) {
  double m[6][6];
  principal_rotation(L.q, m);
  double sd[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sd[k] = (k < 3 ? 1.0 : 2.0) * L.first[vrow(k)][vcol(k)];
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double s = m[0][r] * sd[0] * m[0][c];
#pragma unroll
      for (int k = 1; k < 6; ++k) s = s + m[k][r] * sd[k] * m[k][c];
      P[r][c] = c < 3 ? s : s * 0.5;
    }
}

__device__ __forceinline__ double log_divided2(const LogData& L, int a, int b, int c) {
  int i = a, j = b, k = c;
  if (L.mu[i] < L.mu[j]) { int t = i; i = j; j = t; }
  if (L.mu[j] < L.mu[k]) { int t = j; j = k; k = t; }
  if (L.mu[i] < L.mu[j]) { int t = i; i = j; j = t; }
  double mi = 0, mj = 0, mk = 0, fij = 0, fjk = 0;
  // select by index without dynamic register indexing
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    if (s == i) mi = L.mu[s];
    if (s == j) mj = L.mu[s];
    if (s == k) mk = L.mu[s];
  }
#pragma unroll
  for (int s = 0; s < 3; ++s)
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      if (s == i && t == j) fij = L.first[s][t];
      if (s == j && t == k) fjk = L.first[s][t];
    }
  const double y = 1.0 + 2.0 * mj;
  const double spread = 2.0 * (mi - mk);
  if (spread > 5.0e-3 * y) return (fij - fjk) / spread;
  // the series' quotients as products with 1 / y (y = 1 + 2 mu_j, a
  // principal value of C: positive) and with 1 / (n + 2) rounded, each within
  // an ulp of the quotient
  constexpr double kInv[6] = {0.0, 1.0 / 3.0, 1.0 / 4.0, 1.0 / 5.0, 1.0 / 6.0, 1.0 / 7.0};
  const double iy = recip_positive(y);
  const double u = 2.0 * (mi - mj) * iy;
  const double w = 2.0 * (mk - mj) * iy;
  double h = 1.0, wn = 1.0, sum = -0.5;
#pragma unroll
  for (int n = 1; n <= 5; ++n) {
    wn = wn * w;
    h = u * h + wn;
    sum = sum + ((n % 2 == 1) ? 1.0 : -1.0) * h * kInv[n];
  }
  return sum * (iy * iy);
}

// The stress conjugate to E (S = P^T T, logarithmic_stress) and the tangent
// between Green-Lagrange strains, mat = P^T C P + T:L (IncompatibleModes.cpp:
// 238-243, LogarithmicStrain.cpp:144-197), formed in the principal frame: with
// M the principal rotation of tensorial Voigt components, D = diag(sd) the
// first divided differences (engineering) and Ms = M diag(1,1,1,1/2,1/2,1/2),
// P = M^T D Ms and T:L = Ms^T H Ms, so
//   mat = Ms^T [D (M C M^T) D + H] Ms,
// with H the principal-frame curvature. Four 6 x 6 products instead of seven,
// and at most three 6 x 6 arrays live at a time.
__device__ __forceinline__ void log_material(const LogData& L, const double t[6], const double C[6][6], double S[6],
                                             double mat[6][6]) {
  double tt[3][3];
  tt[0][0] = t[0]; tt[1][1] = t[1]; tt[2][2] = t[2];
  tt[0][1] = tt[1][0] = t[3];
  tt[1][2] = tt[2][1] = t[4];
  tt[2][0] = tt[0][2] = t[5];
  double qt[3][3], tp[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) qt[i][j] = L.q[0][i] * tt[0][j] + L.q[1][i] * tt[1][j] + L.q[2][i] * tt[2][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) tp[i][j] = qt[i][0] * L.q[0][j] + qt[i][1] * L.q[1][j] + qt[i][2] * L.q[2][j];
  {
    // S = Q (tp o first) Q^T
    double pr[3][3], qp[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) pr[i][j] = tp[i][j] * L.first[i][j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) qp[i][j] = L.q[i][0] * pr[0][j] + L.q[i][1] * pr[1][j] + L.q[i][2] * pr[2][j];
    const int si[6] = {0, 1, 2, 0, 1, 2}, sj[6] = {0, 1, 2, 1, 2, 0};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int i = si[k], j = sj[k];
      S[k] = qp[i][0] * L.q[j][0] + qp[i][1] * L.q[j][1] + qp[i][2] * L.q[j][2];
    }
  }
  double m[6][6];
  principal_rotation(L.q, m);
  // X = M C M^T
  double x[6][6];
  {
    double y[6][6];
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int l = 0; l < 6; ++l) {
        double s = m[a][0] * C[0][l];
#pragma unroll
        for (int k = 1; k < 6; ++k) s = s + m[a][k] * C[k][l];
        y[a][l] = s;
      }
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b < 6; ++b) {
        double s = y[a][0] * m[b][0];
#pragma unroll
        for (int l = 1; l < 6; ++l) s = s + y[a][l] * m[b][l];
        x[a][b] = s;
      }
  }
  // X <- D X D + H; H(k, l) = 2 (half(k, l) + half(l, k)), half(x, y) = the sum
  // over the principal unit tensors, sum_{(a,c) in X, (c,b) in Y} tp(a,b) l[a,b,c]
  // with l the second divided differences (ten distinct values, a <= b <= c)
  double l2[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
#pragma unroll
      for (int c = b; c < 3; ++c) {
        const double v = log_divided2(L, a, b, c);
        l2[a][b][c] = l2[a][c][b] = l2[b][a][c] = l2[b][c][a] = l2[c][a][b] = l2[c][b][a] = v;
      }
  double sd[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sd[k] = (k < 3 ? 1.0 : 2.0) * L.first[vrow(k)][vcol(k)];
#pragma unroll
  for (int kx = 0; kx < 6; ++kx)
#pragma unroll
    for (int ky = 0; ky < 6; ++ky) {
      auto half = [&](int u, int v) {
        const int xr[2] = {vrow(u), vcol(u)}, xc[2] = {vcol(u), vrow(u)};
        const int yr[2] = {vrow(v), vcol(v)}, yc[2] = {vcol(v), vrow(v)};
        const int nx = u < 3 ? 1 : 2, ny = v < 3 ? 1 : 2;
        double sum = 0.0;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          if (s >= nx) continue;
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            if (r >= ny) continue;
            if (yr[r] != xc[s]) continue;
            sum = sum + tp[xr[s]][yc[r]] * l2[xr[s]][yc[r]][xc[s]];
          }
        }
        return sum;
      };
      const double h = 2.0 * (half(kx, ky) + half(ky, kx));
      x[kx][ky] = sd[kx] * x[kx][ky] * sd[ky] + h;
    }
  // mat = Ms^T X Ms
  double z[6][6];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      double s = m[0][i] * x[0][b];
#pragma unroll
      for (int a = 1; a < 6; ++a) s = s + m[a][i] * x[a][b];
      z[i][b] = s;
    }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double s = z[i][0] * m[0][j];
#pragma unroll
      for (int b = 1; b < 6; ++b) s = s + z[i][b] * m[b][j];
      const double f = (i < 3 ? 1.0 : 0.5) * (j < 3 ? 1.0 : 0.5);
      mat[i][j] = f * s;
    }
}

// T' = Q^T T Q: a tensorial-Voigt stress in the principal axes.
__device__ __forceinline__ void principal_stress(const LogData& L, const double t[6], double tp[3][3]) {
  double tt[3][3];
  tt[0][0] = t[0]; tt[1][1] = t[1]; tt[2][2] = t[2];
  tt[0][1] = tt[1][0] = t[3];
  tt[1][2] = tt[2][1] = t[4];
  tt[2][0] = tt[0][2] = t[5];
  double qt[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) qt[i][j] = L.q[0][i] * tt[0][j] + L.q[1][i] * tt[1][j] + L.q[2][i] * tt[2][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) tp[i][j] = qt[i][0] * L.q[0][j] + qt[i][1] * L.q[1][j] + qt[i][2] * L.q[2][j];
}

// S = Q (T' o first) Q^T from the principal-axes stress T' (logarithmic_stress,
// LogarithmicStrain.cpp:144-149).
__device__ __forceinline__ void log_stress_from_principal(const LogData& L, const double tp[3][3], double S[6]) {
  double pr[3][3], qp[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) pr[i][j] = tp[i][j] * L.first[i][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) qp[i][j] = L.q[i][0] * pr[0][j] + L.q[i][1] * pr[1][j] + L.q[i][2] * pr[2][j];
  double sm[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) sm[i][j] = qp[i][0] * L.q[j][0] + qp[i][1] * L.q[j][1] + qp[i][2] * L.q[j][2];
  S[0] = sm[0][0]; S[1] = sm[1][1]; S[2] = sm[2][2]; S[3] = sm[0][1]; S[4] = sm[1][2]; S[5] = sm[2][0];
}

// S = P^T T (logarithmic_stress).
__device__ __forceinline__ void log_stress(const LogData& L, const double t[6], double S[6]) {
  double tp[3][3];
  principal_stress(L, t, tp);
  log_stress_from_principal(L, tp, S);
}

// S = P^T T (logarithmic_stress) and T:L (logarithmic_curvature) in one go.
__device__ __forceinline__ void log_push(const LogData& L, const double t[6], double S[6], double curv[6][6]) {
  double tp[3][3];
  principal_stress(L, t, tp);
  log_stress_from_principal(L, tp, S);
  // second divided differences: ten distinct values l[a,b,c], a <= b <= c
  double l2[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
#pragma unroll
      for (int c = b; c < 3; ++c) {
        const double v = log_divided2(L, a, b, c);
        l2[a][b][c] = l2[a][c][b] = l2[b][a][c] = l2[b][c][a] = l2[c][a][b] = l2[c][b][a] = v;
      }
  // hp(k, l) = 2 (half(k, l) + half(l, k)); half(x, y) = sum over the principal
  // unit tensors: sum_{(a,c) in X, (c,b) in Y} tp(a,b) l[a,b,c]
  double hp[6][6];
#pragma unroll
  for (int kx = 0; kx < 6; ++kx)
#pragma unroll
    for (int ky = kx; ky < 6; ++ky) {
      auto half = [&](int x, int y) {
        const int xr[2] = {vrow(x), vcol(x)}, xc[2] = {vcol(x), vrow(x)};
        const int yr[2] = {vrow(y), vcol(y)}, yc[2] = {vcol(y), vrow(y)};
        const int nx = x < 3 ? 1 : 2, ny = y < 3 ? 1 : 2;
        double sum = 0.0;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          if (s >= nx) continue;
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            if (r >= ny) continue;
            if (yr[r] != xc[s]) continue;
            sum = sum + tp[xr[s]][yc[r]] * l2[xr[s]][yc[r]][xc[s]];
          }
        }
        return sum;
      };
      const double v = 2.0 * (half(kx, ky) + half(ky, kx));
      hp[kx][ky] = v;
      hp[ky][kx] = v;
    }
  double m[6][6];
  principal_rotation(L.q, m);
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = 3; c < 6; ++c) m[r][c] = m[r][c] * 0.5;
  double mh[6][6];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double s = m[0][i] * hp[0][j];
#pragma unroll
      for (int k = 1; k < 6; ++k) s = s + m[k][i] * hp[k][j];
      mh[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double s = mh[i][0] * m[0][j];
#pragma unroll
      for (int k = 1; k < 6; ++k) s = s + mh[i][k] * m[k][j];
      curv[i][j] = s;
    }
}

}  // namespace sbk
