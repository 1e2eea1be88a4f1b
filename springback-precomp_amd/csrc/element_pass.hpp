// One integration point per lane: the element set-up and one pass of the
// incompatible-mode Hex8 at the mode parameters alpha.
//
// Reference: src/fem/IncompatibleModes.cpp:134-192 (set_up), :195-222
// (kinematics_at), :227-247 (respond), :252-275 (accumulate_local);
// src/elements/Hex8Incompatible.cpp:73-83 (mode gradients).
#pragma once

#include "device_return.hpp"

namespace sbk {

enum Kin { kSmall = 0, kFinite = 1, kLog = 2 };

// Failure codes written per element (0 = success).
enum ElementStatus {
  kOk = 0,
  kCentreJacobian = 1,     // MeshError: det J <= 0 at the centre
  kPointJacobian = 2,      // MeshError: det J <= 0 at a point
  kInverted = 3,           // SolverError: det F <= 0 (logarithmic kinematics)
  kDegenerate = 4,         // SolverError: non-finite E or 1 + 2 mu <= 0
  kReturnFailed = 5,       // SolverError: the return did not converge
  kSingularModes = 6,      // SolverError: K_aa singular
  kLocalNotConverged = 7,  // SolverError: 25 local iterations
  kNoDescent = 8,          // SolverError: no step lowers |r_alpha|
  kNonFiniteTrial = 9,     // SolverError: non-finite trial stress
  kUnsupported = 10,       // the launch does not fit the kernel
};

// The alpha-independent data of the lane's point.
struct PointGeom {
  double g[3][11];  // [G | G~]
  double hu[3][3];  // U G^T
  double w;         // w_q det J
};

__device__ __forceinline__ void load_element(const double* __restrict__ coords, const int* __restrict__ conn,
                                             const double* __restrict__ u, long long e, double x0[8][3],
                                             double ue[24]) {
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const long long n = conn[e * 8 + a];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      x0[a][k] = coords[n * 3 + k];
      ue[3 * a + k] = u[n * 3 + k];
    }
  }
}

// Returns 0 or kCentreJacobian / kPointJacobian (fail_value: the determinant).
__device__ __forceinline__ int setup_point(const double x0[8][3], const double ue[24], double xi, double eta,
                                           double zeta, double wq, PointGeom& pg, double& fail_value) {
  double dn0[8][3];
  shape_gradients(0.0, 0.0, 0.0, dn0);
  double j0[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = x0[0][i] * dn0[0][j];
#pragma unroll
      for (int a = 1; a < 8; ++a) s = s + x0[a][i] * dn0[a][j];
      j0[i][j] = s;
    }
  const double det0 = det3(j0);
  if (!(det0 > 0.0)) { fail_value = det0; return kCentreJacobian; }
  double ji0[3][3];
  inverse3(j0, ji0);
  double dn[8][3];
  shape_gradients(xi, eta, zeta, dn);
  double jq[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = x0[0][i] * dn[0][j];
#pragma unroll
      for (int a = 1; a < 8; ++a) s = s + x0[a][i] * dn[a][j];
      jq[i][j] = s;
    }
  const double det = det3(jq);
  if (!(det > 0.0)) { fail_value = det; return kPointJacobian; }
  double ji[3][3];
  inverse3(jq, ji);
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) pg.g[i][a] = dn[a][0] * ji[0][i] + dn[a][1] * ji[1][i] + dn[a][2] * ji[2][i];
  const double ratio = det0 / det;
  const double nat[3] = {xi, eta, zeta};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double f = ratio * -2.0 * nat[k];
#pragma unroll
    for (int i = 0; i < 3; ++i) pg.g[i][8 + k] = f * ji0[k][i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = ue[i] * pg.g[j][0];
#pragma unroll
      for (int a = 1; a < 8; ++a) s = s + ue[3 * a + i] * pg.g[j][a];
      pg.hu[i][j] = s;
    }
  pg.w = wq * det;
  return kOk;
}

template <int NT>
__device__ __forceinline__ void load_state(const DevMaterial& m, const double* __restrict__ state, long long stride,
                                           long long q, PointStateT<NT>& st) {
#pragma unroll
  for (int i = 0; i < 6; ++i) st.ep[i] = state[i * stride + q];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
#pragma unroll
    for (int i = 0; i < 6; ++i) st.kin[t][i] = (t < m.kin_fields) ? state[(6 + 6 * t + i) * stride + q] : 0.0;
  }
  st.abar = state[(6 + 6 * m.kin_fields) * stride + q];
  st.loading = state[(7 + 6 * m.kin_fields) * stride + q] != 0.0;
}

template <int NT>
__device__ __forceinline__ void store_state(const DevMaterial& m, double* __restrict__ state, long long stride,
                                            long long q, const PointStateT<NT>& st) {
#pragma unroll
  for (int i = 0; i < 6; ++i) state[i * stride + q] = st.ep[i];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t < m.kin_fields) {
#pragma unroll
      for (int i = 0; i < 6; ++i) state[(6 + 6 * t + i) * stride + q] = st.kin[t][i];
    }
  }
  state[(6 + 6 * m.kin_fields) * stride + q] = st.abar;
  state[(7 + 6 * m.kin_fields) * stride + q] = st.loading ? 1.0 : 0.0;
}

// B_NL(F, [G | G~]) entry (c, 3a + k) (IncompatibleModes.cpp:110-121).
__device__ __forceinline__ double bnl(const double F[3][3], const double g[3][11], int c, int a, int k) {
  switch (c) {
    case 0: return F[k][0] * g[0][a];
    case 1: return F[k][1] * g[1][a];
    case 2: return F[k][2] * g[2][a];
    case 3: return F[k][0] * g[1][a] + F[k][1] * g[0][a];
    case 4: return F[k][1] * g[2][a] + F[k][2] * g[1][a];
    default: return F[k][2] * g[0][a] + F[k][0] * g[2][a];
  }
}
// The same entry for mode b (column 3 (8 + b) + k), from the mode gradients
// gm[i][b] = G~[i][b] alone.
__device__ __forceinline__ double bnl_mode(const double F[3][3], const double gm[3][3], int c, int b, int k) {
  switch (c) {
    case 0: return F[k][0] * gm[0][b];
    case 1: return F[k][1] * gm[1][b];
    case 2: return F[k][2] * gm[2][b];
    case 3: return F[k][0] * gm[1][b] + F[k][1] * gm[0][b];
    case 4: return F[k][1] * gm[2][b] + F[k][2] * gm[1][b];
    default: return F[k][2] * gm[0][b] + F[k][0] * gm[2][b];
  }
}

// What a pass at alpha needs of the lane's point: the mode gradients, the
// nodal displacement gradient and the weight (and for small strain the nodal
// part of the strain, B_u ue); 19 (25) doubles instead of the 43 of PointGeom.
struct ModeGeom {
  double gm[3][3];  // gm[i][b] = G~[i][b]
  double hu[3][3];  // U G^T
  double w;
  double eps_u[6];  // small strain only
};

template <int KIN>
__device__ __forceinline__ void mode_geom(const PointGeom& pg, const double ue[24], ModeGeom& mg) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) { mg.gm[i][j] = pg.g[i][8 + j]; mg.hu[i][j] = pg.hu[i][j]; }
  mg.w = pg.w;
  if constexpr (KIN == kSmall) {
    double I3[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double s1 = bnl(I3, pg.g, c, 0, 0) * ue[0];
#pragma unroll
      for (int kk = 1; kk < 24; ++kk) s1 = s1 + bnl(I3, pg.g, c, kk / 3, kk % 3) * ue[kk];
      mg.eps_u[c] = s1;
    }
  }
}

// A lane's slice of a per-workgroup LDS array laid out [field][64 lanes], so
// that the lanes of a wavefront touch consecutive doubles (no bank conflicts).
// The kernels keep long-lived per-point data here instead of in registers.
struct LaneLds {
  double* p;  // &array[0][lane]
  __device__ __forceinline__ double get(int k) const { return p[k * 64]; }
  __device__ __forceinline__ void set(int k, double v) const { p[k * 64] = v; }
};

// Forces the values kept in LDS to be read again after this point instead of
// being held in registers across it (an empty asm with a memory clobber; no
// instruction is emitted).
__device__ __forceinline__ void reload_barrier() { __asm__ volatile("" ::: "memory"); }

// ModeGeom in LDS: gm (9), hu (9), w, and for small strain eps_u (6).
template <int KIN>
constexpr int mode_geom_fields() { return KIN == kSmall ? 25 : 19; }

template <int KIN>
__device__ __forceinline__ void put_mode_geom(const ModeGeom& mg, LaneLds s) {
#pragma unroll
  for (int i = 0; i < 9; ++i) { s.set(i, mg.gm[i / 3][i % 3]); s.set(9 + i, mg.hu[i / 3][i % 3]); }
  s.set(18, mg.w);
  if constexpr (KIN == kSmall) {
#pragma unroll
    for (int c = 0; c < 6; ++c) s.set(19 + c, mg.eps_u[c]);
  }
}
__device__ __forceinline__ void get_gm(LaneLds s, double gm[3][3]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) gm[i / 3][i % 3] = s.get(i);
}
__device__ __forceinline__ void get_hu(LaneLds s, double hu[3][3]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) hu[i / 3][i % 3] = s.get(9 + i);
}

// H = hu + A G~^T (IncompatibleModes.cpp:195-205); F = I + H except for small
// strain (F = I). Every kernel forms F with this function (the compiler may
// fuse its products differently in each, so copies agree to rounding).
template <int KIN>
__device__ __forceinline__ void mode_kinematics(const double hu[3][3], const double gm[3][3], const double alpha[9],
                                                double h[3][3], double F[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      h[i][j] = hu[i][j] + (alpha[i] * gm[j][0] + alpha[3 + i] * gm[j][1] + alpha[6 + i] * gm[j][2]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if constexpr (KIN == kSmall) F[i][j] = (i == j) ? 1.0 : 0.0;
      else F[i][j] = (i == j) ? 1.0 + h[i][j] : h[i][j];
    }
}

// The response of the lane's point at alpha: the stress conjugate to B_NL,
// the tangent between B_NL's, energy density and yielding; the point's new
// state is written to state_out. Returns 0 or a failure code. The point's
// ModeGeom is read from LDS (geom); for logarithmic strain the principal axes
// and divided differences wait in LDS (logs) while the return runs.
struct PointResponse {
  double S[6];
  double mat[6][6];
  double energy;
  bool yielding;
};

constexpr int kLogStash = 18;  // q (9), mu (3), first divided differences (6)

__device__ __forceinline__ void put_log(const LogData& L, LaneLds s) {
#pragma unroll
  for (int i = 0; i < 9; ++i) s.set(i, L.q[i / 3][i % 3]);
#pragma unroll
  for (int a = 0; a < 3; ++a) s.set(9 + a, L.mu[a]);
  s.set(12, L.first[0][0]); s.set(13, L.first[1][1]); s.set(14, L.first[2][2]);
  s.set(15, L.first[0][1]); s.set(16, L.first[1][2]); s.set(17, L.first[0][2]);
}
__device__ __forceinline__ void get_log(LaneLds s, LogData& L) {
#pragma unroll
  for (int i = 0; i < 9; ++i) L.q[i / 3][i % 3] = s.get(i);
#pragma unroll
  for (int a = 0; a < 3; ++a) L.mu[a] = s.get(9 + a);
  L.first[0][0] = s.get(12); L.first[1][1] = s.get(13); L.first[2][2] = s.get(14);
  L.first[0][1] = L.first[1][0] = s.get(15);
  L.first[1][2] = L.first[2][1] = s.get(16);
  L.first[0][2] = L.first[2][0] = s.get(17);
}

template <int KIN, int NT>
__device__ __forceinline__ int point_response(const DevMaterial& m, LaneLds geom, const double alpha[9],
                                              PointStateT<NT>& st, PointResponse& r, double* __restrict__ state_out,
                                              long long stride, long long q, LaneLds logs) {
  double strain[6];
  {
    double hu[3][3], gm[3][3], h[3][3], F[3][3];
    get_hu(geom, hu);
    get_gm(geom, gm);
    mode_kinematics<KIN>(hu, gm, alpha, h, F);
    if constexpr (KIN == kSmall) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double s2 = bnl_mode(F, gm, c, 0, 0) * alpha[0];
#pragma unroll
        for (int kk = 1; kk < 9; ++kk) s2 = s2 + bnl_mode(F, gm, c, kk / 3, kk % 3) * alpha[kk];
        strain[c] = geom.get(19 + c) + s2;
      }
    } else {
      if constexpr (KIN == kLog) {
        if (!(det3(F) > 0.0)) return kInverted;
      }
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
        LogData L;
        const int lf = logarithmic_strain(green, L);
        if (lf != 0) return kDegenerate;
#pragma unroll
        for (int i = 0; i < 6; ++i) strain[i] = L.strain[i];
        put_log(L, logs);
      } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) strain[i] = green[i];
      }
    }
  }
  reload_barrier();
  ReturnOut ro;
  const int rf = m.general ? return_general(m, strain, st, ro) : return_radial(m, strain, st, ro);
  if (rf == kReturnNotConverged) return kReturnFailed;
  if (rf == kReturnNonFinite) return kNonFiniteTrial;
  r.energy = ro.energy;
  r.yielding = ro.yielding;
  // nothing after the return can fail: the trial state is final here
  store_state(m, state_out, stride, q, st);
  if constexpr (KIN == kLog) {
    LogData L;
    get_log(logs, L);
    log_material(L, ro.stress, ro.C, r.S, r.mat);
  } else {
#pragma unroll
    for (int i = 0; i < 6; ++i) r.S[i] = ro.stress[i];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) r.mat[i][j] = ro.C[i][j];
  }
  return kOk;
}

}  // namespace sbk
