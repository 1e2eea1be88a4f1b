// The records of the element workspace and the row-block products read
// from them, shared by the condensation and the adjoint.
//
// Reference: src/fem/IncompatibleModes.cpp:279-306 (accumulate_coupling).
#pragma once

#include "element_kernels.hpp"
#include "element_pass.hpp"

namespace sbk {

// A point record of the workspace: S (6) then the tangent (36, row-major).
constexpr int kWsStress = 0;
constexpr int kWsTangent = 6;
// An element record: LU factors (81), r_a (9), row and column transpositions.
constexpr int kErFactors = 0;
constexpr int kErResidual = 81;
constexpr int kErRows = 90;
constexpr int kErCols = 99;
// A geometry record: F (9), [G | G~] (3 x 11, row-major), w.
constexpr int kGrF = 0;
constexpr int kGrG = 9;
constexpr int kGrW = 42;

// The helpers below read records through R: a pointer to the record, or
// any reader with operator[] over the record's entries (the team
// condensation reads them from its LDS stage, StagedRecord).

// B_NL(F, [G | G~]) entry (c, 3a + k) with F and G read from a geometry record.
template <class R>
__device__ __forceinline__ double bnl_rec(const R& gr, int c, int a, int k) {
  switch (c) {
    case 0: return gr[kGrF + k * 3 + 0] * gr[kGrG + 0 * 11 + a];
    case 1: return gr[kGrF + k * 3 + 1] * gr[kGrG + 1 * 11 + a];
    case 2: return gr[kGrF + k * 3 + 2] * gr[kGrG + 2 * 11 + a];
    case 3: return gr[kGrF + k * 3 + 0] * gr[kGrG + 1 * 11 + a] + gr[kGrF + k * 3 + 1] * gr[kGrG + 0 * 11 + a];
    case 4: return gr[kGrF + k * 3 + 1] * gr[kGrG + 2 * 11 + a] + gr[kGrF + k * 3 + 2] * gr[kGrG + 1 * 11 + a];
    default: return gr[kGrF + k * 3 + 2] * gr[kGrG + 0 * 11 + a] + gr[kGrF + k * 3 + 0] * gr[kGrG + 2 * 11 + a];
  }
}

// z = B_u[:, 3a..3a+2]^T mat of one point (mat row-major in the point record).
template <class R>
__device__ __forceinline__ void node_times_tangent(const double bu[6][3], const R& pr, double z[3][6]) {
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    double mc[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) mc[c] = pr[kWsTangent + c * 6 + d];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double s = bu[0][k] * mc[0];
#pragma unroll
      for (int c = 1; c < 6; ++c) s = s + bu[c][k] * mc[c];
      z[k][d] = s;
    }
  }
}

template <class R>
__device__ __forceinline__ void node_operator(const R& gr, int a, double bu[6][3]) {
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int k = 0; k < 3; ++k) bu[c][k] = bnl_rec(gr, c, a, k);
}

// (S G)_a: row a of G^T S (3), S tensorial Voigt.
template <class R, class P>
__device__ __forceinline__ void stress_times_gradient(const R& gr, const P& pr, int a, double gsa[3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double S3j0 = j == 0 ? pr[kWsStress + 0] : (j == 1 ? pr[kWsStress + 3] : pr[kWsStress + 5]);
    const double S3j1 = j == 0 ? pr[kWsStress + 3] : (j == 1 ? pr[kWsStress + 1] : pr[kWsStress + 4]);
    const double S3j2 = j == 0 ? pr[kWsStress + 5] : (j == 1 ? pr[kWsStress + 4] : pr[kWsStress + 2]);
    gsa[j] = gr[kGrG + 0 * 11 + a] * S3j0 + gr[kGrG + 1 * 11 + a] * S3j1 + gr[kGrG + 2 * 11 + a] * S3j2;
  }
}

// One point's share of lane a's rows of f_u and K_ua (3 x 9), from its
// geometry record gr and point record pr.
template <int KIN, class R, class P>
__device__ __forceinline__ void coupling_point(const R& gr, const P& pr, int a, double ru[3], double kua[3][9]) {
  const double w = gr[kGrW];
  double bu[6][3];
  node_operator(gr, a, bu);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double s = bu[0][k] * pr[kWsStress];
#pragma unroll
    for (int c = 1; c < 6; ++c) s = s + bu[c][k] * pr[kWsStress + c];
    ru[k] = ru[k] + w * s;
  }
  double z[3][6];
  node_times_tangent(bu, pr, z);
#pragma unroll
  for (int nn = 0; nn < 9; ++nn) {
    double bcol[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) bcol[c] = bnl_rec(gr, c, 8 + nn / 3, nn % 3);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double s = z[k][0] * bcol[0];
#pragma unroll
      for (int d = 1; d < 6; ++d) s = s + z[k][d] * bcol[d];
      kua[k][nn] = kua[k][nn] + w * s;
    }
  }
  if constexpr (KIN != kSmall) {
    // geometric (G_aug^T S G_aug) (x) I: block (a, 8 + bm)
    double gsa[3];
    stress_times_gradient(gr, pr, a, gsa);
#pragma unroll
    for (int bm = 0; bm < 3; ++bm) {
      const double geo = gsa[0] * gr[kGrG + 0 * 11 + 8 + bm] + gsa[1] * gr[kGrG + 1 * 11 + 8 + bm] +
                         gsa[2] * gr[kGrG + 2 * 11 + 8 + bm];
      const double v = w * geo;
#pragma unroll
      for (int k = 0; k < 3; ++k) kua[k][3 * bm + k] = kua[k][3 * bm + k] + v;
    }
  }
}

// Lane a's rows of f_u and K_ua (3 x 9), accumulated point by point in point
// order from the geometry and point records of element points q0..q0+n-1.
template <int KIN>
__device__ __forceinline__ void coupling_rows(const ElementArgs& A, long long q0, int n, int a, double ru[3],
                                              double kua[3][9]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ru[k] = 0.0;
#pragma unroll
    for (int j = 0; j < 9; ++j) kua[k][j] = 0.0;
  }
  for (int q = 0; q < n; ++q) {
    const double* gr = A.geom_ws + (q0 + q) * kGeomRecord;
    const double* pr = A.point_ws + (q0 + q) * kPointRecord;
    coupling_point<KIN>(gr, pr, a, ru, kua);
  }
}

}  // namespace sbk
