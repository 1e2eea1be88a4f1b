// 9 x 9 dense algebra of the local Newton iteration, on a team's LDS slot.
//
// The reference factors K_aa with Eigen's FullPivLU (IncompatibleModes.cpp:335,
// 461), tests the symmetric part with LLT (:491), and in its fallbacks uses a
// 9 x 9 SelfAdjointEigenSolver (:510), partialPivLu (:513) and LLT solves (:527).
// Here the full-pivot LU is team-parallel (lane p owns column/row k+p of each
// step, the pivot is a DPP arg-max with the first-in-column-major tie rule),
// everything else runs redundantly on the team's 8 lanes (identical data and
// identical control flow, so the redundant LDS writes carry identical values).
#pragma once

#include "device_math.hpp"

namespace sbk {

struct LuSlot {
  double a[81];   // row-major; the LU factors after coop_full_lu
  int rt[9], ct[9];
  int nonzero;
  int pad;
  double maxpivot;
};

// Eigen's FullPivLU pivoting: the largest |a_ij| of the remaining corner, the
// first in column-major order on ties; invertible when every pivot exceeds
// 9 eps times the largest pivot.
// The factorisation is a function of its own (not inlined): its unrolled
// code would otherwise change the register allocation of the kernel around
// it. The slot is passed as an LDS pointer so that its accesses stay LDS
// instructions across the call.
using LdsLuSlot = __attribute__((address_space(3))) LuSlot;

// A permutation of 0..8 packed 4 bits per entry (entry i in bits 4i..4i+3).
constexpr unsigned long long kIdentityPerm = 0x876543210ull;
__device__ __forceinline__ int perm_at(unsigned long long P, int i) {
  return static_cast<int>((P >> (4 * i)) & 15ull);
}
__device__ __forceinline__ unsigned long long perm_swap(unsigned long long P, int i, int j) {
  const unsigned long long vi = (P >> (4 * i)) & 15ull, vj = (P >> (4 * j)) & 15ull;
  P &= ~((15ull << (4 * i)) | (15ull << (4 * j)));
  return P | (vi << (4 * j)) | (vj << (4 * i));
}

__device__ __noinline__ bool coop_full_lu_lds(LdsLuSlot* sp, int p) {
  LdsLuSlot& s = *sp;
  // Every loop is unrolled (the steps k included), so that all indices but
  // the pivot's are constants and each phase issues its LDS loads together
  // instead of waiting out one LDS round trip per element. Rows and columns
  // are not swapped in memory: entry (i, j) of the partly factored matrix
  // stays at (R[i], C[j]) of the slot, R and C in registers (the same on the
  // team's lanes), and one gather at the end stores the factors in place;
  // the arithmetic is that of the swapped matrix, operation for operation.
  unsigned long long R = kIdentityPerm, C = kIdentityPerm;
  double maxpivot = 0.0;
  int nonzero = 9;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    // candidates: lane p scans column k + p, lane 0 also column k + 8 (k = 0)
    double best = -1.0;
    int bidx = 1 << 30;
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int j = k + p + 8 * jj;
      if (j < 9 && (jj == 0 || p == 0)) {
        const int cj = perm_at(C, j);
        double v[9];
#pragma unroll
        for (int i = k; i < 9; ++i) v[i] = fabs(s.a[perm_at(R, i) * 9 + cj]);
#pragma unroll
        for (int i = k; i < 9; ++i) {
          const int idx = j * 9 + i;
          if (v[i] > best || (v[i] == best && idx < bidx)) { best = v[i]; bidx = idx; }
        }
      }
    }
    // team arg-max
#pragma unroll
    for (int step = 0; step < 3; ++step) {
      double ov;
      int oi;
      if (step == 0) { ov = dpp_f64<kXor1>(best); oi = dpp_i32<kXor1>(bidx); }
      else if (step == 1) { ov = dpp_f64<kXor2>(best); oi = dpp_i32<kXor2>(bidx); }
      else { ov = dpp_f64<kHalfMirror>(best); oi = dpp_i32<kHalfMirror>(bidx); }
      if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    if (!(best > 0.0)) {
      // a zero corner (or no candidate at all: a NaN matrix)
      nonzero = k;
      if (p == 0)
#pragma unroll
        for (int i = k; i < 9; ++i) { s.rt[i] = i; s.ct[i] = i; }
      break;
    }
    if (best > maxpivot) maxpivot = best;
    const int br = bidx % 9, bc = bidx / 9;
    if (p == 0) { s.rt[k] = br; s.ct[k] = bc; }
    R = perm_swap(R, k, br);
    C = perm_swap(C, k, bc);
    if (k < 8) {
      const int i = k + 1 + p;
      if (i < 9) {
        const int rk = perm_at(R, k), ri = perm_at(R, i);
        double rowk[9], rowi[9];
#pragma unroll
        for (int j = k; j < 9; ++j) {
          const int cj = perm_at(C, j);
          rowk[j] = s.a[rk * 9 + cj];
          rowi[j] = s.a[ri * 9 + cj];
        }
        const double l = rowi[k] / rowk[k];
        s.a[ri * 9 + perm_at(C, k)] = l;
#pragma unroll
        for (int j = k + 1; j < 9; ++j) s.a[ri * 9 + perm_at(C, j)] = rowi[j] - l * rowk[j];
      }
    }
    team_sync();
  }
  // the factors in place: lane p row p (lane 0 also row 8)
  double rows[2][9];
#pragma unroll
  for (int ii = 0; ii < 2; ++ii) {
    const int i = p + 8 * ii;
    if (i < 9 && (ii == 0 || p == 0)) {
      const int ri = perm_at(R, i);
#pragma unroll
      for (int j = 0; j < 9; ++j) rows[ii][j] = s.a[ri * 9 + perm_at(C, j)];
    }
  }
  team_sync();
#pragma unroll
  for (int ii = 0; ii < 2; ++ii) {
    const int i = p + 8 * ii;
    if (i < 9 && (ii == 0 || p == 0))
#pragma unroll
      for (int j = 0; j < 9; ++j) s.a[i * 9 + j] = rows[ii][j];
  }
  if (p == 0) { s.nonzero = nonzero; s.maxpivot = maxpivot; }
  team_sync();
  const double thr = maxpivot * (kEps * 9.0);
  int rank = 0;
  double d[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = s.a[i * 9 + i];
#pragma unroll
  for (int i = 0; i < 9; ++i)
    if (i < nonzero && fabs(d[i]) > thr) ++rank;
  return rank == 9;
}

__device__ __forceinline__ bool coop_full_lu(LuSlot& s, int p) {
  return coop_full_lu_lds((LdsLuSlot*)(&s), p);
}

// Solve A x = b with the factors (every lane, redundantly).
__device__ inline void lu_solve(const LuSlot& s, const double b[9], double x[9]) {
  double c[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) c[i] = b[i];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int r = s.rt[k];
#pragma unroll
    for (int i = k + 1; i < 9; ++i)
      if (i == r) { const double t = c[k]; c[k] = c[i]; c[i] = t; }
  }
#pragma unroll
  for (int i = 1; i < 9; ++i)
#pragma unroll
    for (int k = 0; k < i; ++k) c[i] = c[i] - s.a[i * 9 + k] * c[k];
#pragma unroll
  for (int i = 8; i >= 0; --i) {
#pragma unroll
    for (int k = i + 1; k < 9; ++k) c[i] = c[i] - s.a[i * 9 + k] * c[k];
    c[i] = c[i] / s.a[i * 9 + i];
  }
#pragma unroll
  for (int k = 8; k >= 0; --k) {
    const int q = s.ct[k];
#pragma unroll
    for (int i = k + 1; i < 9; ++i)
      if (i == q) { const double t = c[k]; c[k] = c[i]; c[i] = t; }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) x[i] = c[i];
}

// The composed permutations of full-pivot factors, for solves whose right-hand
// sides stay in memory: row i of P b is row prow[i] of b, and entry j of the
// substituted solution is entry qcol[j] of x. Indexing memory by them replaces
// the transposition swaps of lu_solve, which on registers become select chains.
struct LuPerm {
  int prow[9];
  int qcol[9];
};

// One lane forms the permutations from the row and column transpositions
// (rt, ct: 9 each, as LuSlot holds them).
__device__ inline void lu_perm(const int rt[9], const int ct[9], LuPerm& P) {
  for (int i = 0; i < 9; ++i) P.prow[i] = i;
  for (int k = 0; k < 9; ++k) {
    const int r = rt[k];
    const int t = P.prow[k];
    P.prow[k] = P.prow[r];
    P.prow[r] = t;
  }
  // the column transpositions in reverse order, as lu_solve applies them,
  // composed on an index list kept in P.qcol's storage first
  for (int i = 0; i < 9; ++i) P.qcol[i] = i;
  for (int k = 8; k >= 0; --k) {
    const int q = ct[k];
    const int t = P.qcol[k];
    P.qcol[k] = P.qcol[q];
    P.qcol[q] = t;
  }
  // x[i] = c[idx[i]] -> c[j] goes to x[inverse(idx)[j]]
  int jdx[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) jdx[i] = P.qcol[i];
  for (int i = 0; i < 9; ++i) P.qcol[jdx[i]] = i;
}

// x = A^-1 b in place, for b in column `col` of the row-major matrix m with
// `ld` columns (LDS or global), with the factors f (LuSlot::a's layout):
// lu_solve's arithmetic, in the same order.
__device__ inline void lu_solve_column(const double* f, const LuPerm& P, double* m, int ld, int col) {
  double c[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) c[i] = m[P.prow[i] * ld + col];
#pragma unroll
  for (int i = 1; i < 9; ++i)
#pragma unroll
    for (int k = 0; k < i; ++k) c[i] = c[i] - f[i * 9 + k] * c[k];
#pragma unroll
  for (int i = 8; i >= 0; --i) {
#pragma unroll
    for (int k = i + 1; k < 9; ++k) c[i] = c[i] - f[i * 9 + k] * c[k];
    c[i] = c[i] / f[i * 9 + i];
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) m[P.qcol[j] * ld + col] = c[j];
}

// lu_solve_column on the three adjacent columns col, col + 1, col + 2: each
// column's operations the same and in the same order, but every
// column loaded before any is stored, so that the three substitution chains
// overlap (stores through the permutations would otherwise order each
// column's loads after the previous column's stores) and each factor entry
// is read once for the three.
__device__ inline void lu_solve_column3(const double* f, const LuPerm& P, double* m, int ld, int col) {
  double c[3][9];
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k][i] = m[P.prow[i] * ld + col + k];
#pragma unroll
  for (int i = 1; i < 9; ++i)
#pragma unroll
    for (int kk = 0; kk < i; ++kk) {
      const double fik = f[i * 9 + kk];
#pragma unroll
      for (int k = 0; k < 3; ++k) c[k][i] = c[k][i] - fik * c[k][kk];
    }
#pragma unroll
  for (int i = 8; i >= 0; --i) {
#pragma unroll
    for (int kk = i + 1; kk < 9; ++kk) {
      const double fik = f[i * 9 + kk];
#pragma unroll
      for (int k = 0; k < 3; ++k) c[k][i] = c[k][i] - fik * c[k][kk];
    }
    const double d = f[i * 9 + i];
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k][i] = c[k][i] / d;
  }
#pragma unroll
  for (int j = 0; j < 9; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) m[P.qcol[j] * ld + col + k] = c[k][j];
}

// Solve A^T x = b with the factors of A (P A Q = L U, the row
// transpositions rt applied in order, the column ones ct): x = P^T L^-T U^-T
// Q^T b (every lane, redundantly).
__device__ inline void lu_solve_transposed(const LuSlot& s, const double b[9], double x[9]) {
  double c[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) c[i] = b[i];
  // Q^T b: the column transpositions in order
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int q = s.ct[k];
#pragma unroll
    for (int i = k + 1; i < 9; ++i)
      if (i == q) { const double t = c[k]; c[k] = c[i]; c[i] = t; }
  }
  // U^T y = c (forward)
#pragma unroll
  for (int i = 0; i < 9; ++i) {
#pragma unroll
    for (int k = 0; k < i; ++k) c[i] = c[i] - s.a[k * 9 + i] * c[k];
    c[i] = c[i] / s.a[i * 9 + i];
  }
  // L^T z = y (backward, unit diagonal)
#pragma unroll
  for (int i = 8; i >= 0; --i)
#pragma unroll
    for (int k = i + 1; k < 9; ++k) c[i] = c[i] - s.a[k * 9 + i] * c[k];
  // P^T z: the row transpositions in reverse order
#pragma unroll
  for (int k = 8; k >= 0; --k) {
    const int r = s.rt[k];
#pragma unroll
    for (int i = k + 1; i < 9; ++i)
      if (i == r) { const double t = c[k]; c[k] = c[i]; c[i] = t; }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) x[i] = c[i];
}

// Eigen's unblocked LLT on the lower triangle of w (in place, redundant);
// false at the first non-positive pivot (the rest of w left as it was).
__device__ inline bool llt9(double* w) {
  for (int k = 0; k < 9; ++k) {
    double x = w[k * 9 + k];
    if (k > 0) {
      double s = w[k * 9] * w[k * 9];
      for (int j = 1; j < k; ++j) s = s + w[k * 9 + j] * w[k * 9 + j];
      x = x - s;
    }
    if (x <= 0.0) return false;
    x = sqrt(x);
    w[k * 9 + k] = x;
    if (k > 0)
      for (int i = k + 1; i < 9; ++i) {
        double s = w[i * 9] * w[k * 9];
        for (int j = 1; j < k; ++j) s = s + w[i * 9 + j] * w[k * 9 + j];
        w[i * 9 + k] = w[i * 9 + k] - s;
      }
    const double ix = 1.0 / x;  // one division per column; each quotient within an ulp
    for (int i = k + 1; i < 9; ++i) w[i * 9 + k] = w[i * 9 + k] * ix;
  }
  return true;
}

// Whether sym(K) = (K + K^T) / 2 of the row-major K in LDS is positive
// definite: llt9's operations on it, in the same order, in registers (every
// lane; only the verdict is kept). Not inlined, for the reason coop_full_lu
// is not.
using LdsDouble = __attribute__((address_space(3))) double;
__device__ __noinline__ bool spd9_lds(const LdsDouble* k) {
  double w[9][9];
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) w[i][j] = 0.5 * (k[i * 9 + j] + k[j * 9 + i]);
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    double x = w[c][c];
    if (c > 0) {
      double s = w[c][0] * w[c][0];
#pragma unroll
      for (int j = 1; j < c; ++j) s = s + w[c][j] * w[c][j];
      x = x - s;
    }
    if (x <= 0.0) return false;
    x = sqrt(x);
    w[c][c] = x;
    if (c > 0)
#pragma unroll
      for (int i = c + 1; i < 9; ++i) {
        double s = w[i][0] * w[c][0];
#pragma unroll
        for (int j = 1; j < c; ++j) s = s + w[i][j] * w[c][j];
        w[i][c] = w[i][c] - s;
      }
#pragma unroll
    for (int i = c + 1; i < 9; ++i) w[i][c] = w[i][c] / x;
  }
  return true;
}
__device__ __forceinline__ bool spd9(const double* k) { return spd9_lds((const LdsDouble*)k); }

__device__ inline void llt9_solve(const double* w, const double b[9], double x[9]) {
  double y[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) y[i] = b[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
#pragma unroll
    for (int k = 0; k < i; ++k) y[i] = y[i] - w[i * 9 + k] * y[k];
    y[i] = y[i] / w[i * 9 + i];
  }
#pragma unroll
  for (int i = 8; i >= 0; --i) {
#pragma unroll
    for (int k = i + 1; k < 9; ++k) y[i] = y[i] - w[k * 9 + i] * y[k];
    y[i] = y[i] / w[i * 9 + i];
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) x[i] = y[i];
}

// Eigenvalues (ascending) of the symmetric matrix in w, by cyclic Jacobi in
// place (redundant; the fallback path only).
__device__ inline void jacobi9_values(double* a, double ev[9]) {
  for (int sweep = 0; sweep < 64; ++sweep) {
    bool any = false;
    for (int p = 0; p < 8; ++p)
      for (int q = p + 1; q < 9; ++q) {
        const double apq = a[p * 9 + q];
        if (apq == 0.0) continue;
        const double g = 100.0 * fabs(apq);
        const double app = fabs(a[p * 9 + p]), aqq = fabs(a[q * 9 + q]);
        if (app + g == app && aqq + g == aqq) {
          a[p * 9 + q] = 0.0;
          a[q * 9 + p] = 0.0;
          continue;
        }
        any = true;
        const double h = a[q * 9 + q] - a[p * 9 + p];
        double t;
        if (fabs(h) + g == fabs(h)) {
          t = apq / h;
        } else {
          const double theta = 0.5 * h / apq;
          t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
          if (theta < 0.0) t = -t;
        }
        const double c = 1.0 / sqrt(1.0 + t * t);
        const double s = t * c;
        const double tau = s / (1.0 + c);
        a[p * 9 + p] = a[p * 9 + p] - t * apq;
        a[q * 9 + q] = a[q * 9 + q] + t * apq;
        a[p * 9 + q] = 0.0;
        a[q * 9 + p] = 0.0;
        for (int r = 0; r < 9; ++r) {
          if (r == p || r == q) continue;
          const double gr = a[r * 9 + p], hr = a[r * 9 + q];
          a[r * 9 + p] = gr - s * (hr + gr * tau);
          a[p * 9 + r] = a[r * 9 + p];
          a[r * 9 + q] = hr + s * (gr - hr * tau);
          a[q * 9 + r] = a[r * 9 + q];
        }
      }
    if (!any) break;
  }
  for (int i = 0; i < 9; ++i) ev[i] = a[i * 9 + i];
  for (int i = 1; i < 9; ++i)
    for (int j = i; j > 0 && ev[j] < ev[j - 1]; --j) {
      const double t = ev[j];
      ev[j] = ev[j - 1];
      ev[j - 1] = t;
    }
}

// Partial-pivot LU solve of the 9 x 9 system in w (in place, redundant).
__device__ inline void plu9_solve(double* a, const double b[9], double x[9]) {
  int piv[9];
  for (int k = 0; k < 9; ++k) {
    int br = k;
    double big = fabs(a[k * 9 + k]);
    for (int i = k + 1; i < 9; ++i) {
      const double v = fabs(a[i * 9 + k]);
      if (v > big) { big = v; br = i; }
    }
    piv[k] = br;
    if (big != 0.0) {
      if (br != k)
        for (int j = 0; j < 9; ++j) {
          const double t = a[k * 9 + j];
          a[k * 9 + j] = a[br * 9 + j];
          a[br * 9 + j] = t;
        }
      for (int i = k + 1; i < 9; ++i) a[i * 9 + k] = a[i * 9 + k] / a[k * 9 + k];
    }
    if (k < 8)
      for (int i = k + 1; i < 9; ++i)
        for (int j = k + 1; j < 9; ++j) a[i * 9 + j] = a[i * 9 + j] - a[i * 9 + k] * a[k * 9 + j];
  }
  double c[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) c[i] = b[i];
  for (int k = 0; k < 9; ++k) {
    const int r = piv[k];
#pragma unroll
    for (int i = 0; i < 9; ++i)
      if (i == r && i > k) { const double t = c[k]; c[k] = c[i]; c[i] = t; }
  }
#pragma unroll
  for (int i = 1; i < 9; ++i)
#pragma unroll
    for (int k = 0; k < i; ++k) c[i] = c[i] - a[i * 9 + k] * c[k];
#pragma unroll
  for (int i = 8; i >= 0; --i) {
#pragma unroll
    for (int k = i + 1; k < 9; ++k) c[i] = c[i] - a[i * 9 + k] * c[k];
    c[i] = c[i] / a[i * 9 + i];
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) x[i] = c[i];
}

}  // namespace sbk
