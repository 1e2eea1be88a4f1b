// Oracle (test infrastructure only): scalar and small dense-matrix helpers.
//
// This directory is the CPU restatement of the reference's element-and-assembly
// path. Only tests, the smoke check and the CPU-baseline leg of the benchmark
// may load it, and only as the checker. The product library never links it.
//
// Every algorithm is templated on the scalar type T. `double` is the oracle
// proper; `Counted` charges every floating-point operation to a thread-local
// counter (add/sub/mul/div/sqrt and each libm call = 1 flop, comparisons and
// sign flips = 0), which gives the algorithmic FLOP count of the reference
// algorithm exactly as it is written.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>

namespace oracle {

// ---------------------------------------------------------------------------
// Errors: the reference's exception hierarchy (include/sparlab/core/Exceptions.hpp)
// ---------------------------------------------------------------------------
struct SolverError : std::runtime_error {
  explicit SolverError(const std::string& m) : std::runtime_error(m) {}
};
struct MeshError : std::runtime_error {
  explicit MeshError(const std::string& m) : std::runtime_error(m) {}
};
struct ConfigError : std::runtime_error {
  explicit ConfigError(const std::string& m) : std::runtime_error(m) {}
};
struct ModelError : std::runtime_error {
  explicit ModelError(const std::string& m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------------------
// Planted mistakes (tests only): a test sets one through oracle_set_plant to
// show that its known answer fails when the oracle is wrong in that way, and
// sets 0 again. 0, the default, is the reference's algorithm.
// ---------------------------------------------------------------------------
enum Plant : int {
  kPlantNone = 0,
  kPlantVoceEnergySign = 1,   // isotropic_energy: the Voce term subtracted
  kPlantKinematicEnergy = 2,  // kinematic_energy: 1/2 instead of 3/4
  kPlantBackwardEulerChi = 3, // backward Euler: chi = dl (and dchi = 1), theta kept
  kPlantSlipRadius = 4,       // Coulomb slip limit mu kappa A instead of mu p_N
  kPlantNormalForce = 5,      // contact normal force (1 + 1e-6) kappa A |g|
  kPlantVoceYield = 6,        // Voce saturation rate 1 % high in the yield stress
};
inline int g_plant = kPlantNone;

// ---------------------------------------------------------------------------
// The op-counting scalar
// ---------------------------------------------------------------------------
inline std::uint64_t& flop_counter() {
  static thread_local std::uint64_t count = 0;
  return count;
}
// The counter's value when the local iteration of an element has converged
// (the split between the local solve and the condensation).
inline std::uint64_t& flop_mark() {
  static thread_local std::uint64_t mark = 0;
  return mark;
}

struct Counted {
  double v = 0.0;
  Counted() = default;
  Counted(double x) : v(x) {}  // NOLINT: implicit on purpose (constants)
  explicit operator double() const { return v; }
  Counted operator-() const { return Counted(-v); }
  Counted& operator+=(const Counted& o) { ++flop_counter(); v += o.v; return *this; }
  Counted& operator-=(const Counted& o) { ++flop_counter(); v -= o.v; return *this; }
  Counted& operator*=(const Counted& o) { ++flop_counter(); v *= o.v; return *this; }
  Counted& operator/=(const Counted& o) { ++flop_counter(); v /= o.v; return *this; }
};
inline Counted operator+(Counted a, Counted b) { ++flop_counter(); return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { ++flop_counter(); return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { ++flop_counter(); return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { ++flop_counter(); return Counted(a.v / b.v); }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }

// Scalar functions, overloaded for both scalar types.
inline double val(double x) { return x; }
inline double val(Counted x) { return x.v; }
inline double sqrt_(double x) { return std::sqrt(x); }
inline Counted sqrt_(Counted x) { ++flop_counter(); return Counted(std::sqrt(x.v)); }
inline double exp_(double x) { return std::exp(x); }
inline Counted exp_(Counted x) { ++flop_counter(); return Counted(std::exp(x.v)); }
inline double expm1_(double x) { return std::expm1(x); }
inline Counted expm1_(Counted x) { ++flop_counter(); return Counted(std::expm1(x.v)); }
inline double log1p_(double x) { return std::log1p(x); }
inline Counted log1p_(Counted x) { ++flop_counter(); return Counted(std::log1p(x.v)); }
inline double abs_(double x) { return std::fabs(x); }
inline Counted abs_(Counted x) { return Counted(std::fabs(x.v)); }
inline bool finite_(double x) { return std::isfinite(x); }
inline bool finite_(Counted x) { return std::isfinite(x.v); }
template <class T> inline T max_(T a, T b) { return a < b ? b : a; }   // std::max semantics
template <class T> inline T min_(T a, T b) { return b < a ? b : a; }   // std::min semantics
template <class T> inline T clamp_(T v, T lo, T hi) { return v < lo ? lo : (hi < v ? hi : v); }

constexpr double kEps = std::numeric_limits<double>::epsilon();

// ---------------------------------------------------------------------------
// Small dense matrices, row-major storage, operator()(row, col)
// ---------------------------------------------------------------------------
template <class T, int R, int C>
struct Mat {
  T a[R * C];
  T& operator()(int i, int j) { return a[i * C + j]; }
  const T& operator()(int i, int j) const { return a[i * C + j]; }
  static Mat zero() {
    Mat m;
    for (int i = 0; i < R * C; ++i) m.a[i] = T(0.0);
    return m;
  }
  static Mat identity() {
    Mat m = zero();
    for (int i = 0; i < (R < C ? R : C); ++i) m(i, i) = T(1.0);
    return m;
  }
};

template <class T, int N>
struct Vec {
  T a[N];
  T& operator()(int i) { return a[i]; }
  const T& operator()(int i) const { return a[i]; }
  static Vec zero() {
    Vec v;
    for (int i = 0; i < N; ++i) v.a[i] = T(0.0);
    return v;
  }
};

template <class T, int N> T dot(const Vec<T, N>& x, const Vec<T, N>& y) {
  T s = x(0) * y(0);
  for (int i = 1; i < N; ++i) s = s + x(i) * y(i);
  return s;
}
template <class T, int N> T norm2(const Vec<T, N>& x) { return sqrt_(dot(x, x)); }
template <class T, int N> T max_abs(const Vec<T, N>& x) {
  T m = abs_(x(0));
  for (int i = 1; i < N; ++i) {
    const T a = abs_(x(i));
    if (a > m) m = a;
  }
  return m;
}
template <class T, int N> T max_coeff(const Vec<T, N>& x) {
  T m = x(0);
  for (int i = 1; i < N; ++i)
    if (x(i) > m) m = x(i);
  return m;
}
template <class T, int N> bool all_finite(const Vec<T, N>& x) {
  for (int i = 0; i < N; ++i)
    if (!finite_(x(i))) return false;
  return true;
}

// C = A * B
template <class T, int R, int K, int C>
Mat<T, R, C> mul(const Mat<T, R, K>& A, const Mat<T, K, C>& B) {
  Mat<T, R, C> out;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) {
      T s = A(i, 0) * B(0, j);
      for (int k = 1; k < K; ++k) s = s + A(i, k) * B(k, j);
      out(i, j) = s;
    }
  return out;
}
// C = A^T * B
template <class T, int K, int R, int C>
Mat<T, R, C> mul_tn(const Mat<T, K, R>& A, const Mat<T, K, C>& B) {
  Mat<T, R, C> out;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) {
      T s = A(0, i) * B(0, j);
      for (int k = 1; k < K; ++k) s = s + A(k, i) * B(k, j);
      out(i, j) = s;
    }
  return out;
}
template <class T, int R, int C>
Mat<T, C, R> transpose(const Mat<T, R, C>& A) {
  Mat<T, C, R> out;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) out(j, i) = A(i, j);
  return out;
}
template <class T, int R, int C>
Vec<T, R> mul(const Mat<T, R, C>& A, const Vec<T, C>& x) {
  Vec<T, R> out;
  for (int i = 0; i < R; ++i) {
    T s = A(i, 0) * x(0);
    for (int k = 1; k < C; ++k) s = s + A(i, k) * x(k);
    out(i) = s;
  }
  return out;
}
template <class T, int R, int C>
Vec<T, C> mul_tn(const Mat<T, R, C>& A, const Vec<T, R>& x) {
  Vec<T, C> out;
  for (int j = 0; j < C; ++j) {
    T s = A(0, j) * x(0);
    for (int k = 1; k < R; ++k) s = s + A(k, j) * x(k);
    out(j) = s;
  }
  return out;
}
template <class T, int R, int C> T max_abs(const Mat<T, R, C>& A) {
  T m = abs_(A.a[0]);
  for (int i = 1; i < R * C; ++i) {
    const T a = abs_(A.a[i]);
    if (a > m) m = a;
  }
  return m;
}

// Eigen's 3x3 determinant (bruteforce_det3_helper expansion along row 0).
template <class T> T det3(const Mat<T, 3, 3>& m) {
  return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) -
         m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
         m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
}
// Eigen's 3x3 inverse: cofactors times 1/det with det from the first column.
template <class T> T cof3(const Mat<T, 3, 3>& m, int i, int j) {
  const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m(i1, j1) * m(i2, j2) - m(i1, j2) * m(i2, j1);
}
template <class T> Mat<T, 3, 3> inverse3(const Mat<T, 3, 3>& m) {
  const T c0 = cof3(m, 0, 0), c1 = cof3(m, 1, 0), c2 = cof3(m, 2, 0);
  const T det = c0 * m(0, 0) + c1 * m(1, 0) + c2 * m(2, 0);
  const T inv = T(1.0) / det;
  Mat<T, 3, 3> r;
  r(0, 0) = c0 * inv;
  r(0, 1) = c1 * inv;
  r(0, 2) = c2 * inv;
  for (int rr = 1; rr < 3; ++rr)
    for (int cc = 0; cc < 3; ++cc) r(rr, cc) = cof3(m, cc, rr) * inv;
  return r;
}

// ---------------------------------------------------------------------------
// Dense factorisations with Eigen's pivoting rules (the reference calls
// Eigen's FullPivLU, PartialPivLU, LLT and SelfAdjointEigenSolver)
// ---------------------------------------------------------------------------

// Full-pivot LU of an N x N matrix: the largest remaining entry, first in
// column-major order, is the pivot; invertible when every pivot exceeds
// N eps times the largest pivot (Eigen's default threshold).
template <class T, int N>
struct FullPivLU {
  Mat<T, N, N> lu;
  int row_t[N], col_t[N];
  int nonzero = N;
  T maxpivot = T(0.0);
  void compute(const Mat<T, N, N>& A) {
    lu = A;
    nonzero = N;
    maxpivot = T(0.0);
    for (int k = 0; k < N; ++k) {
      int br = k, bc = k;
      T big = abs_(lu(k, k));
      for (int j = k; j < N; ++j)
        for (int i = k; i < N; ++i) {
          const T a = abs_(lu(i, j));
          if (a > big) { big = a; br = i; bc = j; }
        }
      if (big == T(0.0)) {
        nonzero = k;
        for (int i = k; i < N; ++i) { row_t[i] = i; col_t[i] = i; }
        break;
      }
      if (big > maxpivot) maxpivot = big;
      row_t[k] = br;
      col_t[k] = bc;
      if (br != k)
        for (int j = 0; j < N; ++j) { T t = lu(k, j); lu(k, j) = lu(br, j); lu(br, j) = t; }
      if (bc != k)
        for (int i = 0; i < N; ++i) { T t = lu(i, k); lu(i, k) = lu(i, bc); lu(i, bc) = t; }
      if (k < N - 1) {
        for (int i = k + 1; i < N; ++i) lu(i, k) = lu(i, k) / lu(k, k);
        for (int j = k + 1; j < N; ++j)
          for (int i = k + 1; i < N; ++i) lu(i, j) = lu(i, j) - lu(i, k) * lu(k, j);
      }
    }
  }
  bool invertible() const {
    const T thr = maxpivot * T(kEps * N);
    int rank = 0;
    for (int i = 0; i < nonzero; ++i)
      if (abs_(lu(i, i)) > thr) ++rank;
    return rank == N;
  }
  // Solve A x = b for an invertible A.
  Vec<T, N> solve(const Vec<T, N>& b) const {
    Vec<T, N> c = b;
    for (int k = 0; k < N; ++k)
      if (row_t[k] != k) { T t = c(k); c(k) = c(row_t[k]); c(row_t[k]) = t; }
    for (int i = 1; i < N; ++i)
      for (int k = 0; k < i; ++k) c(i) = c(i) - lu(i, k) * c(k);
    for (int i = N - 1; i >= 0; --i) {
      for (int k = i + 1; k < N; ++k) c(i) = c(i) - lu(i, k) * c(k);
      c(i) = c(i) / lu(i, i);
    }
    for (int k = N - 1; k >= 0; --k)
      if (col_t[k] != k) { T t = c(k); c(k) = c(col_t[k]); c(col_t[k]) = t; }
    return c;
  }
};

// Partial-pivot LU (Eigen's unblocked algorithm: first largest |a_ik| in the
// column is the pivot).
template <class T, int N>
struct PartialPivLU {
  Mat<T, N, N> lu;
  int row_t[N];
  void compute(const Mat<T, N, N>& A) {
    lu = A;
    for (int k = 0; k < N; ++k) {
      int br = k;
      T big = abs_(lu(k, k));
      for (int i = k + 1; i < N; ++i) {
        const T a = abs_(lu(i, k));
        if (a > big) { big = a; br = i; }
      }
      row_t[k] = br;
      if (big != T(0.0)) {
        if (br != k)
          for (int j = 0; j < N; ++j) { T t = lu(k, j); lu(k, j) = lu(br, j); lu(br, j) = t; }
        for (int i = k + 1; i < N; ++i) lu(i, k) = lu(i, k) / lu(k, k);
      }
      if (k < N - 1)
        for (int i = k + 1; i < N; ++i)
          for (int j = k + 1; j < N; ++j) lu(i, j) = lu(i, j) - lu(i, k) * lu(k, j);
    }
  }
  template <int M>
  Mat<T, N, M> solve(const Mat<T, N, M>& B) const {
    Mat<T, N, M> c = B;
    for (int k = 0; k < N; ++k)
      if (row_t[k] != k)
        for (int j = 0; j < M; ++j) { T t = c(k, j); c(k, j) = c(row_t[k], j); c(row_t[k], j) = t; }
    for (int j = 0; j < M; ++j) {
      for (int i = 1; i < N; ++i)
        for (int k = 0; k < i; ++k) c(i, j) = c(i, j) - lu(i, k) * c(k, j);
      for (int i = N - 1; i >= 0; --i) {
        for (int k = i + 1; k < N; ++k) c(i, j) = c(i, j) - lu(i, k) * c(k, j);
        c(i, j) = c(i, j) / lu(i, i);
      }
    }
    return c;
  }
  Vec<T, N> solve(const Vec<T, N>& b) const {
    Mat<T, N, 1> B;
    for (int i = 0; i < N; ++i) B(i, 0) = b(i);
    const Mat<T, N, 1> X = solve<1>(B);
    Vec<T, N> x;
    for (int i = 0; i < N; ++i) x(i) = X(i, 0);
    return x;
  }
};

// Cholesky of the lower triangle (Eigen's unblocked LLT): fails at the first
// non-positive pivot, leaving the rest of the matrix as it was.
template <class T, int N>
struct LLT {
  Mat<T, N, N> l;
  bool ok = false;
  void compute(const Mat<T, N, N>& A) {
    l = A;
    ok = true;
    for (int k = 0; k < N; ++k) {
      T x = l(k, k);
      if (k > 0) {
        T s = l(k, 0) * l(k, 0);
        for (int j = 1; j < k; ++j) s = s + l(k, j) * l(k, j);
        x = x - s;
      }
      if (x <= T(0.0)) { ok = false; return; }
      x = sqrt_(x);
      l(k, k) = x;
      if (k > 0)
        for (int i = k + 1; i < N; ++i) {
          T s = l(i, 0) * l(k, 0);
          for (int j = 1; j < k; ++j) s = s + l(i, j) * l(k, j);
          l(i, k) = l(i, k) - s;
        }
      for (int i = k + 1; i < N; ++i) l(i, k) = l(i, k) / x;
    }
  }
  Vec<T, N> solve(const Vec<T, N>& b) const {
    Vec<T, N> y = b;
    for (int i = 0; i < N; ++i) {
      for (int k = 0; k < i; ++k) y(i) = y(i) - l(i, k) * y(k);
      y(i) = y(i) / l(i, i);
    }
    for (int i = N - 1; i >= 0; --i) {
      for (int k = i + 1; k < N; ++k) y(i) = y(i) - l(k, i) * y(k);
      y(i) = y(i) / l(i, i);
    }
    return y;
  }
};

// Cyclic Jacobi eigen-decomposition of a symmetric N x N matrix; eigenvalues
// ascending, eigenvectors as the columns of V (their signs are arbitrary, as
// in any eigensolver). A rotation is skipped once its off-diagonal entry no
// longer changes either diagonal entry at working precision.
template <class T, int N>
void jacobi_eigen(const Mat<T, N, N>& A, Vec<T, N>& w, Mat<T, N, N>& V, bool vectors) {
  Mat<T, N, N> a = A;
  V = Mat<T, N, N>::identity();
  for (int sweep = 0; sweep < 64; ++sweep) {
    bool any = false;
    for (int p = 0; p < N - 1; ++p)
      for (int q = p + 1; q < N; ++q) {
        const T apq = a(p, q);
        if (apq == T(0.0)) continue;
        const T g = T(100.0) * abs_(apq);
        const T app = abs_(a(p, p)), aqq = abs_(a(q, q));
        if (app + g == app && aqq + g == aqq) {
          a(p, q) = T(0.0);
          a(q, p) = T(0.0);
          continue;
        }
        any = true;
        const T h = a(q, q) - a(p, p);
        T t;
        if (abs_(h) + g == abs_(h)) {
          t = apq / h;
        } else {
          const T theta = T(0.5) * h / apq;
          t = T(1.0) / (abs_(theta) + sqrt_(T(1.0) + theta * theta));
          if (theta < T(0.0)) t = -t;
        }
        const T c = T(1.0) / sqrt_(T(1.0) + t * t);
        const T s = t * c;
        const T tau = s / (T(1.0) + c);
        a(p, p) = a(p, p) - t * apq;
        a(q, q) = a(q, q) + t * apq;
        a(p, q) = T(0.0);
        a(q, p) = T(0.0);
        for (int r = 0; r < N; ++r) {
          if (r == p || r == q) continue;
          const T gr = a(r, p), hr = a(r, q);
          a(r, p) = gr - s * (hr + gr * tau);
          a(p, r) = a(r, p);
          a(r, q) = hr + s * (gr - hr * tau);
          a(q, r) = a(r, q);
        }
        if (vectors)
          for (int r = 0; r < N; ++r) {
            const T gr = V(r, p), hr = V(r, q);
            V(r, p) = gr - s * (hr + gr * tau);
            V(r, q) = hr + s * (gr - hr * tau);
          }
      }
    if (!any) break;
  }
  for (int i = 0; i < N; ++i) w(i) = a(i, i);
  // Ascending order (insertion sort; columns follow).
  for (int i = 1; i < N; ++i)
    for (int j = i; j > 0 && w(j) < w(j - 1); --j) {
      T t = w(j); w(j) = w(j - 1); w(j - 1) = t;
      for (int r = 0; r < N; ++r) { T u = V(r, j); V(r, j) = V(r, j - 1); V(r, j - 1) = u; }
    }
}

}  // namespace oracle
