// The incompatible-mode Hex8 element evaluation on gfx950, FP64.
//
// Reference: src/fem/IncompatibleModes.cpp:320-600 (solve_and_condense,
// incompatible_elastoplastic_element), called for every element by
// src/fem/NonlinearSystem.hpp:192-237.
//
// Two kernels, one wavefront per workgroup, 8 lanes (one per integration
// point) per element:
//   local_solve_kernel  the local Newton iteration of the 9 mode parameters
//                       from the committed ones, as a per-team state machine
//                       that requests one pass at a time; each wavefront owns a
//                       chunk of elements and a team that finishes takes the
//                       next one, so the divergent iteration counts do not idle
//                       the wavefront. Writes alpha and a status per element.
//   condense_kernel     the pass at the converged alpha, K_aa's full-pivot LU,
//                       the coupling blocks K_ua, K_au, K_uu and the static
//                       condensation f* = f_u - K_ua K_aa^-1 r_a,
//                       K* = K_uu - K_ua K_aa^-1 K_au; writes f*, K*, the
//                       points' trial states, energy and yielding count.
// The pass at the converged alpha is the same arithmetic on the same data as
// the last pass of the local iteration, so both kernels agree bit for bit.
#include "element_kernels.hpp"
#include "element_pass.hpp"
#include "team_linalg.hpp"

namespace sbk {

constexpr int kWaveTeams = 8;
constexpr int kChunk = 64;  // elements per wavefront in the local solve

// IncompatibleModes.hpp:98-127
constexpr int kMaxLocalIterations = 25;
constexpr double kFullStepReduction = 1.0 - 1.0e-4;
constexpr int kMaxLineSearch = 8;
constexpr double kLineSearchRatio = 0.5;
constexpr double kLocalShift = 1.0e-3;
constexpr double kLocalMarquardt = 1.0e-6;
constexpr double kLocalTolerance = 1.0e-8;
constexpr double kLocalRoundOff = 1.0e-13;

// ---------------------------------------------------------------------------
// One pass of a team: the local sums r_alpha, scale, floor and K_aa
// (IncompatibleModes.cpp:252-275), reduced over the team.
// ---------------------------------------------------------------------------
struct LocalSums {
  double ra[9];
  double scale_max, floor_max;  // max over the 9 totals
};

template <int KIN>
__device__ __forceinline__ int team_local_pass(const ElementArgs& A, long long e, int p, const PointGeom& pg,
                                               const double ue[24], const double alpha[9], LuSlot& kaa_out,
                                               LocalSums& out) {
  const DevMaterial& m = A.mat;
  const bool has_point = p < A.rule.n;
  PointResponse r;
  int fail = 0;
  if (has_point) {
    PointState st;
    load_state(m, A.state_in, A.state_stride, e * A.rule.n + p, st);
    fail = point_response<KIN>(m, pg, ue, alpha, st, r);
  }
  // The reference evaluates every point's kinematics before any return, so
  // the reported failure is the first kinematic one, else the first return.
  const int stage = (fail == kReturnFailed || fail == kNonFiniteTrial) ? 1 : 0;
  const int first = team_imin(fail != 0 ? (stage << 12) | (p << 8) | fail : 1 << 20);
  if (first != (1 << 20)) return first & 0xFF;
  // B_alpha (6 x 9) of this point and the pieces of the sums
  double ba[6][9];
  double w = 0.0;
  if (has_point) {
    w = pg.w;
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
      for (int n = 0; n < 9; ++n) ba[c][n] = bnl(r.F, pg.g, c, 8 + n / 3, n % 3);
  } else {
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
      for (int n = 0; n < 9; ++n) ba[c][n] = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      r.S[c] = 0.0;
#pragma unroll
      for (int d = 0; d < 6; ++d) r.mat[c][d] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) r.h[i][j] = 0.0;
  }
  double hmax = 0.0, mmax = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) hmax = fmax(hmax, fabs(r.h[i][j]));
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) mmax = fmax(mmax, fabs(r.mat[i][j]));
  const double ff = w * mmax * hmax * (1.0 + hmax);
  double smax = 0.0, fmaxv = 0.0;
#pragma unroll
  for (int n = 0; n < 9; ++n) {
    double sr = ba[0][n] * r.S[0], ss = fabs(ba[0][n]) * fabs(r.S[0]), sf = fabs(ba[0][n]);
#pragma unroll
    for (int c = 1; c < 6; ++c) {
      sr = sr + ba[c][n] * r.S[c];
      ss = ss + fabs(ba[c][n]) * fabs(r.S[c]);
      sf = sf + fabs(ba[c][n]);
    }
    out.ra[n] = team_sum(w * sr);
    smax = fmax(smax, team_sum(w * ss));
    fmaxv = fmax(fmaxv, team_sum(ff * sf));
  }
  out.scale_max = smax;
  out.floor_max = fmaxv;
  // K_aa = sum_q w B_a^T mat B_a (+ the geometric alpha-alpha block)
  double y[6][9];
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int n = 0; n < 9; ++n) {
      double s = r.mat[c][0] * ba[0][n];
#pragma unroll
      for (int d = 1; d < 6; ++d) s = s + r.mat[c][d] * ba[d][n];
      y[c][n] = s;
    }
  double geo[3][3];
  if constexpr (KIN != kSmall) {
    // gt^T S gt with gt = G~ (3 x 3)
    const double S3[3][3] = {{r.S[0], r.S[3], r.S[5]}, {r.S[3], r.S[1], r.S[4]}, {r.S[5], r.S[4], r.S[2]}};
    double gs[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        gs[a][j] = pg.g[0][8 + a] * S3[0][j] + pg.g[1][8 + a] * S3[1][j] + pg.g[2][8 + a] * S3[2][j];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        geo[a][b] = gs[a][0] * pg.g[0][8 + b] + gs[a][1] * pg.g[1][8 + b] + gs[a][2] * pg.g[2][8 + b];
  }
#pragma unroll
  for (int mm = 0; mm < 9; ++mm)
#pragma unroll
    for (int n = 0; n < 9; ++n) {
      double s = ba[0][mm] * y[0][n];
#pragma unroll
      for (int c = 1; c < 6; ++c) s = s + ba[c][mm] * y[c][n];
      double v = w * s;
      if constexpr (KIN != kSmall) {
        if (mm % 3 == n % 3) v = v + w * geo[mm / 3][n / 3];
      }
      const double tot = team_sum(v);
      if (((mm * 9 + n) & 7) == p) kaa_out.a[mm * 9 + n] = tot;
    }
  team_sync();
  return 0;
}

// ---------------------------------------------------------------------------
// The local Newton iteration as a state machine (IncompatibleModes.cpp:320-546)
// ---------------------------------------------------------------------------
enum Phase : int {
  kIdle = 0,
  kFirst,       // pass(alpha), must succeed
  kLsReach,     // line search: try_pass(alpha + reach d)
  kLsFalsi,     // line search: try_pass(alpha + step d), regula falsi
  kLsFinal,     // line search: pass(alpha + s0 d), must succeed
  kSnTrial,     // search on |r_a|: try_pass(alpha + trial d)
  kSnFinal,     // search on |r_a|: pass(alpha + best d), must succeed
  kIndefFull,   // indefinite sym(K_aa): try_pass(alpha + newton)
  kLm,          // Levenberg-Marquardt: try_pass(alpha + d)
};
enum LsContext : int { kLsSpd = 0, kLsShifted = 1 };
enum SnContext : int { kSnFromLs = 0, kSnAfterShift = 1 };

struct TeamSlot {
  LuSlot lu;         // K_aa and its factors
  double k0[81];     // K_aa of the iteration (the fallbacks need it)
  double w[81];      // work matrix of the fallbacks
};

struct Machine {
  int phase, ls_ctx, sn_ctx, iterations, halving, search, side;
  bool last_is_best, moved;
  double alpha[9], d[9], newton[9], r0[9];
  double residual, s;  // s: the trial is alpha + s d
  double phi0, reach, s0, f0, s1, f1, step;
  double best_step, best, trial_step, mu;
};

__device__ __forceinline__ double norm9(const double v[9]) {
  double s = v[0] * v[0];
#pragma unroll
  for (int i = 1; i < 9; ++i) s = s + v[i] * v[i];
  return sqrt(s);
}
__device__ __forceinline__ double maxabs9(const double v[9]) {
  double s = fabs(v[0]);
#pragma unroll
  for (int i = 1; i < 9; ++i) s = fmax(s, fabs(v[i]));
  return s;
}
__device__ __forceinline__ double dot9(const double a[9], const double b[9]) {
  double s = a[0] * b[0];
#pragma unroll
  for (int i = 1; i < 9; ++i) s = s + a[i] * b[i];
  return s;
}
__device__ __forceinline__ void axpy9(double a[9], double s, const double d[9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = a[i] + s * d[i];
}

// The outcome of a step of the machine: keep going (a new trial is
// requested), converged, or failed with a status.
enum StepResult { kContinue = 0, kConverged = 1, kFailed = 2 };

// The top of the local loop (IncompatibleModes.cpp:457-515): convergence
// test, K_aa's LU, Newton's direction and the choice of the search.
__device__ int machine_top(Machine& M, TeamSlot& T, const LocalSums& cur, int p, int& status, double& value) {
  const double largest = maxabs9(cur.ra);
  const double scale = cur.scale_max;
  const double floor = kLocalRoundOff * cur.floor_max;
  // K_aa is in T.lu.a (the pass wrote it); keep a copy for the fallbacks.
  for (int i = p; i < 81; i += 8) T.k0[i] = T.lu.a[i];
  team_sync();
  const bool inv = coop_full_lu(T.lu, p);
  if (!inv) { status = kSingularModes; value = largest; return kFailed; }
  if (largest <= kLocalTolerance * scale + floor || largest == 0.0) return kConverged;
  if (M.iterations == kMaxLocalIterations || !isfinite(largest)) {
    status = kLocalNotConverged;
    value = largest;
    return kFailed;
  }
  ++M.iterations;
#pragma unroll
  for (int i = 0; i < 9; ++i) M.r0[i] = cur.ra[i];
  double nt[9];
  lu_solve(T.lu, M.r0, nt);
#pragma unroll
  for (int i = 0; i < 9; ++i) M.newton[i] = -nt[i];
  // sym(K_aa) positive definite?
  for (int i = p; i < 81; i += 8) {
    const int r = i / 9, c = i % 9;
    T.w[i] = 0.5 * (T.k0[r * 9 + c] + T.k0[c * 9 + r]);
  }
  team_sync();
  const bool spd = llt9(T.w);
  team_sync();
  if (spd) {
    // line_search(newton, r0), SPD context
    M.ls_ctx = kLsSpd;
#pragma unroll
    for (int i = 0; i < 9; ++i) M.d[i] = M.newton[i];
    M.phi0 = dot9(M.d, M.r0);
    M.reach = 1.0;
    M.halving = 0;
    M.s = 1.0;
    M.phase = kLsReach;
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) M.d[i] = M.newton[i];
    M.s = 1.0;
    M.phase = kIndefFull;
  }
  return kContinue;
}

__device__ __forceinline__ void sn_start(Machine& M, double reach) {
  M.best_step = 0.0;
  M.best = M.residual;
  M.last_is_best = false;
  M.trial_step = reach;
  M.search = 0;
  M.s = reach;
  M.phase = kSnTrial;
}

__device__ __forceinline__ void lm_start(Machine& M, TeamSlot& T, int p) {
  // normal = k0^T k0 into w; mu = 1e-6 max diag
  for (int i = p; i < 81; i += 8) {
    const int r = i / 9, c = i % 9;
    double s = T.k0[r] * T.k0[c];
    for (int k = 1; k < 9; ++k) s = s + T.k0[k * 9 + r] * T.k0[k * 9 + c];
    T.w[i] = s;
  }
  team_sync();
  double dmax = T.w[0];
  for (int i = 1; i < 9; ++i) dmax = fmax(dmax, T.w[i * 10]);
  M.mu = kLocalMarquardt * dmax;
  M.moved = false;
  M.search = 0;
}

// The next Levenberg-Marquardt trial, or the end of the fallback.
__device__ int lm_next(Machine& M, TeamSlot& T, int p, int& status, double& value, const LocalSums& cur,
                       bool& to_top) {
  to_top = false;
  if (M.search < kMaxLineSearch && !M.moved) {
    // (normal + mu I) d = -k0^T r0; normal rebuilt from k0 each trial
    double grad[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      double s = T.k0[c] * M.r0[0];
      for (int k = 1; k < 9; ++k) s = s + T.k0[k * 9 + c] * M.r0[k];
      grad[c] = s;
    }
    team_sync();
    for (int i = 0; i < 81; ++i) {
      const int r = i / 9, c = i % 9;
      double s = T.k0[r] * T.k0[c];
      for (int k = 1; k < 9; ++k) s = s + T.k0[k * 9 + r] * T.k0[k * 9 + c];
      T.w[i] = (r == c) ? s + M.mu : s;
    }
    llt9(T.w);
    double x[9];
    llt9_solve(T.w, grad, x);
    team_sync();
#pragma unroll
    for (int i = 0; i < 9; ++i) M.d[i] = -x[i];
    M.s = 1.0;
    M.phase = kLm;
    return kContinue;
  }
  if (!M.moved) {
    status = kNoDescent;
    value = 0.0;
    return kFailed;
  }
  to_top = true;
  return kContinue;
}

// Advance the machine after a pass at alpha + s d (ok: the pass succeeded;
// cur: its sums). Loops internally through the transitions that need no
// new pass (the top of the iteration, the ends of the searches).
__device__ int machine_step(Machine& M, TeamSlot& T, bool ok, const LocalSums& cur, int p, int& status,
                            double& value) {
  // The searches return through these flags.
  enum { kNone, kLsTrue, kLsFalse, kSnTrue, kSnFalse, kTop, kLmNext } next = kNone;
  switch (M.phase) {
    case kFirst:
      if (!ok) return kFailed;  // status set by the pass
      M.residual = norm9(cur.ra);
      next = kTop;
      break;
    case kLsReach:
      if (!ok) {
        if (M.halving < kMaxLineSearch) {
          M.reach = M.reach * 0.5;
          ++M.halving;
          M.s = M.reach;
          return kContinue;
        }
        next = kLsFalse;
        break;
      } else {
        const double full = norm9(cur.ra);
        if (full <= kFullStepReduction * M.residual ||
            maxabs9(cur.ra) <= kLocalTolerance * cur.scale_max + kLocalRoundOff * cur.floor_max) {
          axpy9(M.alpha, M.reach, M.d);
          M.residual = full;
          next = kLsTrue;
          break;
        }
        const double phi1 = dot9(M.d, cur.ra);
        if (!(M.phi0 < 0.0) || !(phi1 > 0.0) || !isfinite(phi1)) {
          M.sn_ctx = kSnFromLs;
          sn_start(M, 0.5 * M.reach);
          return kContinue;
        }
        M.s0 = 0.0; M.f0 = M.phi0; M.s1 = M.reach; M.f1 = phi1; M.step = M.reach;
        M.side = 0;
        M.search = 0;
        const double width = M.s1 - M.s0;
        const double raw = (M.s0 * M.f1 - M.s1 * M.f0) / (M.f1 - M.f0);
        const double lo = M.s0 + 0.1 * width, hi = M.s1 - 0.1 * width;
        M.step = raw < lo ? lo : (hi < raw ? hi : raw);
        M.s = M.step;
        M.phase = kLsFalsi;
        return kContinue;
      }
    case kLsFalsi: {
      const bool evaluated = ok;
      const double f = evaluated ? dot9(M.d, cur.ra) : __builtin_inf();
      bool end = false;
      if (evaluated && fabs(f) <= kLineSearchRatio * fabs(M.phi0)) {
        end = true;
      } else {
        if (!evaluated || f > 0.0) {
          M.s1 = M.step;
          M.f1 = evaluated ? f : 0.5 * M.f1;
          if (M.side == 1) M.f0 = M.f0 * 0.5;
          M.side = 1;
        } else {
          M.s0 = M.step;
          M.f0 = f;
          if (M.side == -1) M.f1 = M.f1 * 0.5;
          M.side = -1;
        }
        ++M.search;
        if (M.search < kMaxLineSearch) {
          const double width = M.s1 - M.s0;
          const double raw = (M.s0 * M.f1 - M.s1 * M.f0) / (M.f1 - M.f0);
          const double lo = M.s0 + 0.1 * width, hi = M.s1 - 0.1 * width;
          M.step = raw < lo ? lo : (hi < raw ? hi : raw);
          M.s = M.step;
          return kContinue;
        }
        end = true;
      }
      if (end) {
        if (!evaluated) {
          if (M.s0 == 0.0) { next = kLsFalse; break; }
          M.step = M.s0;
          M.s = M.s0;
          M.phase = kLsFinal;
          return kContinue;
        }
        axpy9(M.alpha, M.step, M.d);
        M.residual = norm9(cur.ra);
        next = kLsTrue;
      }
      break;
    }
    case kLsFinal:
      if (!ok) return kFailed;
      axpy9(M.alpha, M.step, M.d);
      M.residual = norm9(cur.ra);
      next = kLsTrue;
      break;
    case kSnTrial: {
      bool end = false;
      if (!ok) {
        M.last_is_best = false;
      } else {
        const double n = norm9(cur.ra);
        M.last_is_best = n < M.best;
        if (M.last_is_best) { M.best = n; M.best_step = M.trial_step; }
        if (n <= kFullStepReduction * M.residual) end = true;
      }
      if (!end) {
        ++M.search;
        M.trial_step = M.trial_step * 0.5;
        if (M.search < kMaxLineSearch) {
          M.s = M.trial_step;
          return kContinue;
        }
      }
      if (M.best_step == 0.0) { next = kSnFalse; break; }
      if (!M.last_is_best) {
        M.s = M.best_step;
        M.phase = kSnFinal;
        return kContinue;
      }
      axpy9(M.alpha, M.best_step, M.d);
      M.residual = M.best;
      next = kSnTrue;
      break;
    }
    case kSnFinal:
      if (!ok) return kFailed;
      axpy9(M.alpha, M.best_step, M.d);
      M.residual = M.best;
      next = kSnTrue;
      break;
    case kIndefFull: {
      if (ok) {
        const double full = norm9(cur.ra);
        if (full <= kFullStepReduction * M.residual) {
          axpy9(M.alpha, 1.0, M.newton);
          M.residual = full;
          next = kTop;
          break;
        }
      }
      // shifted direction: (K_aa + mu I) d = -r0, mu from sym's spectrum
      for (int i = 0; i < 81; ++i) {
        const int r = i / 9, c = i % 9;
        T.w[i] = 0.5 * (T.k0[r * 9 + c] + T.k0[c * 9 + r]);
      }
      double ev[9];
      jacobi9_values(T.w, ev);
      double top = fabs(ev[0]);
      for (int i = 1; i < 9; ++i) top = fmax(top, fabs(ev[i]));
      const double mu = kLocalShift * top - ev[0];
      team_sync();
      for (int i = 0; i < 81; ++i) T.w[i] = (i % 10 == 0) ? T.k0[i] + mu : T.k0[i];
      double x[9];
      plu9_solve(T.w, M.r0, x);
      team_sync();
#pragma unroll
      for (int i = 0; i < 9; ++i) M.d[i] = -x[i];
      M.ls_ctx = kLsShifted;
      M.phi0 = dot9(M.d, M.r0);
      M.reach = 1.0;
      M.halving = 0;
      M.s = 1.0;
      M.phase = kLsReach;
      return kContinue;
    }
    case kLm:
      if (ok) {
        const double n = norm9(cur.ra);
        if (n < M.residual) {
          axpy9(M.alpha, 1.0, M.d);
          M.residual = n;
          M.moved = true;
        }
      }
      ++M.search;
      M.mu = M.mu * 10.0;
      next = kLmNext;
      break;
    default:
      return kFailed;
  }
  // Transitions without a pass.
  for (int guard = 0; guard < 8; ++guard) {
    switch (next) {
      case kLsTrue:
        next = kTop;
        break;
      case kLsFalse:
        if (M.ls_ctx == kLsSpd) {
          lm_start(M, T, p);
          next = kLmNext;
        } else {
#pragma unroll
          for (int i = 0; i < 9; ++i) M.d[i] = M.newton[i];
          M.sn_ctx = kSnAfterShift;
          sn_start(M, 0.5);
          return kContinue;
        }
        break;
      case kSnTrue:
        next = (M.sn_ctx == kSnFromLs) ? kLsTrue : kTop;
        break;
      case kSnFalse:
        if (M.sn_ctx == kSnFromLs) {
          next = kLsFalse;
        } else {
          lm_start(M, T, p);
          next = kLmNext;
        }
        break;
      case kLmNext: {
        bool to_top = false;
        const int r = lm_next(M, T, p, status, value, cur, to_top);
        if (r == kFailed) return kFailed;
        if (!to_top) return kContinue;
        next = kTop;
        break;
      }
      case kTop:
        return machine_top(M, T, cur, p, status, value);
      default:
        return kFailed;
    }
  }
  return kFailed;
}

template <int KIN>
__global__ __launch_bounds__(64) void local_solve_kernel(ElementArgs A) {
  __shared__ TeamSlot slots[kWaveTeams];
  const int lane = threadIdx.x;
  const int t = lane >> 3;
  const int p = lane & 7;
  TeamSlot& T = slots[t];
  const long long c0 = static_cast<long long>(blockIdx.x) * kChunk;
  const long long c1 = c0 + kChunk < A.ne ? c0 + kChunk : A.ne;
  long long next = c0;
  long long e = -1;
  Machine M;
  M.phase = kIdle;
  PointGeom pg;
  double ue[24];
  for (;;) {
    // Hand the chunk's next elements to the idle teams, in team order.
    const unsigned long long leaders = __ballot(M.phase == kIdle && p == 0);
    const int before = __popcll(leaders & ((1ull << (t * 8)) - 1ull));
    const int free_teams = __popcll(leaders);
    bool fresh = false;
    if (M.phase == kIdle) {
      const long long cand = next + before;
      if (cand < c1) {
        e = cand;
        fresh = true;
      } else {
        e = -1;
      }
    }
    next += free_teams;
    if (__ballot(M.phase != kIdle || fresh) == 0ull) break;
    if (fresh) {
      double x0[8][3];
      load_element(A.coords, A.conn, A.u, e, x0, ue);
      double fv = 0.0;
      int st = kOk;
      if (p < A.rule.n) st = setup_point(x0, ue, A.rule.xi[p], A.rule.eta[p], A.rule.zeta[p], A.rule.w[p], pg, fv);
      // the first failing point in point order, as the reference's set-up loop
      const int code = team_imin(st != kOk ? (p << 8) | st : 1 << 20);
      if (code != (1 << 20)) {
        if (p == (code >> 8)) { A.status[e] = code & 0xFF; A.fail_value[e] = fv; }
        continue;  // the team stays idle and takes another element
      }
#pragma unroll
      for (int i = 0; i < 9; ++i) M.alpha[i] = A.alpha_in ? A.alpha_in[e * 9 + i] : 0.0;
      M.iterations = 0;
      M.s = 0.0;
      M.phase = kFirst;
    }
    if (M.phase != kIdle) {
      double trial[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) trial[i] = M.alpha[i] + M.s * M.d[i];
      if (M.phase == kFirst) {
#pragma unroll
        for (int i = 0; i < 9; ++i) { trial[i] = M.alpha[i]; M.d[i] = 0.0; }
      }
      LocalSums cur;
      const int fail = team_local_pass<KIN>(A, e, p, pg, ue, trial, T.lu, cur);
      const bool ok = fail == 0;
      int status = ok ? kOk : fail;
      double value = 0.0;
      const int res = machine_step(M, T, ok, cur, p, status, value);
      if (res == kConverged) {
        if (p == 0) {
          A.status[e] = kOk;
          A.iterations[e] = M.iterations;
        }
        if (p < 9) A.alpha_out[e * 9 + p] = M.alpha[p];
        if (p == 0) A.alpha_out[e * 9 + 8] = M.alpha[8];
        M.phase = kIdle;
      } else if (res == kFailed) {
        if (p == 0) {
          A.status[e] = status == kOk ? kUnsupported : status;
          A.fail_value[e] = value;
          A.iterations[e] = M.iterations;
        }
        M.phase = kIdle;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// The final pass and the condensation
// ---------------------------------------------------------------------------
struct CondenseSlot {
  LuSlot lu;
  double region[216];  // staging of one point, then K_aa^-1 K_au, then K rows
};

template <int KIN, bool TANGENT>
__global__ __launch_bounds__(64) void condense_kernel(ElementArgs A) {
  __shared__ CondenseSlot slots[kWaveTeams];
  const int lane = threadIdx.x;
  const int t = lane >> 3;
  const int p = lane & 7;
  CondenseSlot& T = slots[t];
  const long long e = static_cast<long long>(blockIdx.x) * kWaveTeams + t;
  if (e >= A.ne) return;
  if (A.status[e] != kOk) return;
  const DevMaterial& m = A.mat;
  const bool has_point = p < A.rule.n;
  double x0[8][3], ue[24];
  load_element(A.coords, A.conn, A.u, e, x0, ue);
  PointGeom pg;
  double fv = 0.0;
  if (has_point) setup_point(x0, ue, A.rule.xi[p], A.rule.eta[p], A.rule.zeta[p], A.rule.w[p], pg, fv);
  double alpha[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) alpha[i] = A.alpha_out[e * 9 + i];
  // the pass at the converged alpha
  PointResponse r;
  PointState st;
  int fail = 0;
  if (has_point) {
    load_state(m, A.state_in, A.state_stride, e * A.rule.n + p, st);
    fail = point_response<KIN>(m, pg, ue, alpha, st, r);
  } else {
    r.energy = 0.0;
    r.yielding = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      r.S[c] = 0.0;
#pragma unroll
      for (int d = 0; d < 6; ++d) r.mat[c][d] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) { r.F[i][j] = 0.0; r.h[i][j] = 0.0; }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 11; ++j) pg.g[i][j] = 0.0;
    pg.w = 0.0;
  }
  if (team_or(fail != 0 ? 1 : 0)) {
    if (p == 0) A.status[e] = kUnsupported;
    return;
  }
  const double w = has_point ? pg.w : 0.0;
  // sums of the pass: r (33), energy, yielding; K_aa for the LU
  double ba[6][9];
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int n = 0; n < 9; ++n) ba[c][n] = has_point ? bnl(r.F, pg.g, c, 8 + n / 3, n % 3) : 0.0;
  double ra[9];
#pragma unroll
  for (int n = 0; n < 9; ++n) {
    double s = ba[0][n] * r.S[0];
#pragma unroll
    for (int c = 1; c < 6; ++c) s = s + ba[c][n] * r.S[c];
    ra[n] = team_sum(w * s);
  }
  double ru[3];  // f_u rows of this lane's node a = p
#pragma unroll
  for (int k = 0; k < 3; ++k) ru[k] = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double s = has_point ? bnl(r.F, pg.g, 0, a, k) * r.S[0] : 0.0;
#pragma unroll
      for (int c = 1; c < 6; ++c) s = s + (has_point ? bnl(r.F, pg.g, c, a, k) * r.S[c] : 0.0);
      const double tot = team_sum(w * s);
      if (a == p) ru[k] = tot;
    }
  const double energy = team_sum(w * (has_point ? r.energy : 0.0));
  const int yielding = team_isum(has_point && r.yielding ? 1 : 0);
  {
    double y[6][9];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
      for (int n = 0; n < 9; ++n) {
        double s = r.mat[c][0] * ba[0][n];
#pragma unroll
        for (int d = 1; d < 6; ++d) s = s + r.mat[c][d] * ba[d][n];
        y[c][n] = s;
      }
    double geo[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if constexpr (KIN != kSmall) {
      const double S3[3][3] = {{r.S[0], r.S[3], r.S[5]}, {r.S[3], r.S[1], r.S[4]}, {r.S[5], r.S[4], r.S[2]}};
      double gs[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          gs[a][j] = pg.g[0][8 + a] * S3[0][j] + pg.g[1][8 + a] * S3[1][j] + pg.g[2][8 + a] * S3[2][j];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
          geo[a][b] = gs[a][0] * pg.g[0][8 + b] + gs[a][1] * pg.g[1][8 + b] + gs[a][2] * pg.g[2][8 + b];
    }
#pragma unroll
    for (int mm = 0; mm < 9; ++mm)
#pragma unroll
      for (int n = 0; n < 9; ++n) {
        double s = ba[0][mm] * y[0][n];
#pragma unroll
        for (int c = 1; c < 6; ++c) s = s + ba[c][mm] * y[c][n];
        double v = w * s;
        if constexpr (KIN != kSmall) {
          if (mm % 3 == n % 3) v = v + w * geo[mm / 3][n / 3];
        }
        const double tot = team_sum(v);
        if (((mm * 9 + n) & 7) == p) T.lu.a[mm * 9 + n] = tot;
      }
  }
  team_sync();
  coop_full_lu(T.lu, p);
  // trial states of the points
  if (has_point) store_state(m, A.state_out, A.state_stride, e * A.rule.n + p, st);
  if (p == 0) {
    A.energy[e] = energy;
    A.yielding[e] = yielding;
  }
  // Coupling blocks, point by point through the staging region: lane a owns
  // rows 3a..3a+2 of K_ua (and K_uu) and columns 3a..3a+2 of K_au.
  const int a = p;
  double kua[3][9];
  double kau[9][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int n = 0; n < 9; ++n) { kua[k][n] = 0.0; kau[n][k] = 0.0; }
  const bool nonsym = !m.symmetric;
  for (int q = 0; q < A.rule.n; ++q) {
    if (p == q) {
      double* s = T.region;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s[i * 3 + j] = r.F[i][j];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 11; ++j) s[9 + i * 11 + j] = pg.g[i][j];
#pragma unroll
      for (int i = 0; i < 36; ++i) s[42 + i] = r.mat[i / 6][i % 6];
#pragma unroll
      for (int i = 0; i < 6; ++i) s[78 + i] = r.S[i];
      s[84] = pg.w;
    }
    team_sync();
    double Fq[3][3], gq[3][11], Mq[6][6], Sq[6];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Fq[i][j] = T.region[i * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 11; ++j) gq[i][j] = T.region[9 + i * 11 + j];
#pragma unroll
    for (int i = 0; i < 36; ++i) Mq[i / 6][i % 6] = T.region[42 + i];
#pragma unroll
    for (int i = 0; i < 6; ++i) Sq[i] = T.region[78 + i];
    const double wq = T.region[84];
    team_sync();
    // z_k = B_u[:, 3a+k]^T mat (6)
    double z[3][6];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 6; ++d) {
        double s = bnl(Fq, gq, 0, a, k) * Mq[0][d];
#pragma unroll
        for (int c = 1; c < 6; ++c) s = s + bnl(Fq, gq, c, a, k) * Mq[c][d];
        z[k][d] = s;
      }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int n = 0; n < 9; ++n) {
        double s = z[k][0] * bnl(Fq, gq, 0, 8 + n / 3, n % 3);
#pragma unroll
        for (int d = 1; d < 6; ++d) s = s + z[k][d] * bnl(Fq, gq, d, 8 + n / 3, n % 3);
        kua[k][n] = kua[k][n] + wq * s;
      }
    if (nonsym && TANGENT) {
      // K_au[:, 3a+k] = B_a^T (mat B_u[:, 3a+k])
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double y[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          double s = Mq[c][0] * bnl(Fq, gq, 0, a, k);
#pragma unroll
          for (int d = 1; d < 6; ++d) s = s + Mq[c][d] * bnl(Fq, gq, d, a, k);
          y[c] = s;
        }
#pragma unroll
        for (int n = 0; n < 9; ++n) {
          double s = bnl(Fq, gq, 0, 8 + n / 3, n % 3) * y[0];
#pragma unroll
          for (int c = 1; c < 6; ++c) s = s + bnl(Fq, gq, c, 8 + n / 3, n % 3) * y[c];
          kau[n][k] = kau[n][k] + wq * s;
        }
      }
    }
    if constexpr (KIN != kSmall) {
      // geometric (G_aug^T S G_aug) (x) I: block (a, b) for the modes b = 8..10
      double gsa[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double S3j0 = j == 0 ? Sq[0] : (j == 1 ? Sq[3] : Sq[5]);
        const double S3j1 = j == 0 ? Sq[3] : (j == 1 ? Sq[1] : Sq[4]);
        const double S3j2 = j == 0 ? Sq[5] : (j == 1 ? Sq[4] : Sq[2]);
        gsa[j] = gq[0][a] * S3j0 + gq[1][a] * S3j1 + gq[2][a] * S3j2;
      }
#pragma unroll
      for (int bm = 0; bm < 3; ++bm) {
        const double geo = gsa[0] * gq[0][8 + bm] + gsa[1] * gq[1][8 + bm] + gsa[2] * gq[2][8 + bm];
        const double v = wq * geo;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          kua[k][3 * bm + k] = kua[k][3 * bm + k] + v;
          if (nonsym && TANGENT) kau[3 * bm + k][k] = kau[3 * bm + k][k] + v;
        }
      }
    }
  }
  // The force: f* = f_u - K_ua K_aa^-1 r_a
  double corr[9];
  lu_solve(T.lu, ra, corr);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double s = kua[k][0] * corr[0];
#pragma unroll
    for (int n = 1; n < 9; ++n) s = s + kua[k][n] * corr[n];
    A.force[e * 24 + 3 * a + k] = ru[k] - s;
  }
  if constexpr (!TANGENT) return;
  // coupling = K_aa^-1 K_au (or K_ua^T): lane a solves its columns 3a..3a+2
  double cpl[9][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double col[9], x[9];
#pragma unroll
    for (int n = 0; n < 9; ++n) col[n] = nonsym ? kau[n][k] : kua[k][n];
    lu_solve(T.lu, col, x);
#pragma unroll
    for (int n = 0; n < 9; ++n) cpl[n][k] = x[n];
  }
  // K_uu rows of node a, point by point again
  double kuu[3][24];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 24; ++j) kuu[k][j] = 0.0;
  team_sync();
  for (int q = 0; q < A.rule.n; ++q) {
    if (p == q) {
      double* s = T.region;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s[i * 3 + j] = r.F[i][j];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 11; ++j) s[9 + i * 11 + j] = pg.g[i][j];
#pragma unroll
      for (int i = 0; i < 36; ++i) s[42 + i] = r.mat[i / 6][i % 6];
#pragma unroll
      for (int i = 0; i < 6; ++i) s[78 + i] = r.S[i];
      s[84] = pg.w;
    }
    team_sync();
    double Fq[3][3], gq[3][11], Mq[6][6], Sq[6];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Fq[i][j] = T.region[i * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 11; ++j) gq[i][j] = T.region[9 + i * 11 + j];
#pragma unroll
    for (int i = 0; i < 36; ++i) Mq[i / 6][i % 6] = T.region[42 + i];
#pragma unroll
    for (int i = 0; i < 6; ++i) Sq[i] = T.region[78 + i];
    const double wq = T.region[84];
    team_sync();
    double z[3][6];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 6; ++d) {
        double s = bnl(Fq, gq, 0, a, k) * Mq[0][d];
#pragma unroll
        for (int c = 1; c < 6; ++c) s = s + bnl(Fq, gq, c, a, k) * Mq[c][d];
        z[k][d] = s;
      }
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int kb = 0; kb < 3; ++kb) {
        double bcol[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) bcol[d] = bnl(Fq, gq, d, b, kb);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          double s = z[k][0] * bcol[0];
#pragma unroll
          for (int d = 1; d < 6; ++d) s = s + z[k][d] * bcol[d];
          kuu[k][3 * b + kb] = kuu[k][3 * b + kb] + wq * s;
        }
      }
    if constexpr (KIN != kSmall) {
      double gsa[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double S3j0 = j == 0 ? Sq[0] : (j == 1 ? Sq[3] : Sq[5]);
        const double S3j1 = j == 0 ? Sq[3] : (j == 1 ? Sq[1] : Sq[4]);
        const double S3j2 = j == 0 ? Sq[5] : (j == 1 ? Sq[4] : Sq[2]);
        gsa[j] = gq[0][a] * S3j0 + gq[1][a] * S3j1 + gq[2][a] * S3j2;
      }
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const double geo = gsa[0] * gq[0][b] + gsa[1] * gq[1][b] + gsa[2] * gq[2][b];
        const double v = wq * geo;
#pragma unroll
        for (int k = 0; k < 3; ++k) kuu[k][3 * b + k] = kuu[k][3 * b + k] + v;
      }
    }
  }
  // publish the coupling columns, then K* rows = K_uu - K_ua coupling
#pragma unroll
  for (int n = 0; n < 9; ++n)
#pragma unroll
    for (int k = 0; k < 3; ++k) T.region[n * 24 + 3 * a + k] = cpl[n][k];
  team_sync();
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 24; ++j) {
      double s = kua[k][0] * T.region[j];
#pragma unroll
      for (int n = 1; n < 9; ++n) s = s + kua[k][n] * T.region[n * 24 + j];
      kuu[k][j] = kuu[k][j] - s;
    }
  team_sync();
  if (!nonsym) {
    // K = (K + K^T) / 2 (IncompatibleModes.cpp:557), node block by node block.
    // Sweep 1, round rr: lane rr publishes its (untouched) rows; every lane
    // a > rr averages its block (a, rr) with the transpose of block (rr, a).
    // Lane rr's blocks (rr, c) with c > rr are untouched until sweep 2.
#pragma unroll
    for (int rr = 0; rr < 7; ++rr) {
      if (p == rr) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int j = 0; j < 24; ++j) T.region[k * 24 + j] = kuu[k][j];
      }
      team_sync();
      if (a > rr) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int k2 = 0; k2 < 3; ++k2)
            kuu[k][3 * rr + k2] = 0.5 * (kuu[k][3 * rr + k2] + T.region[k2 * 24 + 3 * a + k]);
      }
      team_sync();
    }
    // Diagonal blocks: both halves are this lane's own entries.
    {
      double dgb[3][3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int k2 = 0; k2 < 3; ++k2) {
          double mine = 0.0, other = 0.0;
#pragma unroll
          for (int b = 0; b < 8; ++b)
            if (b == a) { mine = kuu[k][3 * b + k2]; other = kuu[k2][3 * b + k]; }
          dgb[k][k2] = 0.5 * (mine + other);
        }
#pragma unroll
      for (int b = 0; b < 8; ++b)
        if (b == a) {
#pragma unroll
          for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int k2 = 0; k2 < 3; ++k2) kuu[k][3 * b + k2] = dgb[k][k2];
        }
    }
    // Sweep 2, round rr: every lane a > rr publishes its averaged block
    // (a, rr); lane rr copies the transposes into its blocks (rr, a).
#pragma unroll
    for (int rr = 0; rr < 7; ++rr) {
      if (a > rr) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int k2 = 0; k2 < 3; ++k2) T.region[a * 9 + k * 3 + k2] = kuu[k][3 * rr + k2];
      }
      team_sync();
      if (p == rr) {
#pragma unroll
        for (int c = rr + 1; c < 8; ++c)
#pragma unroll
          for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int k2 = 0; k2 < 3; ++k2) kuu[k2][3 * c + k] = T.region[c * 9 + k * 3 + k2];
      }
      team_sync();
    }
  }
  double* out = A.tangent + e * 576 + (3 * a) * 24;
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 24; ++j) out[k * 24 + j] = kuu[k][j];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
int launch_local_solve(const ElementArgs& a, hipStream_t stream) {
  (void)hipGetLastError();
  const unsigned int blocks = static_cast<unsigned int>((a.ne + kChunk - 1) / kChunk);
  if (blocks == 0) return 0;
  switch (a.kinematics) {
    case kSmall: hipLaunchKernelGGL(local_solve_kernel<kSmall>, dim3(blocks), dim3(64), 0, stream, a); break;
    case kFinite: hipLaunchKernelGGL(local_solve_kernel<kFinite>, dim3(blocks), dim3(64), 0, stream, a); break;
    default: hipLaunchKernelGGL(local_solve_kernel<kLog>, dim3(blocks), dim3(64), 0, stream, a); break;
  }
  return static_cast<int>(hipGetLastError());
}

int launch_condense(const ElementArgs& a, bool tangent, hipStream_t stream) {
  (void)hipGetLastError();
  const unsigned int blocks = static_cast<unsigned int>((a.ne + kWaveTeams - 1) / kWaveTeams);
  if (blocks == 0) return 0;
#define SBK_LAUNCH(K, T) hipLaunchKernelGGL((condense_kernel<K, T>), dim3(blocks), dim3(64), 0, stream, a)
  if (tangent) {
    switch (a.kinematics) {
      case kSmall: SBK_LAUNCH(kSmall, true); break;
      case kFinite: SBK_LAUNCH(kFinite, true); break;
      default: SBK_LAUNCH(kLog, true); break;
    }
  } else {
    switch (a.kinematics) {
      case kSmall: SBK_LAUNCH(kSmall, false); break;
      case kFinite: SBK_LAUNCH(kFinite, false); break;
      default: SBK_LAUNCH(kLog, false); break;
    }
  }
#undef SBK_LAUNCH
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
