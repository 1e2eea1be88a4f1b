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
//                       (kept in LDS) that requests one pass at a time; each
//                       wavefront owns a chunk of elements and a team that
//                       finishes takes the next one, so the divergent
//                       iteration counts do not idle the wavefront. Every pass
//                       writes the points' trial states; a pass that meets the
//                       stopping test also writes the points' stress S and
//                       tangent (the point workspace), energy and yielding.
//                       The last such pass of a converged element is its
//                       final pass. Writes alpha and a status per element.
//   condense_*_kernel   reads S and the tangent of the final pass back and
//                       K_aa's factors of the final local iteration, forms the
//                       coupling blocks K_ua, K_au, K_uu and the static
//                       condensation f* = f_u - K_ua K_aa^-1 r_a,
//                       K* = K_uu - K_ua K_aa^-1 K_au (team kernel: rules of
//                       up to 8 points; wide kernel: 9..32 points).
#include <algorithm>

#include "element_kernels.hpp"
#include "element_pass.hpp"
#include "team_linalg.hpp"
#include "workspace_records.hpp"

namespace sbk {

constexpr int kWaveTeams = 8;      // teams of 8 lanes per wavefront
constexpr int kDefaultChunk = 64;  // elements per wavefront in the local solve
// Lanes per element: 8 (a lane per point) for rules of up to 8 points, 32 for
// rules of 9..32 points (the reference's 2 x 2 x m box rules). In a 32-lane
// team lane p still owns point p; the element's 9 x 9 algebra (the local
// Newton's machine, K_aa's factorisation) runs on lanes 0..7 as in an 8-lane
// team, and the condensation's row blocks on lanes 0..7.
template <int TW>
constexpr int teams_per_wave() { return 64 / TW; }

// IncompatibleModes.hpp:98-127
constexpr int kMaxLocalIterations = 25;
constexpr double kFullStepReduction = 1.0 - 1.0e-4;
constexpr int kMaxLineSearch = 8;
constexpr double kLineSearchRatio = 0.5;
constexpr double kLocalShift = 1.0e-3;
constexpr double kLocalMarquardt = 1.0e-6;
constexpr double kLocalTolerance = 1.0e-8;
constexpr double kLocalRoundOff = 1.0e-13;

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

// ---------------------------------------------------------------------------
// One pass of a team: the local sums r_alpha, scale, floor and K_aa
// (IncompatibleModes.cpp:252-275), reduced over the team.
// ---------------------------------------------------------------------------
struct LocalSums {
  double ra[9];
  double scale_max, floor_max;  // max over the 9 totals
  bool within_tol;              // max |r_a| <= tol scale + round-off floor (:398, :469)
  bool stop;                    // within_tol or r_a = 0: the stopping test (:469)
};


// B_a (6 x 9) of a point: column n is mode n / 3, direction n % 3.
__device__ __forceinline__ void mode_operator(const double F[3][3], const double gm[3][3], double ba[6][9]) {
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int n = 0; n < 9; ++n) ba[c][n] = bnl_mode(F, gm, c, n / 3, n % 3);
}

// The team's sums of a pass from each lane's point response (w = 0 and zero
// response on lanes without a point); the K_aa totals go to kaa (row-major).
template <int KIN, int TW>
__device__ __forceinline__ void team_sums(const double F[3][3], const double h[3][3], const PointResponse& r,
                                          const double gm[3][3], double w, int p, double* kaa, LocalSums& out,
                                          bool want_kaa_only) {
  double ba[6][9];
  mode_operator(F, gm, ba);
  if (!want_kaa_only) {
    double hmax = 0.0, mmax = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) hmax = fmax(hmax, fabs(h[i][j]));
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
      out.ra[n] = team_sum<TW>(w * sr);
      smax = fmax(smax, team_sum<TW>(w * ss));
      fmaxv = fmax(fmaxv, team_sum<TW>(ff * sf));
    }
    out.scale_max = smax;
    out.floor_max = fmaxv;
    const double largest = maxabs9(out.ra);
    out.within_tol = largest <= kLocalTolerance * smax + kLocalRoundOff * fmaxv;
    out.stop = out.within_tol || largest == 0.0;
  } else {
#pragma unroll
    for (int n = 0; n < 9; ++n) {
      double sr = ba[0][n] * r.S[0];
#pragma unroll
      for (int c = 1; c < 6; ++c) sr = sr + ba[c][n] * r.S[c];
      out.ra[n] = team_sum<TW>(w * sr);
    }
  }
  // K_aa = sum_q w B_a^T mat B_a (+ the geometric alpha-alpha block), a row
  // at a time: t = B_a[:, mm]^T mat, then K_aa[mm][n] = t B_a[:, n]
  double geo[3][3];
  if constexpr (KIN != kSmall) {
    // gt^T S gt with gt = G~ (3 x 3)
    const double S3[3][3] = {{r.S[0], r.S[3], r.S[5]}, {r.S[3], r.S[1], r.S[4]}, {r.S[5], r.S[4], r.S[2]}};
    double gs[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) gs[a][j] = gm[0][a] * S3[0][j] + gm[1][a] * S3[1][j] + gm[2][a] * S3[2][j];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) geo[a][b] = gs[a][0] * gm[0][b] + gs[a][1] * gm[1][b] + gs[a][2] * gm[2][b];
  }
#pragma unroll
  for (int mm = 0; mm < 9; ++mm) {
    double t[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      double s = ba[0][mm] * r.mat[0][d];
#pragma unroll
      for (int c = 1; c < 6; ++c) s = s + ba[c][mm] * r.mat[c][d];
      t[d] = s;
    }
#pragma unroll
    for (int n = 0; n < 9; ++n) {
      double s = t[0] * ba[0][n];
#pragma unroll
      for (int d = 1; d < 6; ++d) s = s + t[d] * ba[d][n];
      double v = w * s;
      if constexpr (KIN != kSmall) {
        if (mm % 3 == n % 3) v = v + w * geo[mm / 3][n / 3];
      }
      const double tot = team_sum<TW>(v);
      if (((mm * 9 + n) % TW) == p) kaa[mm * 9 + n] = tot;
    }
  }
  team_sync();
}

__device__ __forceinline__ void zero_response(PointResponse& r) {
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    r.S[c] = 0.0;
#pragma unroll
    for (int d = 0; d < 6; ++d) r.mat[c][d] = 0.0;
  }
  r.energy = 0.0;
  r.yielding = false;
}

// One pass at alpha; returns 0 or the team's first failure (a kinematic one
// before any return failure, as the reference evaluates every point's
// kinematics first). Writes the trial states of a successful pass, and the
// workspace, energy and yielding of a pass that meets the stopping test.
// trial(alpha) fills the pass's mode parameters (from the team's LDS machine);
// it is called again after the return rather than holding them across it.
template <int KIN, int NT, int TW, class Trial>
__device__ __forceinline__ int team_local_pass(const ElementArgs& A, const DevMaterial& m, long long e, int p,
                                               LaneLds geom, LaneLds logs, Trial trial, double* kaa,
                                               LocalSums& out) {
  const bool has_point = p < A.rule.n;
  const long long q = e * A.rule.n + p;
  PointResponse r;
  int fail = 0;
  if (has_point) {
    PointStateT<NT> st;
    load_state(m, A.state_in, A.state_stride, q, st);
    double alpha[9];
    trial(alpha);
    fail = point_response<KIN>(m, geom, alpha, st, r, A.state_out, A.state_stride, q, logs);
  } else {
    zero_response(r);
  }
  reload_barrier();
  const int stage = (fail == kReturnFailed || fail == kNonFiniteTrial) ? 1 : 0;
  const int first = team_imin<TW>(fail != 0 ? (stage << 16) | (p << 8) | fail : 1 << 20);
  if (first != (1 << 20)) return first & 0xFF;
  // F and h again from the LDS data and the machine's alpha, rather than held
  // in registers across the return
  double F[3][3], h[3][3], gm[3][3];
  double w = 0.0;
  if (has_point) {
    double alpha[9], hu[3][3];
    trial(alpha);
    get_hu(geom, hu);
    get_gm(geom, gm);
    mode_kinematics<KIN>(hu, gm, alpha, h, F);
    w = geom.get(18);
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) { F[i][j] = 0.0; h[i][j] = 0.0; gm[i][j] = 0.0; }
  }
  const double energy = team_sum<TW>(w * r.energy);
  const int yielding = team_isum<TW>(r.yielding ? 1 : 0);
  team_sums<KIN, TW>(F, h, r, gm, w, p, kaa, out, false);
  if (out.stop) {
    if (has_point) {
      double* rec = A.point_ws + q * kPointRecord;
#pragma unroll
      for (int c = 0; c < 6; ++c) rec[kWsStress + c] = r.S[c];
#pragma unroll
      for (int i = 0; i < 36; ++i) rec[kWsTangent + i] = r.mat[i / 6][i % 6];
    }
    if (p == 0) {
      A.energy[e] = energy;
      A.yielding[e] = yielding;
    }
  }
  return 0;
}

// The outputs of a failed element: no energy, no yielding points, the
// committed modes and history (the reference rejects the trial after a
// SolverError); the condensation zeroes its force and tangent. Every output
// of a failed element is thereby fixed, whatever its team did before.
__device__ __forceinline__ void failed_element_outputs(const ElementArgs& A, long long e, int p) {
  if (p < kTeam) A.alpha_out[e * 9 + p] = A.alpha_in ? A.alpha_in[e * 9 + p] : 0.0;
  if (p == 0) {
    A.alpha_out[e * 9 + 8] = A.alpha_in ? A.alpha_in[e * 9 + 8] : 0.0;
    A.energy[e] = 0.0;
    A.yielding[e] = 0;
  }
  if (p < A.rule.n) {
    const long long q = e * A.rule.n + p;
    const int fields = state_fields(A.mat.kin_fields);
    for (int f = 0; f < fields; ++f) A.state_out[f * A.state_stride + q] = A.state_in[f * A.state_stride + q];
  }
}

// Chunk c of a balanced partition of ne positions into w chunks starts at
// chunk_start(ne, w, c): the first ne % w chunks hold one position more.
__device__ __forceinline__ long long chunk_start(long long ne, long long w, long long c) {
  const long long q = ne / w, r = ne % w;
  return c * q + (c < r ? c : r);
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

// Bits of an element's local path (the phases of the local iteration it went
// through; SPARLAB_LOCAL_* in the C ABI).
constexpr int kPathRegulaFalsi = 1;
constexpr int kPathNormSearch = 2;
constexpr int kPathIndefinite = 4;
constexpr int kPathShifted = 8;
constexpr int kPathMarquardt = 16;
constexpr int kPathHalved = 32;

struct Machine {
  int phase, ls_ctx, sn_ctx, iterations, halving, search, side, path;
  bool last_is_best, moved;
  double alpha[9], d[9], newton[9], r0[9];
  double residual, s;  // s: the trial is alpha + s d
  double phi0, reach, s0, f0, s1, f1, step;
  double best_step, best, trial_step, mu;
};

// A team's LDS: K_aa and its factors; K_aa of the iteration (the fallbacks
// need it); the machine. The fallbacks' work matrix is lu.a: the factors are
// dead once Newton's direction is solved, and a pass rewrites lu.a before the
// next factorisation.
struct TeamSlot {
  LuSlot lu;
  double k0[81];
  Machine m;
  int step;  // the machine's step result, for the lanes of a 32-lane team beyond 8
  __device__ double* w() { return lu.a; }
};

// The outcome of a step of the machine: keep going (a new trial is
// requested), converged, or failed with a status.
enum StepResult { kContinue = 0, kConverged = 1, kFailed = 2 };

// The top of the local loop (IncompatibleModes.cpp:457-515): convergence
// test, K_aa's LU, Newton's direction and the choice of the search.
__device__ int machine_top(Machine& M, TeamSlot& T, const LocalSums& cur, int p, int& status, double& value) {
  const double largest = maxabs9(cur.ra);
  // K_aa is in T.lu.a (the pass wrote it); keep a copy for the fallbacks.
  for (int i = p; i < 81; i += 8) T.k0[i] = T.lu.a[i];
  team_sync();
  const bool inv = coop_full_lu(T.lu, p);
  if (!inv) { status = kSingularModes; value = largest; return kFailed; }
  if (cur.stop) return kConverged;
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
  const bool spd = spd9(T.k0);
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
    M.path |= kPathIndefinite;
  }
  return kContinue;
}

__device__ __forceinline__ void sn_start(Machine& M, double reach) {
  M.path |= kPathNormSearch;
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
    T.w()[i] = s;
  }
  team_sync();
  double dmax = T.w()[0];
  for (int i = 1; i < 9; ++i) dmax = fmax(dmax, T.w()[i * 10]);
  M.mu = kLocalMarquardt * dmax;
  M.moved = false;
  M.path |= kPathMarquardt;
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
      T.w()[i] = (r == c) ? s + M.mu : s;
    }
    llt9(T.w());
    double x[9];
    llt9_solve(T.w(), grad, x);
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
          M.path |= kPathHalved;
          M.reach = M.reach * 0.5;
          ++M.halving;
          M.s = M.reach;
          return kContinue;
        }
        next = kLsFalse;
        break;
      } else {
        const double full = norm9(cur.ra);
        if (full <= kFullStepReduction * M.residual || cur.within_tol) {
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
        M.path |= kPathRegulaFalsi;
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
      M.path |= kPathShifted;
      for (int i = 0; i < 81; ++i) {
        const int r = i / 9, c = i % 9;
        T.w()[i] = 0.5 * (T.k0[r * 9 + c] + T.k0[c * 9 + r]);
      }
      double ev[9];
      jacobi9_values(T.w(), ev);
      double top = fabs(ev[0]);
      for (int i = 1; i < 9; ++i) top = fmax(top, fabs(ev[i]));
      const double mu = kLocalShift * top - ev[0];
      team_sync();
      for (int i = 0; i < 81; ++i) T.w()[i] = (i % 10 == 0) ? T.k0[i] + mu : T.k0[i];
      double x[9];
      plu9_solve(T.w(), M.r0, x);
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

template <int KIN, int NT, int TW>
__global__ __launch_bounds__(64) void local_solve_kernel(ElementArgs A) {
  __shared__ TeamSlot slots[teams_per_wave<TW>()];
  __shared__ double geom_lds[mode_geom_fields<KIN>()][64];
  __shared__ double log_lds[KIN == kLog ? kLogStash : 1][64];
  const int lane = threadIdx.x;
  const LaneLds geom{&geom_lds[0][lane]};
  const LaneLds logs{&log_lds[0][lane]};
  const int t = lane / TW;
  const int p = lane % TW;
  TeamSlot& T = slots[t];
  Machine& M = T.m;
  // the wavefront's positions: a balanced partition of [0, ne) into gridDim.x
  // chunks of at most A.chunk elements (chunk_start)
  const long long c0 = chunk_start(A.ne, gridDim.x, blockIdx.x);
  const long long c1 = chunk_start(A.ne, gridDim.x, blockIdx.x + 1);
  long long next = c0;
  long long e = -1;
  if (p == 0) M.phase = kIdle;
  team_sync();
  for (;;) {
    // Hand the chunk's next elements to the idle teams, in team order.
    const bool idle = M.phase == kIdle;
    const unsigned long long leaders = __ballot(idle && p == 0);
    const int before = __popcll(leaders & ((1ull << (t * TW)) - 1ull));
    const int free_teams = __popcll(leaders);
    bool fresh = false;
    if (idle) {
      const long long cand = next + before;
      if (cand < c1) {
        e = A.ordered ? A.order[cand] : cand;
        fresh = true;
      } else {
        e = -1;
      }
    }
    next += free_teams;
    if (__ballot(!idle || fresh) == 0ull) break;
    if (fresh) {
      double x0[8][3], ue[24];
      load_element(A.coords, A.conn, A.u, e, x0, ue);
      double fv = 0.0;
      int st = kOk;
      PointGeom pg;
      if (p < A.rule.n) {
        st = setup_point(x0, ue, A.rule.xi[p], A.rule.eta[p], A.rule.zeta[p], A.rule.w[p], pg, fv);
        ModeGeom mg;
        mode_geom<KIN>(pg, ue, mg);
        put_mode_geom<KIN>(mg, geom);
      }
      // the first failing point in point order, as the reference's set-up loop
      const int code = team_imin<TW>(st != kOk ? (p << 8) | st : 1 << 20);
      if (code != (1 << 20)) {
        if (p == (code >> 8)) { A.status[e] = code & 0xFF; A.fail_value[e] = fv; }
        if (p == 0 && A.local_path) A.local_path[e] = 0;
        failed_element_outputs(A, e, p);
        continue;  // the team stays idle and takes another element
      }
      if (p == 0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) M.alpha[i] = A.alpha_in ? A.alpha_in[e * 9 + i] : 0.0;
        M.iterations = 0;
        M.path = 0;
        M.s = 0.0;
        M.phase = kFirst;
      }
      team_sync();
    }
    if (!idle || fresh) {
      auto trial = [&M](double a[9]) {
        if (M.phase == kFirst) {
#pragma unroll
          for (int i = 0; i < 9; ++i) a[i] = M.alpha[i];
        } else {
#pragma unroll
          for (int i = 0; i < 9; ++i) a[i] = M.alpha[i] + M.s * M.d[i];
        }
      };
      LocalSums cur;
      const int fail = team_local_pass<KIN, NT, TW>(A, A.mat, e, p, geom, logs, trial, T.lu.a, cur);
      const bool ok = fail == 0;
      int status = ok ? kOk : fail;
      double value = 0.0;
      int res;
      if constexpr (TW == kTeam) {
        res = machine_step(M, T, ok, cur, p, status, value);
      } else {
        // the machine on lanes 0..7; its result to the rest of the team
        if (p < kTeam) {
          res = machine_step(M, T, ok, cur, p, status, value);
          if (p == 0) T.step = res;
        }
        team_sync();
        res = T.step;
      }
      if (res == kConverged) {
        if (p == 0) {
          A.status[e] = kOk;
          A.iterations[e] = M.iterations;
          if (A.local_path) A.local_path[e] = M.path;
        }
        if (p < 9) A.alpha_out[e * 9 + p] = M.alpha[p];
        if (p == 0) A.alpha_out[e * 9 + 8] = M.alpha[8];
        // K_aa's factors and r_a of the final iteration, for the condensation
        double* er = A.elem_ws + e * kElemRecord;
        for (int i = p; i < 81; i += TW) er[kErFactors + i] = T.lu.a[i];
        if (p < kTeam) {
          er[kErResidual + p] = cur.ra[p];
          er[kErRows + p] = static_cast<double>(T.lu.rt[p]);
          er[kErCols + p] = static_cast<double>(T.lu.ct[p]);
        }
        if (p == 0) {
          er[kErResidual + 8] = cur.ra[8];
          er[kErRows + 8] = static_cast<double>(T.lu.rt[8]);
          er[kErCols + 8] = static_cast<double>(T.lu.ct[8]);
        }
        team_sync();
        if (p == 0) M.phase = kIdle;
      } else if (res == kFailed) {
        if (p == 0) {
          A.status[e] = status == kOk ? kUnsupported : status;
          A.fail_value[e] = value;
          A.iterations[e] = M.iterations;
          if (A.local_path) A.local_path[e] = M.path;
        }
        failed_element_outputs(A, e, p);
        team_sync();
        if (p == 0) M.phase = kIdle;
      }
      team_sync();
    }
  }
}

// ---------------------------------------------------------------------------
// The schedule of the local solve
// ---------------------------------------------------------------------------
// A wavefront's pass costs what its most expensive point costs, and most of a
// plastic point's cost is its return; the local solve therefore has each
// wavefront take its elements in descending expected work, so that its 8 teams
// work on elements of like cost at the same time. A launch of many wavefronts
// per SIMD takes the elements in one descending run (each chunk then holds
// elements of one cost, and the cheap chunks fill the end of the launch); a
// launch of few wavefronts, which all run at once, deals them to the chunks
// round-robin in descending order (rank r to chunk r mod w, at slot r / w), so
// that every chunk holds the same mix of work and no wavefront outlasts the
// others. The expected work of an element is the sum over its points of
// return_cost at the committed modes (the elastic trial of the first pass,
// with the Green-Lagrange strain standing in for the logarithmic one).
// The order changes no result: every element's evaluation is independent of
// the others.

// The strain of a point at alpha, for the schedule only: the head of
// point_response, except that logarithmic kinematics take the Green-Lagrange
// strain (E_log's first-order approximation) instead of its eigen-decomposition.
// False for an inverted point, which the local solve itself reports.
template <int KIN>
__device__ __forceinline__ bool trial_strain(const ModeGeom& mg, const double alpha[9], double strain[6]) {
  double h[3][3], F[3][3];
  mode_kinematics<KIN>(mg.hu, mg.gm, alpha, h, F);
  if constexpr (KIN == kSmall) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double s2 = bnl_mode(F, mg.gm, c, 0, 0) * alpha[0];
#pragma unroll
      for (int kk = 1; kk < 9; ++kk) s2 = s2 + bnl_mode(F, mg.gm, c, kk / 3, kk % 3) * alpha[kk];
      strain[c] = mg.eps_u[c] + s2;
    }
    return true;
  } else {
    if constexpr (KIN == kLog) {
      if (!(det3(F) > 0.0)) return false;
    }
    double e3[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double hth = h[0][i] * h[0][j] + h[1][i] * h[1][j] + h[2][i] * h[2][j];
        e3[i][j] = 0.5 * ((h[i][j] + h[j][i]) + hth);
      }
    strain[0] = e3[0][0];
    strain[1] = e3[1][1];
    strain[2] = e3[2][2];
    strain[3] = 2.0 * e3[0][1];
    strain[4] = 2.0 * e3[1][2];
    strain[5] = 2.0 * e3[2][0];
    return true;
  }
}

// One team per element, lane p = point p: cost[e] = the element's expected
// work (0 .. 2 points). Block 0 also clears the class counts and cursors.
template <int KIN, int NT, int TW>
__global__ __launch_bounds__(64) void cost_key_kernel(ElementArgs A) {
  const int lane = threadIdx.x;
  const int t = lane / TW;
  const int p = lane % TW;
  if (blockIdx.x == 0)
    for (int i = lane; i < 2 * kCostClasses; i += 64) A.counts[i] = 0;
  const long long e = static_cast<long long>(blockIdx.x) * teams_per_wave<TW>() + t;
  int c = 0;
  if (e < A.ne && p < A.rule.n) {
    double x0[8][3], ue[24];
    load_element(A.coords, A.conn, A.u, e, x0, ue);
    PointGeom pg;
    double fv = 0.0;
    if (setup_point(x0, ue, A.rule.xi[p], A.rule.eta[p], A.rule.zeta[p], A.rule.w[p], pg, fv) == kOk) {
      ModeGeom mg;
      mode_geom<KIN>(pg, ue, mg);
      double alpha[9], strain[6];
#pragma unroll
      for (int i = 0; i < 9; ++i) alpha[i] = A.alpha_in ? A.alpha_in[e * 9 + i] : 0.0;
      if (trial_strain<KIN>(mg, alpha, strain)) {
        PointStateT<NT> st;
        load_state(A.mat, A.state_in, A.state_stride, e * A.rule.n + p, st);
        c = return_cost(A.mat, strain, st);
      }
    }
  }
  const int key = team_isum<TW>(c);
  if (e < A.ne && p == 0) A.cost[e] = key;
}

// Elements per block of the counting sort: each block owns a contiguous range.
__device__ __forceinline__ void block_range(long long ne, long long& b0, long long& b1) {
  const long long per = (ne + gridDim.x - 1) / gridDim.x;
  b0 = static_cast<long long>(blockIdx.x) * per;
  b1 = b0 + per < ne ? b0 + per : ne;
}

__global__ __launch_bounds__(256) void cost_count_kernel(const int* __restrict__ cost, long long ne, int* counts) {
  __shared__ int h[kCostClasses];
  if (threadIdx.x < kCostClasses) h[threadIdx.x] = 0;
  __syncthreads();
  long long b0, b1;
  block_range(ne, b0, b1);
  for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) atomicAdd(&h[cost[i]], 1);
  __syncthreads();
  if (threadIdx.x < kCostClasses && h[threadIdx.x] != 0) atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

// Cursors: the classes in descending cost, each starting where the costlier
// ones end.
__global__ void cost_scan_kernel(int* counts) {
  if (threadIdx.x == 0) {
    int run = 0;
    for (int c = kCostClasses - 1; c >= 0; --c) {
      counts[kCostClasses + c] = run;
      run += counts[c];
    }
  }
}

// order[r] (deal = 0) or order[chunk_start(ne, chunks, r % chunks) + r / chunks]
// (deal = 1) = the element of rank r in descending cost (ties in no
// particular order).
__global__ __launch_bounds__(256) void cost_scatter_kernel(const int* __restrict__ cost, long long ne, int* counts,
                                                           long long chunks, int deal, int* __restrict__ order) {
  __shared__ int h[kCostClasses], base[kCostClasses];
  if (threadIdx.x < kCostClasses) h[threadIdx.x] = 0;
  __syncthreads();
  long long b0, b1;
  block_range(ne, b0, b1);
  for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) atomicAdd(&h[cost[i]], 1);
  __syncthreads();
  if (threadIdx.x < kCostClasses) {
    base[threadIdx.x] = h[threadIdx.x] != 0 ? atomicAdd(&counts[kCostClasses + threadIdx.x], h[threadIdx.x]) : 0;
    h[threadIdx.x] = 0;
  }
  __syncthreads();
  for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const int c = cost[i];
    const long long r = base[c] + atomicAdd(&h[c], 1);
    order[deal ? chunk_start(ne, chunks, r % chunks) + r / chunks : r] = static_cast<int>(i);
  }
}

// ---------------------------------------------------------------------------
// The condensation (IncompatibleModes.cpp:279-306, 549-561)
// ---------------------------------------------------------------------------
// Two kernels, by the rule. Both write K* row-major or column-major
// (ElementArgs::tangent_column_major).
constexpr int kRhs = 25;

// One point's records as the teams of a wavefront stage them in LDS: the
// geometry record (44 doubles, 22 chunks of 16 bytes) then the point record
// (42, 21 chunks), copied straight from global memory to LDS by 16-byte
// global_load_lds (no registers, no LDS stores): in copy j, lane a of team t
// copies its team's chunk 8 j + a into byte 16 (8 t + a) of the stage's row
// j (64 lanes x 16 bytes). Six copies per point; lanes past chunk 42 copy it
// again into their own unused slots, so every lane issues the same six. Both
// records start 16-byte aligned (the workspace is, and each record holds a
// whole number of chunks).
constexpr int kStagePoint = kGeomRecord + kPointRecord;
constexpr int kStageChunks = kStagePoint / 2;   // 43
constexpr int kStageCopies = (kStageChunks + kTeam - 1) / kTeam;  // 6
constexpr int kStageRow = 2 * 64;               // doubles in one copy's row
constexpr int kStageDoubles = kStageCopies * kStageRow;
constexpr int kStageDepth = 3;                  // points q, q + 1, q + 2
static_assert(kGeomRecord % 2 == 0 && kPointRecord % 2 == 0, "records of whole 16-byte chunks");

// Entry k of team t's staged records (k < kGeomRecord: the geometry record;
// OFF = kGeomRecord reads the point record): chunk k / 2 in row k / 16, at
// lane 8 t + (k % 16) / 2. `col` points at the team's column, stage + 16 t.
template <int OFF>
struct StagedRecord {
  const double* col;
  __device__ __forceinline__ double operator[](int k) const {
    const int i = k + OFF;
    return col[(i >> 4) * kStageRow + (i & 15)];
  }
};

typedef __attribute__((address_space(1))) void GlobalVoid;
typedef __attribute__((address_space(3))) void LdsVoid;

// Lane a of its team: the six copies of point q's records into `stage`
// (the wavefront's stage, uniform across the lanes).
__device__ __forceinline__ void stage_issue(const ElementArgs& A, long long q, int a, double* stage) {
  const double* gr = A.geom_ws + q * kGeomRecord;
  const double* pr = A.point_ws + q * kPointRecord;
#pragma unroll
  for (int j = 0; j < kStageCopies; ++j) {
    int c = kTeam * j + a;
    if (c >= kStageChunks) c = kStageChunks - 1;
    const double* src = c < kGeomRecord / 2 ? gr + 2 * c : pr + 2 * (c - kGeomRecord / 2);
    // address-space casts (C style: the generic pointers name global and LDS memory)
    __builtin_amdgcn_global_load_lds((GlobalVoid*)(src), (LdsVoid*)(stage + j * kStageRow), 16, 0, 0);
  }
}

// Waits until at most `later` points' copies (issued after the one needed
// now) are in flight: vmcnt <= 6 later (s_waitcnt: expcnt and lgkmcnt not
// waited on). Stores count in vmcnt too: a point loop starts from vmcnt 0
// and issues no other vector memory instruction.
__device__ __forceinline__ void stage_wait(int later) {
  static_assert(kStageCopies == 6, "the vmcnt thresholds below assume six copies per point");
  if (later >= 2) __builtin_amdgcn_s_waitcnt(0x0F7C);       // vmcnt(12)
  else if (later == 1) __builtin_amdgcn_s_waitcnt(0x0F76);  // vmcnt(6)
  else __builtin_amdgcn_s_waitcnt(0x0F70);                  // vmcnt(0)
}

// Waits for the wavefront's LDS accesses (lgkmcnt(0); vmcnt and expcnt not
// waited on): before a copy overwrites a stage the team has read.
__device__ __forceinline__ void stage_reads_done() { __builtin_amdgcn_s_waitcnt(0xC07F); }

struct CondenseSlot {
  LuSlot lu;
  LuPerm perm;
  double region[9 * kRhs];
};

// One team (8 lanes) per element. Lane p first forms the geometry of point p
// at the converged alpha (F with the pass's own arithmetic) into the
// workspace; then lane a owns rows 3a..3a+2 of f*, K_ua, K_uu and K* (and
// columns 3a..3a+2 of K_au), each accumulated point by point in point order
// from the geometry records and the point records of the final pass. The
// accumulations run as three point loops (f_u, K_ua and K_au together; K_uu
// in two column halves after the solve): each loop re-reads the records, and
// fewer loops would hold more accumulators than the registers take.
template <int KIN, bool TANGENT, bool SYM, int TW>
__global__ __launch_bounds__(64) void condense_team_kernel(ElementArgs A) {
  static_assert(TW == kTeam, "the team condensation takes rules of up to 8 points");
  __shared__ CondenseSlot slots[teams_per_wave<TW>()];
  // three stages as three LDS objects, each named statically where it is
  // copied to or read: the compiler's wait before an LDS read then covers
  // the copies into that stage only, not every copy in flight
  __shared__ __attribute__((aligned(16))) double stage0[kStageDoubles];
  __shared__ __attribute__((aligned(16))) double stage1[kStageDoubles];
  __shared__ __attribute__((aligned(16))) double stage2[kStageDoubles];
  const int lane = threadIdx.x;
  const int t = lane / TW;
  const int p = lane % TW;
  CondenseSlot& T = slots[t];
  const long long e = static_cast<long long>(blockIdx.x) * teams_per_wave<TW>() + t;
  if (e >= A.ne) return;
  if (A.status[e] != kOk) {
    // a failed element: zero force and tangent (lane a: rows 3a..3a+2)
    if (p >= kTeam) return;
    for (int k = 0; k < 3; ++k) A.force[e * 24 + 3 * p + k] = 0.0;
    if constexpr (TANGENT)
      for (int i = 0; i < 72; ++i) A.tangent[e * 576 + 72 * p + i] = 0.0;
    return;
  }
  const int n = A.rule.n;
  const long long q0 = e * n;
  // geometry of the lane's point
  if (p < n) {
    double x0[8][3], ue[24];
    load_element(A.coords, A.conn, A.u, e, x0, ue);
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
  // K_aa's factors of the final local iteration
  double* er = A.elem_ws + e * kElemRecord;
  for (int i = p; i < 81; i += TW) T.lu.a[i] = er[kErFactors + i];
  if (p < kTeam) {
    T.lu.rt[p] = static_cast<int>(er[kErRows + p]);
    T.lu.ct[p] = static_cast<int>(er[kErCols + p]);
  }
  if (p == 0) {
    T.lu.rt[8] = static_cast<int>(er[kErRows + 8]);
    T.lu.ct[8] = static_cast<int>(er[kErCols + 8]);
  }
  // r_a: column 24 of the right-hand sides
  if (p < kTeam) T.region[p * kRhs + 24] = er[kErResidual + p];
  if (p == 0) T.region[8 * kRhs + 24] = er[kErResidual + 8];
  team_sync_global();  // the geometry records (global) and the factors (LDS)
  if (p == 0) lu_perm(T.lu.rt, T.lu.ct, T.perm);
  // the row blocks of the 8 nodes: lane a of the team
  const int a = p;
  // Every point loop reads the records from the wavefront's LDS stages:
  // while the team works on point q, the copies of the next one or two
  // points are in flight (each lane's own; the team's lanes wait for their
  // copies, then a fence shows them to the whole team).
  auto point_loop = [&](auto&& body) {
    stage_wait(0);  // the stores before the loop
    stage_issue(A, q0, a, stage0);
    if (n > 1) stage_issue(A, q0 + 1, a, stage1);
    // Point p is copied into stage p % 3, two points ahead, over stage
    // (p - 3) % 3, which every lane has read by then. The compiler waits for
    // every copy in flight before the first read of stage 0 in a pass of
    // three points (it cannot tell the stages' copies apart there), so the
    // copy of point q + 2 is issued after point q's reads in that step: the
    // wait then covers point q + 1's copy, issued a whole point earlier.
    // `later`: copies issued after point q's when q is read.
    auto step = [&](int q, const double* cur, double* ahead, bool issue_first) {
      if (issue_first && q + 2 < n) {
        stage_reads_done();
        stage_issue(A, q0 + q + 2, a, ahead);
      }
      const int hi = issue_first ? q + 3 : q + 2;  // points issued so far
      stage_wait((hi < n ? hi : n) - 1 - q);
      team_sync();
      const double* col = cur + 16 * t;
      body(StagedRecord<0>{col}, StagedRecord<kGeomRecord>{col});
      team_sync();
      if (!issue_first && q + 2 < n) {
        stage_reads_done();
        stage_issue(A, q0 + q + 2, a, ahead);
      }
    };
    for (int q = 0; q < n; q += kStageDepth) {
      step(q, stage0, stage2, false);
      if (q + 1 < n) step(q + 1, stage1, stage0, true);
      if (q + 2 < n) step(q + 2, stage2, stage1, true);
    }
  };
  double ru[3], kua[3][9];
  if constexpr (!TANGENT || SYM) {
    // f_u rows and K_ua of node a
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ru[k] = 0.0;
#pragma unroll
      for (int j = 0; j < 9; ++j) kua[k][j] = 0.0;
    }
    point_loop([&](const auto& gr, const auto& pr) { coupling_point<KIN>(gr, pr, a, ru, kua); });
  }
  if constexpr (TANGENT) {
    // the coupling columns 3a..3a+2: K_ua^T, or K_au with
    // K_au[:, 3a + k] = sum_q w B_a^T mat B_u[:, 3a + k] (+ geometric),
    // accumulated in the same point loop as f_u and K_ua
    double col[3][9];
    if constexpr (SYM) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int nn = 0; nn < 9; ++nn) col[k][nn] = kua[k][nn];
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        ru[k] = 0.0;
#pragma unroll
        for (int nn = 0; nn < 9; ++nn) { kua[k][nn] = 0.0; col[k][nn] = 0.0; }
      }
      point_loop([&](const auto& gr, const auto& pr) {
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
        // z = B_u[:, 3a + k]^T mat, y = mat B_u[:, 3a + k]
        double z[3][6], y[3][6];
        node_times_tangent(bu, pr, z);
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double s = pr[kWsTangent + c * 6] * bu[0][k];
#pragma unroll
            for (int d = 1; d < 6; ++d) s = s + pr[kWsTangent + c * 6 + d] * bu[d][k];
            y[k][c] = s;
          }
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
            double t = bcol[0] * y[k][0];
#pragma unroll
            for (int c = 1; c < 6; ++c) t = t + bcol[c] * y[k][c];
            col[k][nn] = col[k][nn] + w * t;
          }
        }
        if constexpr (KIN != kSmall) {
          double gsa[3];
          stress_times_gradient(gr, pr, a, gsa);
#pragma unroll
          for (int bm = 0; bm < 3; ++bm) {
            const double geo = gsa[0] * gr[kGrG + 0 * 11 + 8 + bm] + gsa[1] * gr[kGrG + 1 * 11 + 8 + bm] +
                               gsa[2] * gr[kGrG + 2 * 11 + 8 + bm];
            const double v = w * geo;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              kua[k][3 * bm + k] = kua[k][3 * bm + k] + v;
              col[k][3 * bm + k] = col[k][3 * bm + k] + v;
            }
          }
        }
      });
    }
#pragma unroll
    for (int nn = 0; nn < 9; ++nn)
#pragma unroll
      for (int k = 0; k < 3; ++k) T.region[nn * kRhs + 3 * a + k] = col[k][nn];
  }
  team_sync();  // the permutations and the right-hand sides
  // K_aa^-1 [K_au | r_a] in place: lane a its three columns, lane 0 also r_a
  if (p == 0) lu_solve_column(T.lu.a, T.perm, T.region, kRhs, 24);
  if constexpr (TANGENT) lu_solve_column3(T.lu.a, T.perm, T.region, kRhs, 3 * a);
  team_sync();
  // f* = f_u - K_ua K_aa^-1 r_a
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double s = kua[k][0] * T.region[24];
#pragma unroll
    for (int nn = 1; nn < 9; ++nn) s = s + kua[k][nn] * T.region[nn * kRhs + 24];
    A.force[e * 24 + 3 * a + k] = ru[k] - s;
  }
  if constexpr (!TANGENT) return;
  // K*(r, c) at out[r * 24 + c] (row-major) or out[c * 24 + r]
  double* out = A.tangent + e * 576;
  const int rs = A.tangent_column_major ? 1 : 24, cs = A.tangent_column_major ? 24 : 1;
  // K_uu rows of node a in two column halves, K* = K_uu - K_ua coupling
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    double kuu[3][12];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 12; ++j) kuu[k][j] = 0.0;
    point_loop([&](const auto& gr, const auto& pr) {
      const double w = gr[kGrW];
      double bu[6][3];
      node_operator(gr, a, bu);
      double z[3][6];
      node_times_tangent(bu, pr, z);
#pragma unroll
      for (int bb = 0; bb < 4; ++bb)
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) {
          double bcol[6];
#pragma unroll
          for (int d = 0; d < 6; ++d) bcol[d] = bnl_rec(gr, d, 4 * half + bb, kb);
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double s = z[k][0] * bcol[0];
#pragma unroll
            for (int d = 1; d < 6; ++d) s = s + z[k][d] * bcol[d];
            kuu[k][3 * bb + kb] = kuu[k][3 * bb + kb] + w * s;
          }
        }
      if constexpr (KIN != kSmall) {
        double gsa[3];
        stress_times_gradient(gr, pr, a, gsa);
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
          const int b = 4 * half + bb;
          const double geo =
              gsa[0] * gr[kGrG + 0 * 11 + b] + gsa[1] * gr[kGrG + 1 * 11 + b] + gsa[2] * gr[kGrG + 2 * 11 + b];
          const double v = w * geo;
#pragma unroll
          for (int k = 0; k < 3; ++k) kuu[k][3 * bb + k] = kuu[k][3 * bb + k] + v;
        }
      }
    });
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int jj = 0; jj < 12; ++jj) {
        const int j = 12 * half + jj;
        double s = kua[k][0] * T.region[j];
#pragma unroll
        for (int nn = 1; nn < 9; ++nn) s = s + kua[k][nn] * T.region[nn * kRhs + j];
        out[(3 * a + k) * rs + j * cs] = kuu[k][jj] - s;
      }
  }
  if constexpr (SYM) {
    // K = (K + K^T) / 2 (IncompatibleModes.cpp:557) through the output rows
    team_sync_global();
    double ks[3][24];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 24; ++j) ks[k][j] = 0.5 * (out[(3 * a + k) * rs + j * cs] + out[j * rs + (3 * a + k) * cs]);
    team_sync_global();
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 24; ++j) out[(3 * a + k) * rs + j * cs] = ks[k][j];
  }
}


// The wide kernel (rules of 9..32 points, where the team kernel's lane per
// node would re-read many more point records): one element per workgroup (a
// single wavefront). Over the 11 augmented nodes
// (the 8 nodes, then the 3 mode groups; 3 components each) the element's
// matrix K[(a,k),(b,l)] = sum_q w_q (B_q^T mat_q B_q + delta_kl g_a^T S_q g_b)
// is accumulated in 3 x 3 blocks, point by point in point order, from
// operands staged in LDS kChunk points at a time: B_q (B_NL at the converged
// modes, 6 x 33), W_q = w_q mat_q B_q and w_q S_q g_b. Lane l = 8 a + b owns
// the block K_uu(a, b); lanes 0..23 also own K_ua(a, m) (a = l / 3, m = l % 3),
// lanes 24..47 K_au(m, b) (m = (l - 24) / 8, b = (l - 24) % 8; K_ua^T for a
// symmetric material), lanes 48..55 rows 3a..3a+2 of f_u. K_aa's factors of
// the final local iteration (the element record) then solve [K_au | r_a] in
// place, a column per lane, and each lane forms its block of
// K* = K_uu - K_ua K_aa^-1 K_au in LDS, from where the output is written in
// either layout (with K* = (K* + K*^T) / 2 for a symmetric material,
// IncompatibleModes.cpp:557). Each point record is read once.
constexpr int kChunk = 2;
constexpr int kAug = 33;  // 11 augmented nodes x 3 components
constexpr int kOps = kChunk * 6 * kAug;

// The geometry records (per point: F (9), [G | G~] (33), w) are the
// workgroup's dynamic LDS, rule.n of them (the launch's size), after these.
struct CondenseLds {
  double ops[2 * kOps];         // B_q, then W_q, of the chunk; afterwards K* (576) and K_ua (216)
  double rec[kChunk][kPointRecord];  // the chunk's point records: S, mat
  double sg[kChunk][kAug];      // w_q S_q g_b, b = 0..10
  LuSlot lu;
  LuPerm perm;
  double region[9 * kRhs];      // [K_au | r_a], then K_aa^-1 [K_au | r_a]
};

template <int KIN, bool TANGENT, bool SYM, int TW>
__global__ __launch_bounds__(64) void condense_wide_kernel(ElementArgs A) {
  __shared__ CondenseLds L;
  extern __shared__ __attribute__((aligned(16))) double condense_wide_geo[];
  double(*const geos)[kGeomRecord] = reinterpret_cast<double(*)[kGeomRecord]>(condense_wide_geo);
  const int l = threadIdx.x;
  const long long e = blockIdx.x;
  if (e >= A.ne) return;
  if (A.status[e] != kOk) {
    // a failed element: zero force and tangent
    if (l < 24) A.force[e * 24 + l] = 0.0;
    if constexpr (TANGENT)
      for (int i = l; i < 576; i += 64) A.tangent[e * 576 + i] = 0.0;
    return;
  }
  const int n = A.rule.n;
  const long long q0 = e * n;
  // the geometry of point l at the converged modes (F with the pass's own
  // arithmetic)
  if (l < n) {
    double x0[8][3], ue[24];
    load_element(A.coords, A.conn, A.u, e, x0, ue);
    PointGeom pg;
    double fv = 0.0;
    setup_point(x0, ue, A.rule.xi[l], A.rule.eta[l], A.rule.zeta[l], A.rule.w[l], pg, fv);
    double alpha[9], gm[3][3], h[3][3], F[3][3];
#pragma unroll
    for (int i = 0; i < 9; ++i) alpha[i] = A.alpha_out[e * 9 + i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) gm[i][j] = pg.g[i][8 + j];
    mode_kinematics<KIN>(pg.hu, gm, alpha, h, F);
    double* gr = geos[l];
#pragma unroll
    for (int i = 0; i < 9; ++i) gr[kGrF + i] = F[i / 3][i % 3];
#pragma unroll
    for (int i = 0; i < 33; ++i) gr[kGrG + i] = pg.g[i / 11][i % 11];
    gr[kGrW] = pg.w;
  }
  // K_aa's factors of the final local iteration and r_a (the 25th column)
  const double* er = A.elem_ws + e * kElemRecord;
  for (int i = l; i < 81; i += 64) L.lu.a[i] = er[kErFactors + i];
  if (l < 9) {
    L.lu.rt[l] = static_cast<int>(er[kErRows + l]);
    L.lu.ct[l] = static_cast<int>(er[kErCols + l]);
    L.region[l * kRhs + 24] = er[kErResidual + l];
  }
  // the lane's blocks: K_uu(a, b), and (r2, c2) in augmented nodes
  const int a = l >> 3, b = l & 7;
  int r2 = -1, c2 = 0;
  if (l < 24) {
    r2 = l / 3;
    c2 = 8 + l % 3;
  } else if (TANGENT && !SYM && l < 48) {
    r2 = 8 + (l - 24) / 8;
    c2 = (l - 24) % 8;
  }
  const int fa = l - 48;  // lanes 48..55: f_u rows of node fa
  const bool fu = fa >= 0 && fa < 8;
  double k1[3][3], k2[3][3], f3[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    f3[i] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) k1[i][j] = k2[i][j] = 0.0;
  }
  double* B = L.ops;
  double* W = L.ops + kOps;
  for (int c0 = 0; c0 < n; c0 += kChunk) {
    const int cn = n - c0 < kChunk ? n - c0 : kChunk;
    __syncthreads();  // the geometry is written; the previous chunk's operands are used
    for (int t = l; t < cn * kPointRecord; t += 64)
      L.rec[t / kPointRecord][t % kPointRecord] = A.point_ws[(q0 + c0) * kPointRecord + t];
    for (int t = l; t < cn * 6 * kAug; t += 64) {
      const int q = t / (6 * kAug), r = t % (6 * kAug);
      B[t] = bnl_rec(geos[c0 + q], r / kAug, (r % kAug) / 3, (r % kAug) % 3);
    }
    __syncthreads();
    // W_q = w_q mat_q B_q (mat row-major in the record), w_q S_q g_b
    for (int t = l; t < cn * 6 * kAug; t += 64) {
      const int q = t / (6 * kAug), r = t % (6 * kAug), c = r / kAug, j = r % kAug;
      const double* m = L.rec[q] + kWsTangent + c * 6;
      const double* bq = B + q * 6 * kAug + j;
      double s = m[0] * bq[0];
#pragma unroll
      for (int d = 1; d < 6; ++d) s = s + m[d] * bq[d * kAug];
      W[t] = geos[c0 + q][kGrW] * s;
    }
    if constexpr (KIN != kSmall) {
      for (int t = l; t < cn * 11; t += 64) {
        const int q = t / 11, node = t % 11;
        double gsa[3];
        stress_times_gradient(geos[c0 + q], L.rec[q], node, gsa);
        const double w = geos[c0 + q][kGrW];
#pragma unroll
        for (int j = 0; j < 3; ++j) L.sg[q][3 * node + j] = w * gsa[j];
      }
    }
    __syncthreads();
    for (int q = 0; q < cn; ++q) {
      const double* bq = B + q * 6 * kAug;
      const double* wq = W + q * 6 * kAug;
      const double* g = geos[c0 + q] + kGrG;
      if constexpr (TANGENT) {
        double s[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) s[i][j] = bq[3 * a + i] * wq[3 * b + j];
#pragma unroll
        for (int c = 1; c < 6; ++c)
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) s[i][j] = s[i][j] + bq[c * kAug + 3 * a + i] * wq[c * kAug + 3 * b + j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) k1[i][j] = k1[i][j] + s[i][j];
        if constexpr (KIN != kSmall) {
          const double geo = g[0 * 11 + a] * L.sg[q][3 * b] + g[1 * 11 + a] * L.sg[q][3 * b + 1] +
                             g[2 * 11 + a] * L.sg[q][3 * b + 2];
#pragma unroll
          for (int i = 0; i < 3; ++i) k1[i][i] = k1[i][i] + geo;
        }
      }
      if (r2 >= 0) {
        double s[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) s[i][j] = bq[3 * r2 + i] * wq[3 * c2 + j];
#pragma unroll
        for (int c = 1; c < 6; ++c)
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) s[i][j] = s[i][j] + bq[c * kAug + 3 * r2 + i] * wq[c * kAug + 3 * c2 + j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) k2[i][j] = k2[i][j] + s[i][j];
        if constexpr (KIN != kSmall) {
          const double geo = g[0 * 11 + r2] * L.sg[q][3 * c2] + g[1 * 11 + r2] * L.sg[q][3 * c2 + 1] +
                             g[2 * 11 + r2] * L.sg[q][3 * c2 + 2];
#pragma unroll
          for (int i = 0; i < 3; ++i) k2[i][i] = k2[i][i] + geo;
        }
      }
      if (fu) {
        const double w = geos[c0 + q][kGrW];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double s = bq[3 * fa + i] * L.rec[q][kWsStress];
#pragma unroll
          for (int c = 1; c < 6; ++c) s = s + bq[c * kAug + 3 * fa + i] * L.rec[q][kWsStress + c];
          f3[i] = f3[i] + w * s;
        }
      }
    }
  }
  __syncthreads();  // the operands are used: K_ua takes their place
  double* kua = L.ops + 576;  // K_ua, 24 x 9 row-major
  if (l < 24) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) kua[(3 * r2 + i) * 9 + 3 * (c2 - 8) + j] = k2[i][j];
  } else if (TANGENT && !SYM && l < 48) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) L.region[(3 * (r2 - 8) + i) * kRhs + 3 * c2 + j] = k2[i][j];
  }
  if (l == 0) lu_perm(L.lu.rt, L.lu.ct, L.perm);
  __syncthreads();
  if constexpr (TANGENT && SYM) {
    // K_au = K_ua^T
    for (int t = l; t < 216; t += 64) L.region[(t % 9) * kRhs + t / 9] = kua[t];
    __syncthreads();
  }
  // K_aa^-1 [K_au | r_a] in place, a column per lane
  if (l == 24 || (TANGENT && l < 24)) lu_solve_column(L.lu.a, L.perm, L.region, kRhs, l);
  __syncthreads();
  // f* = f_u - K_ua K_aa^-1 r_a
  if (fu) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double* row = kua + (3 * fa + i) * 9;
      double s = row[0] * L.region[24];
#pragma unroll
      for (int nn = 1; nn < 9; ++nn) s = s + row[nn] * L.region[nn * kRhs + 24];
      A.force[e * 24 + 3 * fa + i] = f3[i] - s;
    }
  }
  if constexpr (!TANGENT) return;
  // K*(a, b) = K_uu(a, b) - K_ua K_aa^-1 K_au
  double* ks = L.ops;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double* row = kua + (3 * a + i) * 9;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = row[0] * L.region[3 * b + j];
#pragma unroll
      for (int nn = 1; nn < 9; ++nn) s = s + row[nn] * L.region[nn * kRhs + 3 * b + j];
      ks[(3 * a + i) * 24 + 3 * b + j] = k1[i][j] - s;
    }
  }
  __syncthreads();
  double* out = A.tangent + e * 576;
  const bool cm = A.tangent_column_major != 0;
  for (int t = l; t < 576; t += 64) {
    const int r = cm ? t % 24 : t / 24, c = cm ? t / 24 : t % 24;
    out[t] = SYM ? 0.5 * (ks[r * 24 + c] + ks[c * 24 + r]) : ks[r * 24 + c];
  }
}

// ---------------------------------------------------------------------------
// Stress recovery (IncompatibleModes.cpp:602-664)
// ---------------------------------------------------------------------------
// One team per element, lane p = point p, at the committed history and modes
// (state_in, alpha_in): the point's return runs again from its committed
// state (an elastic check that reproduces it), then the stress measures of
// F = I + H: Cauchy, the second Piola-Kirchhoff stress and, for logarithmic
// kinematics, the Kirchhoff stress and the log strain, averaged over the
// points, with the extremes of the strain, det F and the equivalent plastic
// strain. out: kStressFields per element in the order of the reference's
// ElastoplasticStress; a failing element (set-up, then kinematics, then the
// returns, each in point order, as the reference's loops) gets zeros and its
// status.
template <int KIN, int NT, int TW>
__global__ __launch_bounds__(64) void stress_kernel(ElementArgs A, double* __restrict__ out) {
  const int lane = threadIdx.x;
  const int t = lane / TW;
  const int p = lane % TW;
  const long long e = static_cast<long long>(blockIdx.x) * teams_per_wave<TW>() + t;
  if (e >= A.ne) return;  // the whole team
  const int n = A.rule.n;
  // cauchy 6, piola_kirchhoff 6, kirchhoff 6, log strain 6 of the lane's point
  double v[24];
#pragma unroll
  for (int i = 0; i < 24; ++i) v[i] = 0.0;
  double max_strain = 0.0, min_jacobian = __builtin_inf(), max_abar = 0.0, fv = 0.0;
  int st = kOk, stage = 0;
  if (p < n) {
    double x0[8][3], ue[24];
    load_element(A.coords, A.conn, A.u, e, x0, ue);
    PointGeom pg;
    st = setup_point(x0, ue, A.rule.xi[p], A.rule.eta[p], A.rule.zeta[p], A.rule.w[p], pg, fv);
    if (st == kOk) {
      stage = 1;
      ModeGeom mg;
      mode_geom<KIN>(pg, ue, mg);
      double alpha[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) alpha[i] = A.alpha_in ? A.alpha_in[e * 9 + i] : 0.0;
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
        const double green[6] = {e3[0][0], e3[1][1], e3[2][2], 2.0 * e3[0][1], 2.0 * e3[1][2], 2.0 * e3[2][0]};
        if constexpr (KIN == kLog) {
          if (!(det3(F) > 0.0)) {
            st = kInverted;
          } else if (logarithmic_strain(green, L) != 0) {
            st = kDegenerate;
          } else {
#pragma unroll
            for (int i = 0; i < 6; ++i) strain[i] = L.strain[i];
          }
        } else {
#pragma unroll
          for (int i = 0; i < 6; ++i) strain[i] = green[i];
        }
      }
      if (st == kOk) {
        stage = 2;
        PointStateT<NT> ps;
        load_state(A.mat, A.state_in, A.state_stride, e * n + p, ps);
        ReturnOut ro;
        const int rf = A.mat.general ? return_general(A.mat, strain, ps, ro) : return_radial(A.mat, strain, ps, ro);
        if (rf == kReturnNotConverged) {
          st = kReturnFailed;
        } else if (rf == kReturnNonFinite) {
          st = kNonFiniteTrial;
        } else {
          double f[3][3], sym[3][3];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              f[i][j] = (i == j ? 1.0 : 0.0) + h[i][j];
              sym[i][j] = 0.5 * (h[i][j] + h[j][i]);
            }
          double pk[6];
#pragma unroll
          for (int i = 0; i < 6; ++i) pk[i] = ro.stress[i];
          if constexpr (KIN == kSmall) {
#pragma unroll
            for (int i = 0; i < 6; ++i) v[i] = ro.stress[i];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) max_strain = fmax(max_strain, fabs(sym[i][j]));
          } else {
            if constexpr (KIN == kLog) log_stress(L, ro.stress, pk);
            // F S F^T (tensorial Voigt)
            const double s3[3][3] = {{pk[0], pk[3], pk[5]}, {pk[3], pk[1], pk[4]}, {pk[5], pk[4], pk[2]}};
            double fs[3][3], fsf[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) fs[i][j] = f[i][0] * s3[0][j] + f[i][1] * s3[1][j] + f[i][2] * s3[2][j];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) fsf[i][j] = fs[i][0] * f[j][0] + fs[i][1] * f[j][1] + fs[i][2] * f[j][2];
            const double tau[6] = {fsf[0][0], fsf[1][1], fsf[2][2], fsf[0][1], fsf[1][2], fsf[2][0]};
            if constexpr (KIN == kLog) {
              double e6[6];
#pragma unroll
              for (int i = 0; i < 6; ++i) e6[i] = i < 3 ? strain[i] : strain[i] * 0.5;  // tensor components
              const double j = exp(e6[0] + e6[1] + e6[2]);
#pragma unroll
              for (int i = 0; i < 6; ++i) {
                v[i] = tau[i] / j;
                v[12 + i] = tau[i];
                v[18 + i] = e6[i];
                max_strain = fmax(max_strain, fabs(e6[i]));
              }
              min_jacobian = det3(f);
            } else {
              const double j = det3(f);
#pragma unroll
              for (int i = 0; i < 6; ++i) v[i] = tau[i] / j;
#pragma unroll
              for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                  const double hth = h[0][a] * h[0][b] + h[1][a] * h[1][b] + h[2][a] * h[2][b];
                  max_strain = fmax(max_strain, fabs(sym[a][b] + 0.5 * hth));
                }
              min_jacobian = j;
            }
          }
#pragma unroll
          for (int i = 0; i < 6; ++i) v[6 + i] = pk[i];
          max_abar = ps.abar;
        }
      }
    }
  }
  double* o = out + e * kStressFields;
  const int code = team_imin<TW>(st != kOk ? (stage << 16) | (p << 8) | st : 1 << 20);
  if (code != (1 << 20)) {
    for (int i = p; i < kStressFields; i += TW) o[i] = 0.0;
    if (p == 0 && A.status) A.status[e] = code & 0xFF;
    if (p == ((code >> 8) & 0xFF) && A.fail_value) A.fail_value[e] = fv;
    return;
  }
  const double count = static_cast<double>(n);
  double avg[24];
#pragma unroll
  for (int i = 0; i < 24; ++i) avg[i] = team_sum<TW>(v[i]) / count;
  max_strain = team_max<TW>(max_strain);
  min_jacobian = KIN == kSmall ? 1.0 : team_min<TW>(min_jacobian);
  max_abar = team_max<TW>(max_abar);
  const double a = avg[0] - avg[1], b = avg[1] - avg[2], c = avg[2] - avg[0];
  const double vm = sqrt(0.5 * (a * a + b * b + c * c) + 3.0 * (avg[3] * avg[3] + avg[4] * avg[4] + avg[5] * avg[5]));
  const double tail[4] = {vm, max_abar, max_strain, min_jacobian};
#pragma unroll
  for (int i = 0; i < 24; ++i)
    if ((i % TW) == p) o[i] = avg[i];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (((24 + i) % TW) == p) o[24 + i] = tail[i];
  if (p == 0 && A.status) A.status[e] = kOk;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// Wavefront slots of the current device at one wavefront per SIMD (the local
// solve's occupancy): 4 SIMDs per CU.
int resident_waves() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
  if (cached[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    cached[dev] = 4 * cus;
  }
  return cached[dev];
}

template <int KIN, int NT, int TW>
void launch_local_solve_nt(const ElementArgs& a, unsigned int blocks, hipStream_t stream) {
  if (a.ordered) {
    const int deal = static_cast<long long>(blocks) < 4ll * resident_waves() ? 1 : 0;
    constexpr int kTeams = teams_per_wave<TW>();
    const unsigned int teams = static_cast<unsigned int>((a.ne + kTeams - 1) / kTeams);
    const long long span = 2048;  // elements per block of the counting sort
    const unsigned int sort_blocks = static_cast<unsigned int>(std::min<long long>(1024, (a.ne + span - 1) / span));
    hipLaunchKernelGGL((cost_key_kernel<KIN, NT, TW>), dim3(teams), dim3(64), 0, stream, a);
    hipLaunchKernelGGL(cost_count_kernel, dim3(sort_blocks), dim3(256), 0, stream, a.cost, a.ne, a.counts);
    hipLaunchKernelGGL(cost_scan_kernel, dim3(1), dim3(64), 0, stream, a.counts);
    hipLaunchKernelGGL(cost_scatter_kernel, dim3(sort_blocks), dim3(256), 0, stream, a.cost, a.ne, a.counts,
                       static_cast<long long>(blocks), deal, a.order);
  }
  hipLaunchKernelGGL((local_solve_kernel<KIN, NT, TW>), dim3(blocks), dim3(64), 0, stream, a);
}

template <int KIN, int TW>
void launch_local_solve_kin(const ElementArgs& a, bool two_terms, unsigned int blocks, hipStream_t stream) {
  if (two_terms)
    launch_local_solve_nt<KIN, 2, TW>(a, blocks, stream);
  else
    launch_local_solve_nt<KIN, kMaxTerms, TW>(a, blocks, stream);
}

template <int TW>
void launch_local_solve_tw(const ElementArgs& a, bool two, unsigned int blocks, hipStream_t stream) {
  switch (a.kinematics) {
    case kSmall: launch_local_solve_kin<kSmall, TW>(a, two, blocks, stream); break;
    case kFinite: launch_local_solve_kin<kFinite, TW>(a, two, blocks, stream); break;
    default: launch_local_solve_kin<kLog, TW>(a, two, blocks, stream); break;
  }
}

// Lanes per element of a rule: 8 up to 8 points, 32 up to 32.
int team_width(int points) { return points <= kTeam ? kTeam : kWideTeam; }

int launch_local_solve(const ElementArgs& args, hipStream_t stream) {
  (void)hipGetLastError();
  ElementArgs a = args;
  if (a.rule.n < 1 || a.rule.n > kMaxPoints) return static_cast<int>(hipErrorInvalidValue);
  const int tw = team_width(a.rule.n);
  const int teams = 64 / tw;
  if (a.chunk <= 0) {
    // Elements per wavefront when the caller sets none: 64 (16 in 32-lane
    // teams), or fewer for a launch too small to give every SIMD's wavefront
    // slot two chunks (a small mesh would otherwise run on a few wavefronts,
    // each working through its elements in turn); a multiple of the teams per
    // wavefront, at least one per team.
    const int full = tw == kTeam ? kDefaultChunk : kDefaultChunk / 4;
    const long long per = a.ne / (2ll * resident_waves());
    a.chunk = per >= full ? full : (per < teams ? teams : static_cast<int>(per / teams) * teams);
  }
  if (a.chunk % teams != 0) return static_cast<int>(hipErrorInvalidValue);
  const unsigned int blocks = static_cast<unsigned int>((a.ne + a.chunk - 1) / a.chunk);
  if (blocks == 0) return 0;
  // the smallest instantiation that holds the material's kinematic history
  const int fields = a.mat.terms > a.mat.kin_fields ? a.mat.terms : a.mat.kin_fields;
  if (fields > kMaxTerms) return static_cast<int>(hipErrorInvalidValue);
  const bool two = fields <= 2;
  if (tw == kTeam)
    launch_local_solve_tw<kTeam>(a, two, blocks, stream);
  else
    launch_local_solve_tw<kWideTeam>(a, two, blocks, stream);
  return static_cast<int>(hipGetLastError());
}

template <int KIN, bool TANGENT, bool SYM, int TW>
void launch_condense_one(const ElementArgs& a, hipStream_t stream) {
  if constexpr (TW == kTeam) {
    const unsigned int blocks = static_cast<unsigned int>((a.ne + kTeam - 1) / kTeam);  // 8 teams per wavefront
    hipLaunchKernelGGL((condense_team_kernel<KIN, TANGENT, SYM, TW>), dim3(blocks), dim3(64), 0, stream, a);
  } else {
    const unsigned int blocks = static_cast<unsigned int>(a.ne);  // an element per workgroup
    // the geometry records in dynamic LDS: the rule's points, not the most a rule may hold
    const size_t geo_bytes = static_cast<size_t>(a.rule.n) * kGeomRecord * sizeof(double);
    hipLaunchKernelGGL((condense_wide_kernel<KIN, TANGENT, SYM, TW>), dim3(blocks), dim3(64), geo_bytes, stream, a);
  }
}

template <int KIN, int TW>
void launch_condense_kin(const ElementArgs& a, bool tangent, hipStream_t stream) {
  if (!tangent)
    launch_condense_one<KIN, false, true, TW>(a, stream);
  else if (a.mat.symmetric)
    launch_condense_one<KIN, true, true, TW>(a, stream);
  else
    launch_condense_one<KIN, true, false, TW>(a, stream);
}

template <int KIN, int TW>
void launch_stress_kin(const ElementArgs& a, bool two_terms, double* out, unsigned int blocks, hipStream_t stream) {
  if (two_terms)
    hipLaunchKernelGGL((stress_kernel<KIN, 2, TW>), dim3(blocks), dim3(64), 0, stream, a, out);
  else
    hipLaunchKernelGGL((stress_kernel<KIN, kMaxTerms, TW>), dim3(blocks), dim3(64), 0, stream, a, out);
}

template <int TW>
void launch_stress_tw(const ElementArgs& a, bool two, double* out, unsigned int blocks, hipStream_t stream) {
  switch (a.kinematics) {
    case kSmall: launch_stress_kin<kSmall, TW>(a, two, out, blocks, stream); break;
    case kFinite: launch_stress_kin<kFinite, TW>(a, two, out, blocks, stream); break;
    default: launch_stress_kin<kLog, TW>(a, two, out, blocks, stream); break;
  }
}

int launch_element_stress(const ElementArgs& a, double* out, hipStream_t stream) {
  (void)hipGetLastError();
  if (a.rule.n < 1 || a.rule.n > kMaxPoints) return static_cast<int>(hipErrorInvalidValue);
  const int tw = team_width(a.rule.n);
  const int teams = 64 / tw;
  const unsigned int blocks = static_cast<unsigned int>((a.ne + teams - 1) / teams);
  if (blocks == 0) return 0;
  const int fields = a.mat.terms > a.mat.kin_fields ? a.mat.terms : a.mat.kin_fields;
  if (fields > kMaxTerms) return static_cast<int>(hipErrorInvalidValue);
  const bool two = fields <= 2;
  if (tw == kTeam)
    launch_stress_tw<kTeam>(a, two, out, blocks, stream);
  else
    launch_stress_tw<kWideTeam>(a, two, out, blocks, stream);
  return static_cast<int>(hipGetLastError());
}

template <int TW>
void launch_condense_tw(const ElementArgs& a, bool tangent, hipStream_t stream) {
  switch (a.kinematics) {
    case kSmall: launch_condense_kin<kSmall, TW>(a, tangent, stream); break;
    case kFinite: launch_condense_kin<kFinite, TW>(a, tangent, stream); break;
    default: launch_condense_kin<kLog, TW>(a, tangent, stream); break;
  }
}

int launch_condense(const ElementArgs& a, bool tangent, hipStream_t stream) {
  (void)hipGetLastError();
  if (a.rule.n < 1 || a.rule.n > kMaxPoints) return static_cast<int>(hipErrorInvalidValue);
  if (a.ne == 0) return 0;
  if (team_width(a.rule.n) == kTeam)
    launch_condense_tw<kTeam>(a, tangent, stream);
  else
    launch_condense_tw<kWideTeam>(a, tangent, stream);
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
