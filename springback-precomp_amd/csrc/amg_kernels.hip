// Smoothed-aggregation AMG on the device: the numeric set-up of each level
// (diagonal, power estimate, Chebyshev coefficients, P, A P, P^T, the next
// level's values), the coarsest level's dense inverse and the V-cycle.
//
// Row-wise kernels give each row 32 lanes (two rows per wavefront); a lane
// takes the row's entries l, l + 32, ... and a butterfly adds the lanes, so
// every sum has one order. The sparse products C = X Y (pattern of C from
// the host) give each row of C one 64-lane workgroup: the row's columns and
// sums sit in LDS; the (X entry, Y entry) pairs of up to 64 of the row's X
// entries are loaded 64 at a time, each lane finding its column's slot by
// binary search, and added in the order of X's entries; two entries of one
// row of Y never share a slot, so each slot's sum runs in that order.

#include "amg_kernels.hpp"

#include <cmath>

namespace sbk {

namespace {
// Every launch is grouped: it covers one level of each of a set of
// hierarchies (one per problem; a plain solve's set has one), each entry of
// the group table taking its own range of the grid's blocks with exactly the
// block decomposition a launch of its own would use. A workgroup finds its
// entry in the table's block map and, when its problem rests (its active
// flag is 0), does nothing.
struct GroupBlock {
  const AmgGroupEntry* e;
  long long block;    // the block within the entry's range
  long long nblocks;  // the range's length
};

__device__ __forceinline__ bool group_block(const AmgGroup& G, GroupBlock& out) {
  const long long b = blockIdx.x;
  const int j = G.map[b];
  out.e = G.e + j;
  out.block = b - G.e[j].first;
  out.nblocks = G.e[j + 1].first - G.e[j].first;
  return !(G.active && G.active[G.e[j].problem] == 0);
}

// Kernels copy their entry's level descriptor (const AmgLevel L = e.L) so
// that stores through its pointers cannot alias it: the compiler keeps its
// fields in registers instead of reloading them after every store.
//
// The V-cycle's vectors of an entry's level: the caller's (at the problem's
// offset) on level 0, the level's own above it.
__device__ __forceinline__ const double* level_b(const AmgGroupEntry& e, const double* b) {
  return e.offset >= 0 ? b + e.offset : e.L.b;
}
__device__ __forceinline__ double* level_x(const AmgGroupEntry& e, double* x) {
  return e.offset >= 0 ? x + e.offset : e.L.x;
}

constexpr int kWaves = 4;
constexpr int kLanesPerRow = 16;
constexpr int kGroups = 64 / kLanesPerRow;
constexpr int kRowsPerBlock = kWaves * kGroups;
constexpr int kCap = 512;  // columns of C a product workgroup holds at a time

unsigned int grid_of(long long n, long long per) { return static_cast<unsigned int>((n + per - 1) / per); }

// sum over the row of values[k] x[index[k]] (a row >= n sums to 0), in every
// lane of the row's 16-lane group: lane l adds entries l, l + 16, ... in
// order, then a butterfly over the group. The lane's first kSlots entries
// have every index and value load issued before their gathers of x, and
// those before the sums (a row of level 0 holds at most 81 entries); longer
// rows finish in the loop after, in the same order.
constexpr int kSlots = 4;
__device__ __forceinline__ double row_dot(const int* __restrict__ ptr, const int* __restrict__ index,
                                          const double* __restrict__ values, const double* __restrict__ x,
                                          long long row, long long n, int l) {
  double acc = 0.0;
  if (row < n) {
    const int k0 = ptr[row] + l, k1 = ptr[row + 1];
    int col[kSlots];
    double v[kSlots], xv[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int k = k0 + s * kLanesPerRow;
      col[s] = k < k1 ? index[k] : -1;
      v[s] = k < k1 ? values[k] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < kSlots; ++s) xv[s] = col[s] >= 0 ? x[col[s]] : 0.0;
#pragma unroll
    for (int s = 0; s < kSlots; ++s)
      if (col[s] >= 0) acc = fma(v[s], xv[s], acc);
    for (int k = k0 + kSlots * kLanesPerRow; k < k1; k += kLanesPerRow) acc = fma(values[k], x[index[k]], acc);
  }
#pragma unroll
  for (int off = kLanesPerRow / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  return acc;
}

// Row-wise launches deal the rows to an entry's blocks round-robin, kGroups
// rows per wavefront at a time (a block's prologue, its entry's lookup, is
// paid once for many rows); a row's result does not depend on the dealing.
// Row r0 of the calling lane's first turn, the stride between its turns, and
// the lane within its row's group.
struct RowLoop {
  long long first, stride;
  int g, l;
};

__device__ __forceinline__ RowLoop row_loop(const GroupBlock& gb) {
  const int lane = threadIdx.x & 63;
  RowLoop r;
  r.first = (gb.block * kWaves + (threadIdx.x >> 6)) * kGroups;
  r.stride = gb.nblocks * kRowsPerBlock;
  r.g = lane / kLanesPerRow;
  r.l = lane % kLanesPerRow;
  return r;
}

__global__ __launch_bounds__(256) void diag_kernel(AmgGroup G) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  for (long long i = gb.block * 256 + threadIdx.x; i < L.n; i += gb.nblocks * 256) {
    double d = 0.0;
    for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k)
      if (L.index[k] == i) d = L.values[k];
    L.dinv[i] = (d != 0.0 && std::isfinite(d)) ? 1.0 / d : 0.0;
  }
}

// A fixed start vector in [-1, 1) from a hash of the row (into d0).
__global__ __launch_bounds__(256) void power_init_kernel(AmgGroup G) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  if (gb.block == 0 && threadIdx.x == 0) L.coef[kAmgScale] = 1.0;
  for (long long i = gb.block * 256 + threadIdx.x; i < L.n; i += gb.nblocks * 256) {
    unsigned long long h = static_cast<unsigned long long>(i) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    L.d0[i] = static_cast<double>(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

// y = D^-1 A (x / ||x||) (x = d0, y = d1 on even steps, the other way round
// on odd ones), rows dealt round-robin over the level's nparts blocks;
// partial[block] = the block's sum of y^2 (groups, then wavefronts in order).
__global__ __launch_bounds__(256) void power_step_kernel(AmgGroup G, int odd) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  const double* __restrict__ x = odd ? L.d1 : L.d0;
  double* __restrict__ y = odd ? L.d0 : L.d1;
  __shared__ double ws[kRowsPerBlock];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane / kLanesPerRow, l = lane % kLanesPerRow;
  const double scale = L.coef[kAmgScale];
  const long long stride = gb.nblocks * kRowsPerBlock;
  double sq = 0.0;
  for (long long base = (gb.block * kWaves + wave) * kGroups; base < L.n; base += stride) {
    const long long row = base + g;
    const double v = row_dot(L.ptr, L.index, L.values, x, row, L.n, l);
    if (l == 0 && row < L.n) {
      const double yi = L.dinv[row] * scale * v;
      y[row] = yi;
      sq = fma(yi, yi, sq);
    }
  }
  if (l == 0) ws[wave * kGroups + g] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int i = 0; i < kRowsPerBlock; ++i) sum += ws[i];
    L.partial[gb.block] = sum;
  }
}

// ||y|| (the estimate, x having been normalised) and 1 / ||y||.
__global__ __launch_bounds__(256) void power_finish_kernel(AmgGroup G) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < L.nparts; i += 256) acc += L.partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double norm = std::sqrt(s[0]);
    L.coef[kAmgRho] = norm;
    L.coef[kAmgScale] = (norm > 0.0 && std::isfinite(norm)) ? 1.0 / norm : 0.0;
  }
}

// Chebyshev on [lmax / ratio, lmax], lmax = boost rho (degree 2: d0 = r0 /
// theta, d1 = c1 d0 + c2 r1), and the prolongator's omega = 4 / (3 rho); no
// smoothing at all when rho is not a positive finite number.
__global__ void coef_kernel(AmgGroup G) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  if (threadIdx.x != 0) return;
  const double rho = L.coef[kAmgRho];
  double* c = L.coef;
  if (!(rho > 0.0) || !std::isfinite(rho)) {
    c[kAmgThetaInv] = c[kAmgC1] = c[kAmgC2] = c[kAmgOmega] = 0.0;
    return;
  }
  const double lmax = kAmgEigenBoost * rho, lmin = lmax / kAmgEigenRatio;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  const double rho0 = 1.0 / sigma;
  const double rho1 = 1.0 / (2.0 * sigma - rho0);
  c[kAmgThetaInv] = 1.0 / theta;
  c[kAmgC1] = rho1 * rho0;
  c[kAmgC2] = 2.0 * rho1 / delta;
  c[kAmgOmega] = 4.0 / (3.0 * rho);
}

struct Csr {
  const int* ptr;
  const int* index;
  const double* values;
};

// The slot of column c among cols[0, mw) (ascending), or -1.
__device__ __forceinline__ int find_slot(const int* cols, int mw, int c) {
  int lo = 0, hi = mw;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cols[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < mw && cols[lo] == c) ? lo : -1;
}

// C = X Y on C's pattern, one 64-lane workgroup (a single wavefront) per row
// of C, in windows of kCap columns. X's entries of the row are taken 64 at a
// time: their (X entry, Y entry) pairs are numbered by a scan of their rows
// of Y, loaded 64 at a time with each lane's slot (its column's place among
// C's columns, by binary search), and added in the order of X's entries. A
// wavefront's LDS accesses run in program order, so an entry's additions see
// the previous entry's without a barrier; two entries of one row of Y never
// share a slot, so each slot's sum runs in the order of X's entries.
// kind kProductP: C = P, the smoothed prolongator, written as
// T - omega D^-1 (A T) (X = A, Y = T); kProductAP: C = A P (X = A, Y = P);
// kProductPtAP: C = the next level's A = P^T (A P) (X = P^T, Y = A P), over
// the coarse rows. One block per row of C.
enum ProductKind : int { kProductP = 0, kProductAP = 1, kProductPtAP = 2 };

__global__ __launch_bounds__(64) void product_kernel(AmgGroup G, int kind) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  const Csr A{L.ptr, L.index, L.values};
  Csr X = A, Y{L.tp_ptr, L.tp_index, L.tp_values};
  const int* __restrict__ cptr = L.p_ptr;
  const int* __restrict__ cindex = L.p_index;
  double* __restrict__ cvalues = L.p_values;
  long long rows = L.n;
  if (kind == kProductAP) {
    Y = Csr{L.p_ptr, L.p_index, L.p_values};
    cptr = L.ap_ptr;
    cindex = L.ap_index;
    cvalues = L.ap_values;
  } else if (kind == kProductPtAP) {
    const AmgLevel N = *gb.e->N;
    X = Csr{L.pt_ptr, L.pt_index, L.pt_values};
    Y = Csr{L.ap_ptr, L.ap_index, L.ap_values};
    cptr = N.ptr;
    cindex = N.index;
    cvalues = const_cast<double*>(N.values);
    rows = L.nc;
  }
  const int smoothed = kind == kProductP;
  __shared__ int cols[kCap];
  __shared__ double acc[kCap];
  __shared__ int xs_start[64], xs_yb[64];
  __shared__ double xs_a[64];
  const long long row = gb.block;
  if (row >= rows) return;
  const int lane = threadIdx.x;
  const int c0 = cptr[row], m = cptr[row + 1] - c0;
  const int x0 = X.ptr[row], x1 = X.ptr[row + 1];
  double scale = 0.0;
  int own = 0, t0 = 0, tn = 0;
  if (smoothed) {
    scale = L.coef[kAmgOmega] * L.dinv[row];
    own = L.p_own[row];
    t0 = L.tp_ptr[row];
    tn = L.tp_ptr[row + 1] - t0;
  }
  for (int w0 = 0; w0 < m; w0 += kCap) {
    const int mw = m - w0 < kCap ? m - w0 : kCap;
    for (int s = lane; s < mw; s += 64) {
      cols[s] = cindex[c0 + w0 + s];
      acc[s] = 0.0;
    }
    __syncthreads();
    const int lo_col = cols[0], hi_col = cols[mw - 1];
    for (int b0 = x0; b0 < x1; b0 += 64) {
      // up to 64 entries of X's row: each lane loads one, with its row of Y's
      // bounds; a scan of the rows' lengths numbers their (X entry, Y entry)
      // pairs
      const int t = b0 + lane;
      const bool in = t < x1;
      const int k = in ? X.index[t] : 0;
      const double a_l = in ? X.values[t] : 0.0;
      const int yb_l = in ? Y.ptr[k] : 0;
      const int len_l = in ? Y.ptr[k + 1] - yb_l : 0;
      int incl = len_l;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
      }
      const int total = __shfl(incl, 63, 64);
      const int cnt = x1 - b0 < 64 ? x1 - b0 : 64;
      xs_start[lane] = incl - len_l;
      xs_a[lane] = a_l;
      xs_yb[lane] = yb_l;
      __syncthreads();
      for (int f0 = 0; f0 < total; f0 += 64) {
        // 64 pairs at a time: every load issued together, each lane's slot
        // found, then the pairs added in X's order (the lanes of one X entry
        // hold distinct columns), so each slot's sum runs in X's order
        const int f = f0 + lane;
        int jj = 1 << 30, slot = -1;
        double a = 0.0, v = 0.0;
        if (f < total) {
          int lo = 0, hi = cnt - 1;  // the last X entry whose pairs start at or before f
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (xs_start[mid] <= f) lo = mid; else hi = mid - 1;
          }
          jj = lo;
          const int u = xs_yb[lo] + (f - xs_start[lo]);
          const int c = Y.index[u];
          v = Y.values[u];
          a = xs_a[lo];
          if (c >= lo_col && c <= hi_col) slot = find_slot(cols, mw, c);
        }
        const int j_first = __shfl(jj, 0, 64);
        const int last_lane = total - f0 < 64 ? total - f0 - 1 : 63;
        const int j_last = __shfl(jj, last_lane, 64);
        for (int j = j_first; j <= j_last; ++j)
          if (jj == j && slot >= 0) acc[slot] = fma(a, v, acc[slot]);
      }
      __syncthreads();  // before the next chunk rewrites xs_*
    }
    __syncthreads();
    for (int s = lane; s < mw; s += 64) {
      double v = acc[s];
      if (smoothed) {
        const int j = w0 + s - own;  // the tentative column's place in T's row
        const double tv = (j >= 0 && j < tn) ? L.tp_values[t0 + j] : 0.0;
        v = tv - scale * v;
      }
      cvalues[c0 + w0 + s] = v;
    }
    __syncthreads();
  }
}

// P^T's values from P's (pt_values[i] = p_values[pt_perm[i]]).
__global__ __launch_bounds__(256) void gather_kernel(AmgGroup G) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  for (long long i = gb.block * 256 + threadIdx.x; i < L.pt_nnz; i += gb.nblocks * 256)
    L.pt_values[i] = L.p_values[L.pt_perm[i]];
}

// [A | I] -> [I | A^-1] by Gauss-Jordan with partial pivoting (the largest
// |entry| of the column, the lowest row among equals); one workgroup.
__global__ __launch_bounds__(1024) void dense_inverse_kernel(AmgGroup G) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  __shared__ double best[16];
  __shared__ int besti[16];
  __shared__ double factor[kAmgDenseMax];
  const int n = static_cast<int>(L.n);
  const int w = 2 * n;
  double* M = L.dense;
  const int tid = threadIdx.x;
  for (long long i = tid; i < static_cast<long long>(n) * w; i += 1024) M[i] = 0.0;
  if (tid == 0) L.coef[kAmgSingular] = 0.0;
  __syncthreads();
  for (int row = tid; row < n; row += 1024) {
    for (int k = L.ptr[row]; k < L.ptr[row + 1]; ++k) M[static_cast<long long>(row) * w + L.index[k]] = L.values[k];
    M[static_cast<long long>(row) * w + n + row] = 1.0;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    double v = -1.0;
    int vi = n;
    for (int i = j + tid; i < n; i += 1024) {
      const double a = fabs(M[static_cast<long long>(i) * w + j]);
      if (a > v) {
        v = a;
        vi = i;
      }
    }
    // the largest within each wavefront (lowest row among equals), then
    // across the 16 wavefronts
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off, 64);
      const int oi = __shfl_xor(vi, off, 64);
      if (o > v || (o == v && oi < vi)) {
        v = o;
        vi = oi;
      }
    }
    if ((tid & 63) == 0) {
      best[tid >> 6] = v;
      besti[tid >> 6] = vi;
    }
    __syncthreads();
    double pv = best[0];
    int p = besti[0];
    for (int wv = 1; wv < 16; ++wv)
      if (best[wv] > pv || (best[wv] == pv && besti[wv] < p)) {
        pv = best[wv];
        p = besti[wv];
      }
    if (!(pv > 0.0) || !std::isfinite(pv)) {  // every thread sees the same pivot
      if (tid == 0) L.coef[kAmgSingular] = 1.0;
      return;
    }
    if (p != j)
      for (int c = tid; c < w; c += 1024) {
        const double t = M[static_cast<long long>(p) * w + c];
        M[static_cast<long long>(p) * w + c] = M[static_cast<long long>(j) * w + c];
        M[static_cast<long long>(j) * w + c] = t;
      }
    __syncthreads();
    const double inv = 1.0 / M[static_cast<long long>(j) * w + j];
    __syncthreads();
    for (int c = tid; c < w; c += 1024) M[static_cast<long long>(j) * w + c] *= inv;
    for (int i = tid; i < n; i += 1024) factor[i] = M[static_cast<long long>(i) * w + j];
    __syncthreads();
    // a wavefront per row (16 at a time); the pivot row's columns left of j
    // are zero already
    const double* pj = M + static_cast<long long>(j) * w;
    for (int i = tid >> 6; i < n; i += 16) {
      const double f = factor[i];
      if (i == j || f == 0.0) continue;
      double* mi = M + static_cast<long long>(i) * w;
      for (int c = j + (tid & 63); c < w; c += 64) mi[c] = fma(-f, pj[c], mi[c]);
    }
    __syncthreads();
  }
}

// Level 0's symmetric part (A + A^T) / 2 from the entry's block values.
__global__ __launch_bounds__(256) void symmetric_part_kernel(AmgGroup G) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgGroupEntry& e = *gb.e;
  const double* __restrict__ v = e.src;
  for (long long k = gb.block * 256 + threadIdx.x; k < e.count; k += gb.nblocks * 256) {
    const int m = e.mirror[k];
    e.dst[k] = m >= 0 ? 0.5 * (v[k] + v[m]) : 0.5 * v[k];
  }
}

// --- the V-cycle ------------------------------------------------------------

// Pre-smoothing from x = 0, first half: r = D^-1 b, d0 = r / theta.
__global__ __launch_bounds__(256) void pre_first_kernel(AmgGroup G, const double* __restrict__ b0) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  const double* __restrict__ b = level_b(*gb.e, b0);
  const double theta_inv = L.coef[kAmgThetaInv];
  for (long long i = gb.block * 256 + threadIdx.x; i < L.n; i += gb.nblocks * 256) {
    const double ri = L.dinv[i] * b[i];
    L.r[i] = ri;
    L.d0[i] = theta_inv * ri;
  }
}

// Second half: r -= D^-1 A d0, d1 = c1 d0 + c2 r, x = d0 + d1.
__global__ __launch_bounds__(256) void pre_step_kernel(AmgGroup G, double* __restrict__ x0) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  double* __restrict__ x = level_x(*gb.e, x0);
  const double* __restrict__ d0 = L.d0;
  const double* __restrict__ dinv = L.dinv;
  double* __restrict__ r = L.r;
  const double c1 = L.coef[kAmgC1], c2 = L.coef[kAmgC2];
  const RowLoop q = row_loop(gb);
  for (long long base = q.first; base < L.n; base += q.stride) {
    const long long i = base + q.g;
    const double v = row_dot(L.ptr, L.index, L.values, d0, i, L.n, q.l);
    if (q.l != 0 || i >= L.n) continue;
    const double ri = fma(-dinv[i], v, r[i]);
    const double d1 = fma(c1, d0[i], c2 * ri);
    r[i] = ri;
    x[i] = d0[i] + d1;
  }
}

// r = b - A x.
__global__ __launch_bounds__(256) void residual_kernel(AmgGroup G, const double* __restrict__ b0,
                                                       const double* __restrict__ x0) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  const double* __restrict__ b = level_b(*gb.e, b0);
  const double* __restrict__ x = level_x(*gb.e, const_cast<double*>(x0));
  double* __restrict__ r = L.r;
  const RowLoop q = row_loop(gb);
  for (long long base = q.first; base < L.n; base += q.stride) {
    const long long i = base + q.g;
    const double v = row_dot(L.ptr, L.index, L.values, x, i, L.n, q.l);
    if (q.l == 0 && i < L.n) r[i] = b[i] - v;
  }
}

// The next level's b = P^T r.
__global__ __launch_bounds__(256) void restrict_kernel(AmgGroup G) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  double* __restrict__ nb = gb.e->N->b;
  const double* __restrict__ r = L.r;
  const RowLoop q = row_loop(gb);
  for (long long base = q.first; base < L.nc; base += q.stride) {
    const long long i = base + q.g;
    const double v = row_dot(L.pt_ptr, L.pt_index, L.pt_values, r, i, L.nc, q.l);
    if (q.l == 0 && i < L.nc) nb[i] = v;
  }
}

// x += P x_next.
__global__ __launch_bounds__(256) void prolong_kernel(AmgGroup G, double* __restrict__ x0) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  double* __restrict__ x = level_x(*gb.e, x0);
  const double* __restrict__ nx = gb.e->N->x;
  const RowLoop q = row_loop(gb);
  for (long long base = q.first; base < L.n; base += q.stride) {
    const long long i = base + q.g;
    const double v = row_dot(L.p_ptr, L.p_index, L.p_values, nx, i, L.n, q.l);
    if (q.l == 0 && i < L.n) x[i] += v;
  }
}

// Post-smoothing from x, first half: r = D^-1 (b - A x), d0 = r / theta.
__global__ __launch_bounds__(256) void post_first_kernel(AmgGroup G, const double* __restrict__ b0,
                                                         const double* __restrict__ x0) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  const double* __restrict__ b = level_b(*gb.e, b0);
  const double* __restrict__ x = level_x(*gb.e, const_cast<double*>(x0));
  const double* __restrict__ dinv = L.dinv;
  double* __restrict__ r = L.r;
  double* __restrict__ d0 = L.d0;
  const double theta_inv = L.coef[kAmgThetaInv];
  const RowLoop q = row_loop(gb);
  for (long long base = q.first; base < L.n; base += q.stride) {
    const long long i = base + q.g;
    const double v = row_dot(L.ptr, L.index, L.values, x, i, L.n, q.l);
    if (q.l != 0 || i >= L.n) continue;
    const double ri = dinv[i] * (b[i] - v);
    r[i] = ri;
    d0[i] = theta_inv * ri;
  }
}

// Second half: r -= D^-1 A d0, d1 = c1 d0 + c2 r, x += d0 + d1.
__global__ __launch_bounds__(256) void post_step_kernel(AmgGroup G, double* __restrict__ x0) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  double* __restrict__ x = level_x(*gb.e, x0);
  const double* __restrict__ d0 = L.d0;
  const double* __restrict__ dinv = L.dinv;
  double* __restrict__ r = L.r;
  const double c1 = L.coef[kAmgC1], c2 = L.coef[kAmgC2];
  const RowLoop q = row_loop(gb);
  for (long long base = q.first; base < L.n; base += q.stride) {
    const long long i = base + q.g;
    const double v = row_dot(L.ptr, L.index, L.values, d0, i, L.n, q.l);
    if (q.l != 0 || i >= L.n) continue;
    const double ri = fma(-dinv[i], v, r[i]);
    const double d1 = fma(c1, d0[i], c2 * ri);
    r[i] = ri;
    x[i] += d0[i] + d1;
  }
}

// The coarsest level: x = A^-1 b from the dense inverse (each row's sum in
// column order).
__global__ __launch_bounds__(256) void dense_apply_kernel(AmgGroup G, const double* __restrict__ b0,
                                                          double* __restrict__ x0) {
  GroupBlock gb;
  if (!group_block(G, gb)) return;
  const AmgLevel L = gb.e->L;
  const double* __restrict__ b = level_b(*gb.e, b0);
  double* __restrict__ x = level_x(*gb.e, x0);
  const int n = static_cast<int>(L.n);
  const long long w = 2ll * n;
  for (long long i = gb.block * 256 + threadIdx.x; i < n; i += gb.nblocks * 256) {
    const double* row = L.dense + i * w + n;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc = fma(row[j], b[j], acc);
    x[i] = acc;
  }
}

int launch(void (*kernel)(AmgGroup), const AmgTable& t, const int* active, int threads, hipStream_t s) {
  if (t.blocks > 0)
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned int>(t.blocks)), dim3(threads), 0, s,
                       AmgGroup{t.entries, t.map, active});
  return 0;
}

}  // namespace

int amg_product_blocks(long long n) {
  const long long b = (n + 4 * kRowsPerBlock - 1) / (4 * kRowsPerBlock);
  return static_cast<int>(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

// An entry's blocks: one turn per thread (elementwise) or per row group
// (row-wise) up to 2048 blocks, so that a small level's launch is one short
// turn (its latency), a large one's blocks take many turns each (their
// prologue paid once).
long long elem_blocks(long long n) {
  const long long b = grid_of(n, 256);
  return b < 1 ? 1 : (b > 2048 ? 2048 : b);
}
long long row_blocks(long long n) {
  const long long b = grid_of(n, kRowsPerBlock);
  return b < 1 ? 1 : (b > 2048 ? 2048 : b);
}

long long amg_table_blocks(AmgTableKind kind, const AmgLevel& L, long long nnz0) {
  switch (kind) {
    case kAmgTableElem: return elem_blocks(L.n);
    case kAmgTableRows: return row_blocks(L.n);
    case kAmgTableRestrict: return row_blocks(L.nc);
    case kAmgTablePower: return L.nparts;
    case kAmgTableOne: return 1;
    case kAmgTableProduct: return L.n;
    case kAmgTableCoarseProduct: return L.nc;
    case kAmgTableGather: return elem_blocks(L.pt_nnz);
    case kAmgTableSymmetric: return elem_blocks(nnz0);
    default: return 0;
  }
}

int launch_amg_setup(const AmgTables& T, bool symmetrise, hipStream_t s) {
  (void)hipGetLastError();
  if (symmetrise) launch(symmetric_part_kernel, T.symmetric, nullptr, 256, s);
  for (std::size_t l = 0; l < T.levels.size(); ++l) {
    const AmgLevelTables& t = T.levels[l];
    launch(diag_kernel, t.elem, nullptr, 256, s);
    launch(power_init_kernel, t.elem, nullptr, 256, s);
    if (t.power.blocks > 0)
      for (int k = 0; k < kAmgPowerSteps; ++k) {
        hipLaunchKernelGGL(power_step_kernel, dim3(static_cast<unsigned int>(t.power.blocks)), dim3(256), 0, s,
                           AmgGroup{t.power.entries, t.power.map, nullptr}, k & 1);
        launch(power_finish_kernel, t.one, nullptr, 256, s);
      }
    launch(coef_kernel, t.one, nullptr, 64, s);
    if (t.product.blocks > 0) {
      const AmgGroup g{t.product.entries, t.product.map, nullptr};
      hipLaunchKernelGGL(product_kernel, dim3(static_cast<unsigned int>(t.product.blocks)), dim3(64), 0, s, g,
                         static_cast<int>(kProductP));
      hipLaunchKernelGGL(product_kernel, dim3(static_cast<unsigned int>(t.product.blocks)), dim3(64), 0, s, g,
                         static_cast<int>(kProductAP));
    }
    launch(gather_kernel, t.gather, nullptr, 256, s);
    if (t.coarse_product.blocks > 0)
      hipLaunchKernelGGL(product_kernel, dim3(static_cast<unsigned int>(t.coarse_product.blocks)), dim3(64), 0, s,
                         AmgGroup{t.coarse_product.entries, t.coarse_product.map, nullptr},
                         static_cast<int>(kProductPtAP));
    launch(dense_inverse_kernel, t.dense_one, nullptr, 1024, s);
  }
  return static_cast<int>(hipGetLastError());
}

int launch_amg_vcycle(const AmgTables& T, const int* active, const double* b, double* x, hipStream_t s) {
  (void)hipGetLastError();
  const int count = static_cast<int>(T.levels.size());
  auto grid = [](const AmgTable& t) { return dim3(static_cast<unsigned int>(t.blocks)); };
  auto group = [active](const AmgTable& t) { return AmgGroup{t.entries, t.map, active}; };
  const double* cb = b;
  // down: pre-smoothing, residual and restriction where a level is not the
  // problem's coarsest, the dense solve where it is
  for (int l = 0; l < count; ++l) {
    const AmgLevelTables& t = T.levels[l];
    if (t.elem.blocks > 0)
      hipLaunchKernelGGL(pre_first_kernel, grid(t.elem), dim3(256), 0, s, group(t.elem), cb);
    if (t.rows.blocks > 0) {
      hipLaunchKernelGGL(pre_step_kernel, grid(t.rows), dim3(256), 0, s, group(t.rows), x);
      hipLaunchKernelGGL(residual_kernel, grid(t.rows), dim3(256), 0, s, group(t.rows), cb,
                         static_cast<const double*>(x));
    }
    launch(restrict_kernel, t.restrict_, active, 256, s);
    if (t.dense.blocks > 0)
      hipLaunchKernelGGL(dense_apply_kernel, grid(t.dense), dim3(256), 0, s, group(t.dense), cb, x);
  }
  // up: prolongation and post-smoothing
  for (int l = count - 1; l >= 0; --l) {
    const AmgLevelTables& t = T.levels[l];
    if (t.rows.blocks == 0) continue;
    hipLaunchKernelGGL(prolong_kernel, grid(t.rows), dim3(256), 0, s, group(t.rows), x);
    hipLaunchKernelGGL(post_first_kernel, grid(t.rows), dim3(256), 0, s, group(t.rows), cb,
                       static_cast<const double*>(x));
    hipLaunchKernelGGL(post_step_kernel, grid(t.rows), dim3(256), 0, s, group(t.rows), x);
  }
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
