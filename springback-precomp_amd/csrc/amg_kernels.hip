// Smoothed-aggregation AMG on the device: the numeric set-up of each level
// (diagonal, power estimate, Chebyshev coefficients, P, A P, P^T, the next
// level's values), the coarsest level's dense inverse and the V-cycle.
//
// Row-wise kernels give each row 32 lanes (two rows per wavefront); a lane
// takes the row's entries l, l + 32, ... and a butterfly adds the lanes, so
// every sum has one order. The sparse products C = X Y (pattern of C from
// the host) give each row of C one 64-lane workgroup: the row's columns and
// sums sit in LDS, X's entries of the row are taken in order, and the lanes
// share out the entries of Y's matching row, finding each column's slot by
// binary search; two entries of one row of Y never share a slot, so each
// slot's sum runs in the order of X's entries.

#include "amg_kernels.hpp"

#include <cmath>

namespace sbk {

namespace {

constexpr int kWaves = 4;
constexpr int kLanesPerRow = 32;
constexpr int kGroups = 64 / kLanesPerRow;
constexpr int kRowsPerBlock = kWaves * kGroups;
constexpr int kCap = 512;  // columns of C a product workgroup holds at a time

unsigned int grid_of(long long n, long long per) { return static_cast<unsigned int>((n + per - 1) / per); }

// sum over the row of values[k] x[index[k]] (a row >= n sums to 0), in every
// lane of the row's 32-lane group.
__device__ __forceinline__ double row_dot(const int* __restrict__ ptr, const int* __restrict__ index,
                                          const double* __restrict__ values, const double* __restrict__ x,
                                          long long row, long long n, int l) {
  double acc = 0.0;
  if (row < n) {
    const int k1 = ptr[row + 1];
    for (int k = ptr[row] + l; k < k1; k += kLanesPerRow) acc = fma(values[k], x[index[k]], acc);
  }
#pragma unroll
  for (int off = kLanesPerRow / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  return acc;
}

// The row of the calling lane in a row-wise launch (grid: ceil(n / 8) blocks
// of 256) and whether the lane writes it.
struct RowLane {
  long long row;
  int l;
};

__device__ __forceinline__ RowLane row_lane() {
  const int lane = threadIdx.x & 63;
  RowLane r;
  r.row = (static_cast<long long>(blockIdx.x) * kWaves + (threadIdx.x >> 6)) * kGroups + lane / kLanesPerRow;
  r.l = lane % kLanesPerRow;
  return r;
}

__global__ __launch_bounds__(256) void diag_kernel(AmgLevel L) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= L.n) return;
  double d = 0.0;
  for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k)
    if (L.index[k] == i) d = L.values[k];
  L.dinv[i] = (d != 0.0 && std::isfinite(d)) ? 1.0 / d : 0.0;
}

// A fixed start vector in [-1, 1) from a hash of the row.
__global__ __launch_bounds__(256) void power_init_kernel(AmgLevel L, double* __restrict__ x) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i == 0) L.coef[kAmgScale] = 1.0;
  if (i >= L.n) return;
  unsigned long long h = static_cast<unsigned long long>(i) * 0x9E3779B97F4A7C15ull;
  h ^= h >> 29;
  x[i] = static_cast<double>(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// y = D^-1 A (x / ||x||), rows dealt round-robin over the nparts blocks;
// partial[block] = the block's sum of y^2 (groups, then wavefronts in order).
__global__ __launch_bounds__(256) void power_step_kernel(AmgLevel L, const double* __restrict__ x,
                                                         double* __restrict__ y) {
  __shared__ double ws[kRowsPerBlock];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane / kLanesPerRow, l = lane % kLanesPerRow;
  const double scale = L.coef[kAmgScale];
  const long long stride = static_cast<long long>(gridDim.x) * kRowsPerBlock;
  double sq = 0.0;
  for (long long base = (static_cast<long long>(blockIdx.x) * kWaves + wave) * kGroups; base < L.n;
       base += stride) {
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
    L.partial[blockIdx.x] = sum;
  }
}

// ||y|| (the estimate, x having been normalised) and 1 / ||y||.
__global__ __launch_bounds__(256) void power_finish_kernel(AmgLevel L) {
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
__global__ void coef_kernel(AmgLevel L) {
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
// of C, in windows of kCap columns. X's entries of the row are read 64 at a
// time (one per lane, with the bounds of their rows of Y) and taken in order;
// the first 64 entries of the next entry's row of Y are loaded while the
// current one is added. A wavefront's LDS accesses run in program order, so
// an entry's additions see the previous entry's without a barrier.
// smoothed: C is the smoothed prolongator, written as T - omega D^-1 (A T)
// (X = A, Y = T).
__global__ __launch_bounds__(64) void product_kernel(long long rows, Csr X, Csr Y, const int* __restrict__ cptr,
                                                     const int* __restrict__ cindex, double* __restrict__ cvalues,
                                                     int smoothed, AmgLevel L) {
  __shared__ int cols[kCap];
  __shared__ double acc[kCap];
  const long long row = blockIdx.x;
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
      const int t = b0 + lane;
      const bool in = t < x1;
      const int k = in ? X.index[t] : 0;
      const double a_l = in ? X.values[t] : 0.0;
      const int yb_l = in ? Y.ptr[k] : 0;
      const int ye_l = in ? Y.ptr[k + 1] : 0;
      const int cnt = x1 - b0 < 64 ? x1 - b0 : 64;
      int yb = __shfl(yb_l, 0, 64), ye = __shfl(ye_l, 0, 64);
      int c_next = yb + lane < ye ? Y.index[yb + lane] : -1;
      double v_next = yb + lane < ye ? Y.values[yb + lane] : 0.0;
      for (int j = 0; j < cnt; ++j) {
        const double a = __shfl(a_l, j, 64);
        const int cb = yb, ce = ye;
        const int c = c_next;
        const double v = v_next;
        if (j + 1 < cnt) {
          yb = __shfl(yb_l, j + 1, 64);
          ye = __shfl(ye_l, j + 1, 64);
          c_next = yb + lane < ye ? Y.index[yb + lane] : -1;
          v_next = yb + lane < ye ? Y.values[yb + lane] : 0.0;
        }
        if (c >= lo_col && c <= hi_col) {
          const int slot = find_slot(cols, mw, c);
          if (slot >= 0) acc[slot] = fma(a, v, acc[slot]);
        }
        for (int u = cb + 64 + lane; u < ce; u += 64) {  // rows of Y longer than 64
          const int cu = Y.index[u];
          if (cu < lo_col || cu > hi_col) continue;
          const int slot = find_slot(cols, mw, cu);
          if (slot >= 0) acc[slot] = fma(a, Y.values[u], acc[slot]);
        }
      }
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

__global__ __launch_bounds__(256) void gather_kernel(long long n, const int* __restrict__ perm,
                                                     const double* __restrict__ src, double* __restrict__ dst) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

// [A | I] -> [I | A^-1] by Gauss-Jordan with partial pivoting (the largest
// |entry| of the column, the lowest row among equals); one workgroup.
__global__ __launch_bounds__(1024) void dense_inverse_kernel(AmgLevel L) {
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

__global__ __launch_bounds__(256) void symmetric_part_kernel(long long nnz, const double* __restrict__ v,
                                                             const int* __restrict__ mirror,
                                                             double* __restrict__ out) {
  const long long k = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (k >= nnz) return;
  const int m = mirror[k];
  out[k] = m >= 0 ? 0.5 * (v[k] + v[m]) : 0.5 * v[k];
}

// --- the V-cycle ------------------------------------------------------------

// Pre-smoothing from x = 0, first half: r = D^-1 b, d0 = r / theta.
__global__ __launch_bounds__(256) void pre_first_kernel(AmgLevel L, const double* __restrict__ b) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= L.n) return;
  const double ri = L.dinv[i] * b[i];
  L.r[i] = ri;
  L.d0[i] = L.coef[kAmgThetaInv] * ri;
}

// Second half: r -= D^-1 A d0, d1 = c1 d0 + c2 r, x = d0 + d1.
__global__ __launch_bounds__(256) void pre_step_kernel(AmgLevel L, double* __restrict__ x) {
  const RowLane q = row_lane();
  const double v = row_dot(L.ptr, L.index, L.values, L.d0, q.row, L.n, q.l);
  if (q.l != 0 || q.row >= L.n) return;
  const long long i = q.row;
  const double ri = fma(-L.dinv[i], v, L.r[i]);
  const double d1 = fma(L.coef[kAmgC1], L.d0[i], L.coef[kAmgC2] * ri);
  L.r[i] = ri;
  x[i] = L.d0[i] + d1;
}

// r = b - A x.
__global__ __launch_bounds__(256) void residual_kernel(AmgLevel L, const double* __restrict__ b,
                                                       const double* __restrict__ x) {
  const RowLane q = row_lane();
  const double v = row_dot(L.ptr, L.index, L.values, x, q.row, L.n, q.l);
  if (q.l == 0 && q.row < L.n) L.r[q.row] = b[q.row] - v;
}

// The next level's b = P^T r.
__global__ __launch_bounds__(256) void restrict_kernel(AmgLevel L, double* __restrict__ nb) {
  const RowLane q = row_lane();
  const double v = row_dot(L.pt_ptr, L.pt_index, L.pt_values, L.r, q.row, L.nc, q.l);
  if (q.l == 0 && q.row < L.nc) nb[q.row] = v;
}

// x += P x_next.
__global__ __launch_bounds__(256) void prolong_kernel(AmgLevel L, double* __restrict__ x,
                                                      const double* __restrict__ nx) {
  const RowLane q = row_lane();
  const double v = row_dot(L.p_ptr, L.p_index, L.p_values, nx, q.row, L.n, q.l);
  if (q.l == 0 && q.row < L.n) x[q.row] += v;
}

// Post-smoothing from x, first half: r = D^-1 (b - A x), d0 = r / theta.
__global__ __launch_bounds__(256) void post_first_kernel(AmgLevel L, const double* __restrict__ b,
                                                         const double* __restrict__ x) {
  const RowLane q = row_lane();
  const double v = row_dot(L.ptr, L.index, L.values, x, q.row, L.n, q.l);
  if (q.l != 0 || q.row >= L.n) return;
  const double ri = L.dinv[q.row] * (b[q.row] - v);
  L.r[q.row] = ri;
  L.d0[q.row] = L.coef[kAmgThetaInv] * ri;
}

// Second half: r -= D^-1 A d0, d1 = c1 d0 + c2 r, x += d0 + d1.
__global__ __launch_bounds__(256) void post_step_kernel(AmgLevel L, double* __restrict__ x) {
  const RowLane q = row_lane();
  const double v = row_dot(L.ptr, L.index, L.values, L.d0, q.row, L.n, q.l);
  if (q.l != 0 || q.row >= L.n) return;
  const long long i = q.row;
  const double ri = fma(-L.dinv[i], v, L.r[i]);
  const double d1 = fma(L.coef[kAmgC1], L.d0[i], L.coef[kAmgC2] * ri);
  L.r[i] = ri;
  x[i] += L.d0[i] + d1;
}

// The coarsest level: x = A^-1 b from the dense inverse (each row's sum in
// column order).
__global__ __launch_bounds__(256) void dense_apply_kernel(AmgLevel L, const double* __restrict__ b,
                                                          double* __restrict__ x) {
  const int n = static_cast<int>(L.n);
  const long long w = 2ll * n;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double* row = L.dense + i * w + n;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc = fma(row[j], b[j], acc);
    x[i] = acc;
  }
}

int vcycle(const AmgLevel* lv, int l, int count, const double* b, double* x, hipStream_t s) {
  const AmgLevel& L = lv[l];
  if (L.n == 0) return 0;
  if (l == count - 1) {
    hipLaunchKernelGGL(dense_apply_kernel, dim3(grid_of(L.n, 256)), dim3(256), 0, s, L, b, x);
    return 0;
  }
  const AmgLevel& N = lv[l + 1];
  const dim3 rows(grid_of(L.n, kRowsPerBlock)), blk(256);
  hipLaunchKernelGGL(pre_first_kernel, dim3(grid_of(L.n, 256)), blk, 0, s, L, b);
  hipLaunchKernelGGL(pre_step_kernel, rows, blk, 0, s, L, x);
  hipLaunchKernelGGL(residual_kernel, rows, blk, 0, s, L, b, static_cast<const double*>(x));
  if (L.nc > 0) {
    hipLaunchKernelGGL(restrict_kernel, dim3(grid_of(L.nc, kRowsPerBlock)), blk, 0, s, L, N.b);
    vcycle(lv, l + 1, count, N.b, N.x, s);
    hipLaunchKernelGGL(prolong_kernel, rows, blk, 0, s, L, x, static_cast<const double*>(N.x));
  }
  hipLaunchKernelGGL(post_first_kernel, rows, blk, 0, s, L, b, static_cast<const double*>(x));
  hipLaunchKernelGGL(post_step_kernel, rows, blk, 0, s, L, x);
  return 0;
}

}  // namespace

int amg_product_blocks(long long n) {
  const long long b = (n + 4 * kRowsPerBlock - 1) / (4 * kRowsPerBlock);
  return static_cast<int>(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

int launch_amg_level_setup(const AmgLevel& L, const AmgLevel& next, hipStream_t s) {
  (void)hipGetLastError();
  if (L.n == 0) return 0;
  const dim3 blk(256);
  hipLaunchKernelGGL(diag_kernel, dim3(grid_of(L.n, 256)), blk, 0, s, L);
  hipLaunchKernelGGL(power_init_kernel, dim3(grid_of(L.n, 256)), blk, 0, s, L, L.d0);
  double* x = L.d0;
  double* y = L.d1;
  for (int k = 0; k < kAmgPowerSteps; ++k) {
    hipLaunchKernelGGL(power_step_kernel, dim3(L.nparts), blk, 0, s, L, static_cast<const double*>(x), y);
    hipLaunchKernelGGL(power_finish_kernel, dim3(1), blk, 0, s, L);
    double* t = x;
    x = y;
    y = t;
  }
  hipLaunchKernelGGL(coef_kernel, dim3(1), dim3(64), 0, s, L);
  const Csr A{L.ptr, L.index, L.values};
  hipLaunchKernelGGL(product_kernel, dim3(static_cast<unsigned int>(L.n)), dim3(64), 0, s, L.n, A,
                     Csr{L.tp_ptr, L.tp_index, L.tp_values}, L.p_ptr, L.p_index, L.p_values, 1, L);
  hipLaunchKernelGGL(product_kernel, dim3(static_cast<unsigned int>(L.n)), dim3(64), 0, s, L.n, A,
                     Csr{L.p_ptr, L.p_index, L.p_values}, L.ap_ptr, L.ap_index, L.ap_values, 0, L);
  if (L.nc > 0) {
    const long long pt_nnz = L.pt_nnz;
    if (pt_nnz > 0)
      hipLaunchKernelGGL(gather_kernel, dim3(grid_of(pt_nnz, 256)), blk, 0, s, pt_nnz, L.pt_perm,
                         static_cast<const double*>(L.p_values), L.pt_values);
    hipLaunchKernelGGL(product_kernel, dim3(static_cast<unsigned int>(L.nc)), dim3(64), 0, s, L.nc,
                       Csr{L.pt_ptr, L.pt_index, L.pt_values}, Csr{L.ap_ptr, L.ap_index, L.ap_values}, next.ptr,
                       next.index, const_cast<double*>(next.values), 0, L);
  }
  return static_cast<int>(hipGetLastError());
}

int launch_amg_dense_setup(const AmgLevel& L, hipStream_t s) {
  (void)hipGetLastError();
  if (L.n == 0) return 0;
  if (L.n > kAmgDenseMax) return 1;
  hipLaunchKernelGGL(dense_inverse_kernel, dim3(1), dim3(1024), 0, s, L);
  return static_cast<int>(hipGetLastError());
}

int launch_amg_symmetric_part(long long nnz, const double* values, const int* mirror, double* out, hipStream_t s) {
  (void)hipGetLastError();
  if (nnz > 0)
    hipLaunchKernelGGL(symmetric_part_kernel, dim3(grid_of(nnz, 256)), dim3(256), 0, s, nnz, values, mirror, out);
  return static_cast<int>(hipGetLastError());
}

int launch_amg_vcycle(const AmgLevel* levels, int count, const double* b, double* x, hipStream_t s) {
  (void)hipGetLastError();
  vcycle(levels, 0, count, b, x, s);
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
