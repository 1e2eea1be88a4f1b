// Penalty node-to-tool contact with Coulomb friction on gfx950, FP64.
//
// Reference: RigidTool::gap (src/fem/RigidTool.cpp:237-270) and
// ToolContact::evaluate (:398-555) for one tool over a list of slave nodes;
// add_node_block (src/fem/Forming.cpp:280-297).
//
// contact_node_kernel: one thread per slave node: gap, normal and its
//   derivative; the penalty force p_N n and its block kappa A n n^T - p_N dn/dx;
//   the friction return (the committed tangential force transported onto the
//   tangent plane, the elastic trial, stick or slip on mu p_N) with the exact
//   non-symmetric derivative or the symmetric approximation; the residual
//   -= force; per-node terms of the norms and resultants; the trial history.
// contact_reduce_kernel: one block, fixed-order tree: gross and round-off
//   norms, the resultant, counts, the largest penetration, symmetry.
// add_node_blocks_kernel: each active node's 3 x 3 block into its diagonal
//   block of the assembled values (row- or column-major).
// candidate_*_kernel: the increment's candidate slaves in slave order (flags,
//   block offsets, an in-block scan); anticipate_kernel: the penalty of the
//   nodes a Newton correction would drive into the tool.
#include "contact_kernels.hpp"

namespace sbk {

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ double norm3(const double a[3]) { return sqrt(dot3(a, a)); }

// RigidTool::gap (RigidTool.cpp:237-270) at x for the tool at c: the gap g,
// the normal n and its derivative dn; false when the node lies less than half
// the radius from a sphere's centre or a cylinder's axis (the tool is through
// the surface: the reference's SolverError).
__device__ __forceinline__ bool tool_gap(const ContactTool& T, const double x[3], const double c[3], double n[3],
                                         double dn[3][3], double& g) {
  if (T.shape == 0) {
    const double nn = norm3(T.normal);
#pragma unroll
    for (int k = 0; k < 3; ++k) n[k] = T.normal[k] / nn;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) dn[a][b] = 0.0;
    g = n[0] * (x[0] - c[0]) + n[1] * (x[1] - c[1]) + n[2] * (x[2] - c[2]);
    return true;
  }
  double r[3] = {x[0] - c[0], x[1] - c[1], x[2] - c[2]};
  double base[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (T.shape == 2) {
    const double an = norm3(T.axis);
    const double ax[3] = {T.axis[0] / an, T.axis[1] / an, T.axis[2] / an};
    const double ar = dot3(ax, r);
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = r[k] - ar * ax[k];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) base[a][b] = base[a][b] - ax[a] * ax[b];
  }
  const double d = norm3(r);
  if (!(d > 0.5 * T.radius)) return false;
#pragma unroll
  for (int k = 0; k < 3; ++k) n[k] = r[k] / d;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) dn[a][b] = (base[a][b] - n[a] * n[b]) / d;
  g = d - T.radius;
  return true;
}

__global__ __launch_bounds__(256) void contact_node_kernel(ContactArgs A) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= A.num_slaves) return;
  const ContactTool& T = A.tool;
  const long long node = A.nodes[i];
  double x[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) x[k] = A.coords[node * 3 + k] + A.u[node * 3 + k];
  A.active[i] = 0;
  A.trial_present[i] = 0;
  A.node_terms[i * 8 + 0] = 0.0;  // sum |f_k|^2
  A.node_terms[i * 8 + 1] = 0.0;  // position scale
  A.node_terms[i * 8 + 2] = 0.0;  // -f_x, -f_y, -f_z (resultant share)
  A.node_terms[i * 8 + 3] = 0.0;
  A.node_terms[i * 8 + 4] = 0.0;
  A.node_terms[i * 8 + 5] = 0.0;  // penetration
  A.node_terms[i * 8 + 6] = 0.0;  // slipping
  A.node_terms[i * 8 + 7] = 0.0;  // asymmetric block
  A.status[i] = 0;
  double n[3], dn[3][3];
  double g;
  const double* c = T.centre;
  if (!tool_gap(T, x, c, n, dn, g)) {
    A.status[i] = 1;  // through the tool: SolverError
    return;
  }
  if (!(g < 0.0)) return;
  const double ka = A.kappa_area[i];
  const double pn = -ka * g;
  double force[3] = {pn * n[0], pn * n[1], pn * n[2]};
  double kn[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) kn[a][b] = ka * (n[a] * n[b]) - pn * dn[a][b];
  bool slipping = false;
  const double rel[3] = {x[0] - c[0], x[1] - c[1], x[2] - c[2]};
  const double mu = T.friction;
  if (mu > 0.0) {
    const double kt = T.tangential_ratio * ka;
    double f0[3] = {0.0, 0.0, 0.0}, q[3];
    if (A.hist_present && A.hist_present[i]) {
#pragma unroll
      for (int k = 0; k < 3; ++k) { f0[k] = A.hist_force[i * 3 + k]; q[k] = A.hist_relative[i * 3 + k]; }
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) q[k] = (A.coords[node * 3 + k] + A.start[node * 3 + k]) - T.centre0[k];
    }
    const double s0 = norm3(f0);
    const double fn = dot3(f0, n);
    double ft_[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) ft_[k] = f0[k] - fn * n[k];
    const double ftn = norm3(ft_);
    double f_hat[3] = {0.0, 0.0, 0.0}, d_hat[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (s0 > 0.0 && ftn > 1.0e-12 * s0) {
      double e[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) { e[k] = ft_[k] / ftn; f_hat[k] = s0 * e[k]; }
      double f0dn[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) f0dn[b] = f0[0] * dn[0][b] + f0[1] * dn[1][b] + f0[2] * dn[2][b];
      double dt[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) dt[a][b] = -(n[a] * f0dn[b] + fn * dn[a][b]);
      const double sc = s0 / ftn;
      double pe[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) pe[a][b] = sc * ((a == b ? 1.0 : 0.0) - e[a] * e[b]);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) d_hat[a][b] = pe[a][0] * dt[0][b] + pe[a][1] * dt[1][b] + pe[a][2] * dt[2][b];
    }
    const double w[3] = {rel[0] - q[0], rel[1] - q[1], rel[2] - q[2]};
    double p[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) p[a][b] = (a == b ? 1.0 : 0.0) - n[a] * n[b];
    double delta[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) delta[a] = p[a][0] * w[0] + p[a][1] * w[1] + p[a][2] * w[2];
    const double nw = dot3(n, w);
    double wdn[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) wdn[b] = w[0] * dn[0][b] + w[1] * dn[1][b] + w[2] * dn[2][b];
    double trial[3], dtr[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) trial[a] = f_hat[a] - kt * delta[a];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) dtr[a][b] = d_hat[a][b] - kt * (p[a][b] - nw * dn[a][b] - n[a] * wdn[b]);
    const double tn = norm3(trial);
    const double limit = mu * pn;
    double ft[3], dft[3][3], ks[3][3];
    if (tn <= limit) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        ft[a] = trial[a];
#pragma unroll
        for (int b = 0; b < 3; ++b) { dft[a][b] = dtr[a][b]; ks[a][b] = kt * p[a][b]; }
      }
    } else {
      slipping = true;
      double e[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) { e[a] = trial[a] / tn; ft[a] = limit * e[a]; }
      const double mk = mu * ka, lt = limit / tn;
      double pe[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) pe[a][b] = lt * ((a == b ? 1.0 : 0.0) - e[a] * e[b]);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const double pd = pe[a][0] * dtr[0][b] + pe[a][1] * dtr[1][b] + pe[a][2] * dtr[2][b];
          dft[a][b] = -mk * (e[a] * n[b]) + pd;
          ks[a][b] = lt * kt * (p[a][b] - e[a] * e[b]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) force[a] = force[a] + ft[a];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) kn[a][b] = A.friction_symmetric ? kn[a][b] + ks[a][b] : kn[a][b] - dft[a][b];
    A.trial_present[i] = 1;
#pragma unroll
    for (int a = 0; a < 3; ++a) { A.trial_force[i * 3 + a] = ft[a]; A.trial_relative[i * 3 + a] = rel[a]; }
  }
  const double ps = ka * (norm3(x) + norm3(c));
  double gsq = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    A.residual[node * 3 + k] = A.residual[node * 3 + k] - force[k];
    gsq = gsq + fabs(force[k]) * fabs(force[k]);
  }
  A.active[i] = 1;
  A.node_terms[i * 8 + 0] = gsq;
  A.node_terms[i * 8 + 1] = 3.0 * ps * ps;
  A.node_terms[i * 8 + 2] = -force[0];
  A.node_terms[i * 8 + 3] = -force[1];
  A.node_terms[i * 8 + 4] = -force[2];
  A.node_terms[i * 8 + 5] = -g;
  A.node_terms[i * 8 + 6] = slipping ? 1.0 : 0.0;
  if (A.want_tangent) {
    if (A.friction_symmetric) {
      double sym[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) sym[a][b] = 0.5 * (kn[a][b] + kn[b][a]);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) kn[a][b] = sym[a][b];
    }
    double top = 0.0, asym = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        top = fmax(top, fabs(kn[a][b]));
        asym = fmax(asym, fabs(kn[a][b] - kn[b][a]));
      }
    A.node_terms[i * 8 + 7] = asym > 1.0e-12 * top ? 1.0 : 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) A.blocks[i * 9 + a * 3 + b] = kn[a][b];
  }
}

// One block of 256 threads; thread t sums slaves t, t + 256, ... then a tree.
__global__ __launch_bounds__(256) void contact_reduce_kernel(const double* __restrict__ terms, const int* __restrict__ status,
                                                             const int* __restrict__ active, long long ns,
                                                             double* __restrict__ out) {
  __shared__ double s[9][256];
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long long i = threadIdx.x; i < ns; i += 256) {
    acc[0] = acc[0] + terms[i * 8 + 0];
    acc[1] = acc[1] + terms[i * 8 + 1];
    acc[2] = acc[2] + terms[i * 8 + 2];
    acc[3] = acc[3] + terms[i * 8 + 3];
    acc[4] = acc[4] + terms[i * 8 + 4];
    acc[5] = fmax(acc[5], terms[i * 8 + 5]);
    acc[6] = acc[6] + terms[i * 8 + 6];
    acc[7] = acc[7] + terms[i * 8 + 7];
    acc[8] = acc[8] + (active[i] ? 1.0 : 0.0) + (status[i] ? 1.0e9 : 0.0);
  }
  for (int k = 0; k < 9; ++k) s[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 9; ++k)
        s[k][threadIdx.x] = (k == 5) ? fmax(s[k][threadIdx.x], s[k][threadIdx.x + w])
                                     : s[k][threadIdx.x] + s[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sqrt(s[0][0]);      // gross
    out[1] = sqrt(s[1][0]);      // round-off
    out[2] = s[2][0];            // resultant
    out[3] = s[3][0];
    out[4] = s[4][0];
    out[5] = s[5][0];            // max penetration
    out[6] = s[6][0];            // slipping nodes
    out[7] = s[7][0];            // asymmetric blocks
    const double a = s[8][0];
    out[8] = fmod(a, 1.0e9);     // active nodes
    out[9] = floor(a / 1.0e9);   // nodes through the tool
  }
}

__global__ __launch_bounds__(256) void add_node_blocks_kernel(const double* __restrict__ blocks,
                                                              const int* __restrict__ active,
                                                              const int* __restrict__ first_value,
                                                              const int* __restrict__ row_len, long long ns,
                                                              int column_major, double* __restrict__ values) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ns || !active[i]) return;
  const long long f = first_value[i];
  const long long rl = row_len[i];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const long long pos = column_major ? f + b * rl + a : f + a * rl + b;
      values[pos] = values[pos] + blocks[i * 9 + a * 3 + b];
    }
}

// ---------------------------------------------------------------------------
// Increment set-up (ToolContact::begin_increment, RigidTool.cpp:345-393) and
// the semismooth anticipation (ToolContact::anticipate, :557-603)
// ---------------------------------------------------------------------------
// RigidTool.cpp:44-50: the distance from p to the segment [a, b].
__device__ __forceinline__ double segment_distance(const double p[3], const double a[3], const double b[3]) {
  const double s[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double len2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
  double t = 0.0;
  if (len2 > 0.0)
    t = fmin(fmax(((p[0] - a[0]) * s[0] + (p[1] - a[1]) * s[1] + (p[2] - a[2]) * s[2]) / len2, 0.0), 1.0);
  const double r[3] = {p[0] - a[0] - t * s[0], p[1] - a[1] - t * s[1], p[2] - a[2] - t * s[2]};
  return sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

// A node at x is a candidate of the increment whose tool moves from centre0
// to centre: within radius + margin of that path (a cylinder: in the plane
// normal to its axis), or for a plane within margin of either position.
__device__ __forceinline__ bool is_candidate(const ContactTool& T, const double x[3], double margin) {
  const double* c0 = T.centre0;
  const double* c1 = T.centre;
  if (T.shape == 1) return segment_distance(x, c0, c1) <= T.radius + margin;
  if (T.shape == 2) {
    const double an = norm3(T.axis);
    const double a[3] = {T.axis[0] / an, T.axis[1] / an, T.axis[2] / an};
    const double d0[3] = {x[0] - c0[0], x[1] - c0[1], x[2] - c0[2]};
    const double d1[3] = {c1[0] - c0[0], c1[1] - c0[1], c1[2] - c0[2]};
    double px[3], pc[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      px[r] = 0.0;
      pc[r] = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double pk = (r == k ? 1.0 : 0.0) - a[r] * a[k];
        px[r] = px[r] + pk * d0[k];
        pc[r] = pc[r] + pk * d1[k];
      }
    }
    const double zero[3] = {0.0, 0.0, 0.0};
    return segment_distance(px, zero, pc) <= T.radius + margin;
  }
  const double nn = norm3(T.normal);
  const double nv[3] = {T.normal[0] / nn, T.normal[1] / nn, T.normal[2] / nn};
  const double g0 = nv[0] * (x[0] - c0[0]) + nv[1] * (x[1] - c0[1]) + nv[2] * (x[2] - c0[2]);
  const double g1 = nv[0] * (x[0] - c1[0]) + nv[1] * (x[1] - c1[1]) + nv[2] * (x[2] - c1[2]);
  return fmin(g0, g1) <= margin;
}

// Candidate flags and each block's count (blocks of 256 slaves).
__global__ __launch_bounds__(256) void candidate_flags_kernel(ContactTool T, const double* __restrict__ coords,
                                                              const double* __restrict__ u0, long long ns,
                                                              const int* __restrict__ nodes, double margin,
                                                              int* __restrict__ flags, int* __restrict__ block_counts) {
  __shared__ int cnt[256];
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  int f = 0;
  if (i < ns) {
    const long long node = nodes[i];
    double x[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) x[k] = coords[node * 3 + k] + u0[node * 3 + k];
    f = is_candidate(T, x, margin) ? 1 : 0;
    flags[i] = f;
  }
  cnt[threadIdx.x] = f;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) cnt[threadIdx.x] += cnt[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_counts[blockIdx.x] = cnt[0];
}

// Exclusive offsets of the block counts in place, and the total (one block:
// each thread scans a contiguous run of the counts).
__global__ __launch_bounds__(256) void candidate_offsets_kernel(int* __restrict__ block_counts, int nb,
                                                                int* __restrict__ count) {
  __shared__ int part[256];
  const int per = (nb + 255) / 256;
  const int b0 = static_cast<int>(threadIdx.x) * per;
  const int b1 = b0 + per < nb ? b0 + per : nb;
  int run = 0;
  for (int b = b0; b < b1; ++b) run += block_counts[b];
  part[threadIdx.x] = run;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < 256; ++k) {
      const int v = part[k];
      part[k] = acc;
      acc += v;
    }
    count[0] = acc;
  }
  __syncthreads();
  run = part[threadIdx.x];
  for (int b = b0; b < b1; ++b) {
    const int v = block_counts[b];
    block_counts[b] = run;
    run += v;
  }
}

// The candidates in slave order: each block places its flagged slaves after
// the earlier blocks' (an in-block exclusive scan of the flags).
__global__ __launch_bounds__(256) void candidate_scatter_kernel(const int* __restrict__ flags, long long ns,
                                                                const int* __restrict__ block_offsets,
                                                                int* __restrict__ candidates) {
  __shared__ int scan[256];
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  const int f = i < ns ? flags[i] : 0;
  scan[threadIdx.x] = f;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int add = static_cast<int>(threadIdx.x) >= off ? scan[threadIdx.x - off] : 0;
    __syncthreads();
    scan[threadIdx.x] += add;
    __syncthreads();
  }
  if (f) candidates[block_offsets[blockIdx.x] + scan[threadIdx.x] - 1] = static_cast<int>(i);
}

// One thread per listed slave: a node outside the tool at u that the
// correction du drives into it (or through it) gets residual += kappa A g n
// and the block kappa A n n^T; the others get active = 0 and a zero block.
__global__ __launch_bounds__(256) void anticipate_kernel(ContactTool T, const double* __restrict__ coords,
                                                         const double* __restrict__ u, const double* __restrict__ du,
                                                         long long ns, const int* __restrict__ nodes,
                                                         const double* __restrict__ kappa_area,
                                                         double* __restrict__ residual, double* __restrict__ blocks,
                                                         int* __restrict__ active) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  active[i] = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) blocks[i * 9 + k] = 0.0;
  const long long node = nodes[i];
  double x[3], xn[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x[k] = coords[node * 3 + k] + u[node * 3 + k];
    xn[k] = x[k] + du[node * 3 + k];
  }
  double n[3], dn[3][3], g;
  if (!tool_gap(T, x, T.centre, n, dn, g)) return;  // deep in the tool
  if (!(g > 0.0)) return;                           // already active: in the tangent
  double n2[3], g_next;
  if (!tool_gap(T, xn, T.centre, n2, dn, g_next)) g_next = -__builtin_inf();  // through the tool
  if (!(g_next < 0.0)) return;
  const double ka = kappa_area[i];
#pragma unroll
  for (int k = 0; k < 3; ++k) residual[node * 3 + k] = residual[node * 3 + k] + ka * g * n[k];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) blocks[i * 9 + 3 * a + b] = ka * (n[a] * n[b]);
  active[i] = 1;
}

// count[0] = sum of flags[0, n) (one block, fixed order).
__global__ __launch_bounds__(256) void count_flags_kernel(const int* __restrict__ flags, long long n,
                                                          int* __restrict__ count) {
  __shared__ int s[256];
  int acc = 0;
  for (long long i = threadIdx.x; i < n; i += 256) acc += flags[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) count[0] = s[0];
}

int launch_contact_candidates(const ContactTool& t, const double* coords, const double* u0, long long ns,
                              const int* nodes, double margin, int* candidates, int* count, int* work,
                              hipStream_t s) {
  (void)hipGetLastError();
  const long long nb = (ns + 255) / 256;
  int* flags = work;
  int* block_counts = work + ns;
  if (nb > 0)
    hipLaunchKernelGGL(candidate_flags_kernel, dim3(static_cast<unsigned int>(nb)), dim3(256), 0, s, t, coords, u0,
                       ns, nodes, margin, flags, block_counts);
  hipLaunchKernelGGL(candidate_offsets_kernel, dim3(1), dim3(256), 0, s, block_counts, static_cast<int>(nb), count);
  if (nb > 0)
    hipLaunchKernelGGL(candidate_scatter_kernel, dim3(static_cast<unsigned int>(nb)), dim3(256), 0, s, flags, ns,
                       block_counts, candidates);
  return static_cast<int>(hipGetLastError());
}

int launch_contact_anticipate(const ContactTool& t, const double* coords, const double* u, const double* du,
                              long long ns, const int* nodes, const double* kappa_area, double* residual,
                              double* blocks, int* active, int* added, hipStream_t s) {
  (void)hipGetLastError();
  if (ns > 0)
    hipLaunchKernelGGL(anticipate_kernel, dim3(static_cast<unsigned int>((ns + 255) / 256)), dim3(256), 0, s, t,
                       coords, u, du, ns, nodes, kappa_area, residual, blocks, active);
  hipLaunchKernelGGL(count_flags_kernel, dim3(1), dim3(256), 0, s, active, ns, added);
  return static_cast<int>(hipGetLastError());
}

int launch_contact(const ContactArgs& a, double* scalars, hipStream_t s) {
  (void)hipGetLastError();
  if (a.num_slaves > 0) {
    const unsigned int grid = static_cast<unsigned int>((a.num_slaves + 255) / 256);
    hipLaunchKernelGGL(contact_node_kernel, dim3(grid), dim3(256), 0, s, a);
  }
  hipLaunchKernelGGL(contact_reduce_kernel, dim3(1), dim3(256), 0, s, a.node_terms, a.status, a.active,
                     a.num_slaves, scalars);
  return static_cast<int>(hipGetLastError());
}

int launch_add_node_blocks(const double* blocks, const int* active, const int* first_value, const int* row_len,
                           long long ns, int column_major, double* values, hipStream_t s) {
  (void)hipGetLastError();
  if (ns == 0) return 0;
  const unsigned int grid = static_cast<unsigned int>((ns + 255) / 256);
  hipLaunchKernelGGL(add_node_blocks_kernel, dim3(grid), dim3(256), 0, s, blocks, active, first_value, row_len, ns,
                     column_major, values);
  return static_cast<int>(hipGetLastError());
}

}  // namespace sbk
