// Conjugate gradients on the free-free block of a symmetric tangent (and
// BiCGStab on a non-symmetric one): the device solve of K_ff x = b beside the
// reference's direct factorisations (PartitionedFactor::factorize and solve,
// src/fem/Forming.cpp:104-172). Every launch below acts on the problems whose
// active flag is set (all of them in a plain solve).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace sbk {

// Device scalars of the iteration (doubles).
enum CgScalar : int {
  kCgRz = 0,      // r . z
  kCgAlpha = 1,
  kCgBeta = 2,
  kCgRr = 3,      // r . r
  kCgBb = 4,      // b . b
  kCgBreak = 5,   // 1: p . K p <= 0 (not positive definite) or a non-finite step
  kCgScalars = 8
};

struct CgArgs {
  const int* ptr;       // n + 1 row pointers of the block
  const int* index;     // the block's column indices
  const double* values; // the block's values (rows; K_ff symmetric)
  long long n;
  int nparts;           // vector blocks: fixed contiguous ranges, one partial sum each
  int nrows;            // product blocks (4 wavefronts each, rows dealt round-robin)
  double* dinv;         // n: 1 / K_ii
  double* r;            // n
  double* z;            // n
  double* p;            // n
  double* q;            // n: K p
  double* pq;           // nrows: partial sums of p . q
  double* rz;           // 2 nparts: partial sums of r . z, by iteration parity
  double* rr;           // nparts: partial sums of r . r
  double* partial;      // 2 nparts (start and residual)
  double* scalars;      // kCgScalars
  // BiCGStab (a non-symmetric block) besides: r^ (the shadow residual), v =
  // K y, y = D^-1 p, s, t = K z (z = D^-1 s in z, the product's input), and
  // partial sums of r^ . v (nrows), t . s and t . t (2 nrows) and of r^ . r
  // by iteration parity (2 nparts)
  double *rhat, *v, *y, *s, *t;
  double *pv, *pts, *rho;
  // Segments: a block-diagonal block of nseg independent problems (rows
  // [seg_row[p], seg_row[p + 1])), each iterated with its own scalars
  // (scalars + p kCgScalars) and its own partial sums, cut into the vector
  // and product blocks a plain solve of its rows would use (vfirst, pfirst:
  // each problem's first block; vseg, pseg: each block's problem), so every
  // problem's sums run in the order of its own plain solve. active[p] = 0
  // leaves problem p as it is; cycles[p] is its BiCGStab cycle start. A plain
  // solve is one segment with these arrays null.
  int nseg;
  const long long* seg_row;
  const int *vseg, *vfirst, *pseg, *pfirst;
  int* active;
  long long* cycles;
  // An external preconditioner in place of the diagonal: z = M^-1 r (and
  // BiCGStab's y = M^-1 p, z = M^-1 s) by precond(precond_ctx, in, out,
  // stream) between the kernels, which then leave those vectors to it
  // (external = 1). A plain solve only.
  int external;
  int (*precond)(const void* ctx, const double* in, double* out, hipStream_t s);
  const void* precond_ctx;
};

// The blocks of the problems of a segmented solve (host side).
struct CgSegments {
  std::vector<long long> rows;  // nseg + 1 row offsets
  std::vector<int> vfirst, pfirst, vseg, pseg;
};

// Product blocks for a block of n rows, and the doubles of work a solve of n
// unknowns takes (vectors, partial sums, scalars).
int cg_vector_blocks(long long n);
int cg_product_blocks(long long n);
long long cg_work_doubles(long long n);
// The work's pieces laid out from `work` (a plain solve: one segment).
void cg_layout(CgArgs& a, double* work);
// The blocks of problems whose rows start at rows[p] (rows: nseg + 1
// ascending offsets, the last the block's row count).
CgSegments cg_segments(const std::vector<long long>& rows);
// Doubles of work and bytes of device metadata a segmented solve takes.
long long cg_segment_work_doubles(const CgSegments& g);
long long cg_segment_meta_bytes(const CgSegments& g);
// The segmented solve's pieces from `work` and `meta` (device), the metadata
// copied in on stream s (every problem active, cycles 0); waits for the
// copies. 0 on success.
int cg_segment_layout(CgArgs& a, const CgSegments& g, double* work, void* meta, hipStream_t s);
// flag[0] = 1 (device) when a row has a column outside its problem's rows.
int launch_segment_check(const CgArgs& a, int* flag, hipStream_t s);

// The diagonal's inverse (a missing or non-positive diagonal sets kCgBreak),
// x = 0, r = b, z = D^-1 r, p = z and the first r . z, r . r, b . b.
int launch_cg_start(const CgArgs& a, const double* b, double* x, hipStream_t s);
// Iterations [first, first + count): q = K p, alpha = rz / p.q, x += alpha p,
// r -= alpha q, z = D^-1 r, beta = r.z / rz, p = z + beta p; three launches
// each (the product with p . q, the update with r . z and r . r, the
// direction), every block deriving alpha and beta from the partial sums in
// the same fixed order.
int launch_cg_iterations(const CgArgs& a, double* x, long long first, int count, hipStream_t s);
// A restart from the residual in r (after launch_cg_residual): z = D^-1 r,
// p = z, with r . z where iteration `next` reads it, and r . z, r . r in the
// scalars.
int launch_cg_restart(const CgArgs& a, long long next, hipStream_t s);
// BiCGStab, preconditioned with the diagonal, for a non-symmetric block:
// x = 0, r = r^ = b and the first r^ . r, r . r, b . b (after the diagonal's
// inverse, as launch_cg_start).
int launch_bicg_start(const CgArgs& a, const double* b, double* x, hipStream_t s);
// Iterations [first, first + count) of the cycle that began at `cycle`
// (whose first iteration takes p = r): five launches each (the direction
// with y = D^-1 p, v = K y with r^ . v, the half step with z = D^-1 s, t = K z
// with t . s and t . t, the full step with r^ . r and r . r), every block
// deriving alpha, omega and beta itself from the partial sums; then r . r
// into the scalars.
int launch_bicg_iterations(const CgArgs& a, double* x, long long cycle, long long first, int count,
                           hipStream_t s);
// A new cycle from the residual in r (after launch_cg_residual): r^ = r,
// with r^ . r where iteration `next` reads it; clears a breakdown.
int launch_bicg_restart(const CgArgs& a, long long next, hipStream_t s);
// The true residual r = b - K x into r and its r . r into scalars[kCgRr].
int launch_cg_residual(const CgArgs& a, const double* b, const double* x, hipStream_t s);
// y = K x over the block's rows (fixed-order sums).
int launch_block_spmv(const int* ptr, const int* index, const double* values, long long n, const double* x,
                      double* y, hipStream_t s);

}  // namespace sbk
