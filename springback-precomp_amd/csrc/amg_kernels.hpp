// Smoothed-aggregation algebraic multigrid for the free-free block: the
// numeric set-up of a hierarchy whose patterns the host built (amg.cpp) and
// its V-cycle, the preconditioner of the block solve where the reference
// factorises the block directly (PartitionedFactor::factorize and solve,
// src/fem/Forming.cpp:104-172).
//
// Level l holds A_l (CSR), the tentative prolongator T_l (each row: its
// aggregate's orthonormalised near-nullspace), the smoothed prolongator
// P_l = T_l - omega_l D_l^-1 A_l T_l, its transpose and A_l P_l; the next
// level's matrix is A_{l+1} = P_l^T (A_l P_l). The products are row-wise with
// a fixed order of every sum (one wavefront per row, the row's columns and
// sums in LDS), so a set-up and a V-cycle repeat bit for bit. The smoother is
// Chebyshev of degree 2 in D^-1 A on [lmax / 30, lmax], lmax = 1.2 times a
// 25-step power estimate of the spectral radius of D^-1 A; the coarsest
// level is solved by its dense inverse (Gauss-Jordan, partial pivoting).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace sbk {

enum AmgCoef : int {
  kAmgRho = 0,       // power estimate of the spectral radius of D^-1 A
  kAmgThetaInv = 1,  // Chebyshev: 1 / theta
  kAmgC1 = 2,        // Chebyshev: d1 = c1 d0 + c2 r1
  kAmgC2 = 3,
  kAmgOmega = 4,     // prolongator smoothing: 4 / (3 rho)
  kAmgScale = 5,     // power iteration: 1 / ||x||
  kAmgSingular = 6,  // coarsest: 1 when the dense inverse met a zero pivot
  kAmgCoefs = 8
};

constexpr int kAmgPowerSteps = 25;
constexpr double kAmgEigenBoost = 1.2;
constexpr double kAmgEigenRatio = 30.0;
constexpr int kAmgDenseMax = 1024;  // rows of the coarsest level at most

struct AmgLevel {
  long long n = 0;  // rows
  const int* ptr = nullptr;  // A_l
  const int* index = nullptr;
  const double* values = nullptr;
  double* dinv = nullptr;     // n: 1 / a_ii (0 for a zero or non-finite diagonal)
  double* coef = nullptr;     // kAmgCoefs
  double* partial = nullptr;  // power iteration: a partial sum per product block
  int nparts = 0;             // product blocks over the level's rows
  // to the next level (nc = 0: the coarsest)
  long long nc = 0;
  const int* tp_ptr = nullptr;  // tentative prolongator T (n rows)
  const int* tp_index = nullptr;
  const double* tp_values = nullptr;
  const int* p_own = nullptr;  // n: where the row's own aggregate starts in P's row
  const int* p_ptr = nullptr;  // smoothed prolongator P (n rows)
  const int* p_index = nullptr;
  double* p_values = nullptr;
  const int* pt_ptr = nullptr;  // P^T (nc rows) and, per entry, its position in P
  const int* pt_index = nullptr;
  const int* pt_perm = nullptr;
  double* pt_values = nullptr;
  long long pt_nnz = 0;
  const int* ap_ptr = nullptr;  // A P (n rows)
  const int* ap_index = nullptr;
  double* ap_values = nullptr;
  // the V-cycle's vectors (level 0: b and x are the caller's)
  double* b = nullptr;
  double* x = nullptr;
  double* r = nullptr;
  double* d0 = nullptr;
  double* d1 = nullptr;
  // the coarsest level: [A | I] reduced to [I | A^-1], n x 2n row-major
  double* dense = nullptr;
};

// One problem's part of a grouped launch (device memory). A launch covers
// one level of each of a set of hierarchies: entry k takes the grid's blocks
// [first, next entry's first) with the decomposition a launch of that level
// alone would use; a table ends with an entry whose first is the launch's
// block count.
struct AmgGroupEntry {
  AmgLevel L;                   // the problem's level (a copy)
  const AmgLevel* N = nullptr;  // its next level (device; null at its coarsest)
  long long offset = -1;        // level 0: its rows' offset in the caller's vectors; -1 above
  long long first = 0;          // its first block
  int problem = 0;              // its segment (the active flags)
  int level = 0;
  // the symmetric-part table only: (A + A^T) / 2 of count block values
  const double* src = nullptr;
  const int* mirror = nullptr;
  double* dst = nullptr;
  long long count = 0;
};

struct AmgGroup {
  const AmgGroupEntry* e;  // count + 1 entries
  const int* map;          // per block of the launch: its entry
  const int* active;       // per problem: 0 rests (its blocks do nothing); null: all work
};

// A group table on the device.
struct AmgTable {
  const AmgGroupEntry* entries = nullptr;
  const int* map = nullptr;
  int count = 0;
  long long blocks = 0;
};

// The blocks an entry of each kind of table takes for its level.
enum AmgTableKind : int {
  kAmgTableElem = 0,        // a thread per row (at most 2048 blocks of 256, the rows dealt round-robin)
  kAmgTableRows,            // 32 lanes per row (at most 2048 blocks of 8 rows, dealt round-robin)
  kAmgTableRestrict,        // the same over the coarse rows
  kAmgTablePower,           // the level's nparts blocks
  kAmgTableOne,             // one block
  kAmgTableProduct,         // a block per row
  kAmgTableCoarseProduct,   // a block per coarse row
  kAmgTableGather,          // a thread per entry of P^T, as kAmgTableElem
  kAmgTableSymmetric        // a thread per entry of level 0, as kAmgTableElem
};
long long amg_table_blocks(AmgTableKind kind, const AmgLevel& L, long long nnz0);

// The tables of level l: the problems for which l is not the coarsest level
// (smoothing, restriction, set-up) and those for which it is (dense).
struct AmgLevelTables {
  AmgTable elem, rows, restrict_, power, one, product, coarse_product, gather;
  AmgTable dense_one, dense;
};

struct AmgTables {
  std::vector<AmgLevelTables> levels;
  AmgTable symmetric;  // level 0 of every problem
};

// Product blocks of a level of n rows (the power iteration's partial sums).
int amg_product_blocks(long long n);
// The numeric set-up of every hierarchy of the tables from their level 0
// values (the symmetric part first when symmetrise): per level, the
// diagonal, the power estimate and the coefficients, P, A P, P^T and the
// next level's values; the dense inverse of each coarsest level
// (kAmgSingular on a zero pivot).
int launch_amg_setup(const AmgTables& T, bool symmetrise, hipStream_t s);
// x = M^-1 b: one V-cycle of every hierarchy whose problem is active (null:
// all), level 0 reading b and writing x at each problem's offset.
int launch_amg_vcycle(const AmgTables& T, const int* active, const double* b, double* x, hipStream_t s);

}  // namespace sbk
