/* sparlab_gpu.h - the C ABI of the MI355X element-and-assembly path.
 *
 * This library replaces, behind plain pointers and sizes, the element loop of
 * sparlab_form's Newton iteration:
 *   - the incompatible-mode Hex8 element evaluation
 *       detail::incompatible_elastoplastic_element
 *       (reference src/fem/IncompatibleModes.cpp:577-600, dispatched from
 *       src/fem/Elastoplastic.cpp:226-235, declared at
 *       include/sparlab/fem/Elastoplastic.hpp:196-200), for a whole batch of
 *       elements at once;
 *   - the element loop and gather of NonlinearSystem::evaluate
 *       (src/fem/NonlinearSystem.hpp:192-268) and the sparse assembly of
 *       Assembler::build_pattern / scatter / assemble_elementwise
 *       (src/fem/Assembler.cpp:86-215, 271-292);
 *   - the penalty node-to-tool contact of ToolContact::evaluate
 *       (src/fem/RigidTool.cpp:398-555) and add_node_block
 *       (src/fem/Forming.cpp:280-297).
 *
 * Conventions (the reference's, include/sparlab/core/Types.hpp:31-48):
 * FP64 throughout; node-major DOFs {ux, uy, uz} (DOF = 3 node + k); VTK
 * hexahedron node order; Voigt order {11, 22, 33, 12, 23, 31} with engineering
 * shear strains and tensorial stresses.
 *
 * Device pointers: every pointer argument of a *_device entry point is HIP
 * device memory on the current device. The caller owns all memory.
 *
 * Errors: every entry point returns 0 on success or a negative
 * SPARLAB_E_* code; sparlab_last_error() returns the message of the calling
 * thread's last failure. Per-element failures are reported per element in a
 * status array (SPARLAB_ELEMENT_* codes), never as a return code of the call,
 * as the reference records a failure per element and raises one SolverError
 * after its element loop (src/fem/NonlinearSystem.hpp:231-238).
 */
#ifndef SPARLAB_GPU_H
#define SPARLAB_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes of the calls. */
#define SPARLAB_OK 0
#define SPARLAB_E_ARGUMENT (-1)   /* ConfigError / ModelError in the reference */
#define SPARLAB_E_DEVICE (-2)     /* a HIP call failed */
#define SPARLAB_E_UNSUPPORTED (-3)
#define SPARLAB_E_SOLVER (-4)     /* SolverError (an element or a contact node failed) */
#define SPARLAB_E_MESH (-5)       /* MeshError */

/* Per-element status codes (status arrays). */
#define SPARLAB_ELEMENT_OK 0
#define SPARLAB_ELEMENT_CENTRE_JACOBIAN 1   /* MeshError: det J <= 0 at the centre */
#define SPARLAB_ELEMENT_POINT_JACOBIAN 2    /* MeshError: det J <= 0 at a point */
#define SPARLAB_ELEMENT_INVERTED 3          /* SolverError: det F <= 0 */
#define SPARLAB_ELEMENT_DEGENERATE 4        /* SolverError: 1 + 2 mu <= 0 or non-finite E */
#define SPARLAB_ELEMENT_RETURN_FAILED 5     /* SolverError: the return did not converge */
#define SPARLAB_ELEMENT_SINGULAR_MODES 6    /* SolverError: K_aa singular */
#define SPARLAB_ELEMENT_NOT_CONVERGED 7     /* SolverError: 25 local iterations */
#define SPARLAB_ELEMENT_NO_DESCENT 8        /* SolverError: no step lowers |r_alpha| */
#define SPARLAB_ELEMENT_NONFINITE_TRIAL 9   /* SolverError: non-finite trial stress */
#define SPARLAB_ELEMENT_INTERNAL 10

/* Kinematics (include/sparlab/fem/Elastoplastic.hpp: Kinematics). */
#define SPARLAB_SMALL_STRAIN 0
#define SPARLAB_FINITE 1
#define SPARLAB_FINITE_LOGARITHMIC 2

/* Material: Material + PlasticityParameters
 * (include/sparlab/material/Material.hpp:148-257). yield_matrix is found by
 * sparlab_material_finalize (Material::set_plasticity). */
typedef struct sparlab_material {
  double youngs_modulus;              /* E [Pa] */
  double poisson_ratio;               /* nu */
  double yield_stress;                /* sigma_y0 [Pa]; 0: elastic */
  double hardening_modulus;           /* H, linear isotropic [Pa] */
  double kinematic_hardening_modulus; /* H_kin, Prager [Pa] */
  double saturation_stress;           /* Q, Voce [Pa] */
  double saturation_rate;             /* delta, Voce [-] */
  int32_t criterion;                  /* 0 von Mises, 1 Hill48 */
  int32_t num_backstresses;           /* Chaboche terms, 0..4 */
  double backstress_modulus[4];       /* C_i [Pa] */
  double backstress_recovery[4];      /* gamma_i [-] */
  int32_t kinematic_integration;      /* 0 exponential, 1 backward Euler */
  int32_t hill_calibration;           /* 0 r-values, 1 stress ratios, 2 coefficients */
  double hill_r[3];                   /* r0, r45, r90 */
  double hill_stress_ratio[3];        /* sigma45/sigma0, sigma90/sigma0, sigma_b/sigma0 */
  double hill_shear_lm[2];            /* L, M of the r-value and ratio calibrations */
  double hill_coefficients[6];        /* F, G, H, L, M, N (in for 2, out otherwise) */
  double rolling_direction[3];
  double sheet_normal[3];
  double yield_matrix[36];            /* out: P, tensorial Voigt, global axes */
} sparlab_material;

/* Integration rule: the n x n x n Gauss cube, or the n x n x m box rule
 * through thickness_axis (src/elements/Hex8Kernels.hpp:125-130). */
typedef struct sparlab_rule {
  int32_t stiffness_points;  /* n: 1..4 (2 by default) */
  int32_t thickness_points;  /* m: 0 (= n) or 1..7 */
  int32_t thickness_axis;    /* 0 xi, 1 eta, 2 zeta */
} sparlab_rule;

/* Validates the material and fills yield_matrix and hill_coefficients
 * (replaces Material::set_plasticity, src/material/Material.cpp:221-271 with
 * derive_hill :135-217). */
int sparlab_material_finalize(sparlab_material* material);

/* Points of the rule (elastoplastic_points, src/fem/Elastoplastic.cpp:222-224),
 * or a negative code. The element kernels take rules of up to 32 points (8-lane teams up to
 * 8 points, 32-lane teams above: the 2 x 2 x 5 and 2 x 2 x 7 box rules). */
int sparlab_rule_points(const sparlab_rule* rule);

/* Doubles of history per integration point in the library's layout: plastic
 * strain (6), kinematic history (6 per 6-vector: the radial return's back
 * stress, or the general return's backstresses), equivalent plastic strain,
 * loading flag. */
int sparlab_state_fields(const sparlab_material* material);

/* A batch of elements on a mesh (device pointers).
 * coords: 3 num_nodes; conn: 8 num_elements (int32, VTK order); u: 3 num_nodes.
 * state_in / state_out: fields x (num_elements x points), structure of
 *   arrays: field f of point q of element e at [f * num_elements * points +
 *   e * points + q] (the committed and the trial history,
 *   src/fem/NonlinearSystem.hpp:328).
 * alpha_in: committed mode parameters, 9 per element (may be null: zero);
 * alpha_out: 9 per element.
 * force: 24 per element; tangent: 576 per element, row-major (may be null for
 *   a residual-only evaluation); energy: per element; yielding, iterations,
 *   status: int32 per element; fail_value: per element (the refused Jacobian
 *   determinant for SPARLAB_ELEMENT_*_JACOBIAN).
 *   A failed element (status != 0) outputs zero force, tangent, energy and
 *   yielding, alpha_in as alpha_out and state_in as state_out (its trial is
 *   rejected, as after the reference's SolverError).
 * workspace: sparlab_element_workspace_doubles(num_elements, rule) doubles
 *   of device scratch, 16-byte aligned (each point's stress and tangent of its final local
 *   pass and each element's K_aa factors, passed from the local solve to the
 *   condensation); the reference keeps the same per-call data in a
 *   thread_local workspace (src/fem/IncompatibleModes.cpp:66-71). Calls that
 *   share a workspace must not overlap. */
typedef struct sparlab_element_batch {
  const double* coords;
  const int32_t* conn;
  const double* u;
  int64_t num_nodes;
  int64_t num_elements;
  const double* state_in;
  const double* alpha_in;
  double* state_out;
  double* alpha_out;
  double* force;
  double* tangent;
  double* energy;
  int32_t* yielding;
  int32_t* iterations;
  int32_t* status;
  double* fail_value;
  double* workspace;
  int32_t* local_path;  /* optional (may be null): per element, the SPARLAB_LOCAL_* phases of
                           its local iteration (a diagnostic; the reference keeps none) */
} sparlab_element_batch;

/* Bits of sparlab_element_batch.local_path: the fallbacks of the local
 * iteration of the modes (src/fem/IncompatibleModes.cpp:398-546) an element
 * went through. */
#define SPARLAB_LOCAL_REGULA_FALSI 1   /* the line search's regula falsi on d . r_alpha */
#define SPARLAB_LOCAL_NORM_SEARCH 2    /* search_norm: halvings on |r_alpha| */
#define SPARLAB_LOCAL_INDEFINITE 4     /* sym(K_aa) not positive definite */
#define SPARLAB_LOCAL_SHIFTED 8        /* the shifted direction (K_aa + mu I) d = -r_alpha */
#define SPARLAB_LOCAL_MARQUARDT 16     /* Levenberg-Marquardt */
#define SPARLAB_LOCAL_HALVED 32        /* a refused trial halved by the line search */

/* Doubles of workspace an element batch needs (86 per integration point, 109
 * per element and 17 more), or a negative code for a rule the kernels do not
 * take. */
int64_t sparlab_element_workspace_doubles(int64_t num_elements, const sparlab_rule* rule);

/* Evaluates every element of the batch (replaces the calls of
 * incompatible_elastoplastic_element in the element loop of
 * src/fem/NonlinearSystem.hpp:208-237; want_tangent as there). Asynchronous
 * on `stream` (a hipStream_t; null: the default stream). Returns 0 once the
 * work is queued; per-element outcomes are in status. */
int sparlab_elements_evaluate_device(const sparlab_material* material, const sparlab_rule* rule,
                                     int kinematics, int want_tangent,
                                     const sparlab_element_batch* batch, void* stream);

/* The same evaluation in its two stages (stages: 1 the local iteration of
 * the modes, src/fem/IncompatibleModes.cpp:320-546; 2 the coupling blocks and
 * the condensation, :549-561, which needs stage 1 of the same batch to have
 * run; 3 both), so that each stage can be timed or overlapped on its own.
 * The local iteration takes the elements in the order of their expected work
 * (a trial of every point at the committed modes, most work first), which
 * changes no result; adding SPARLAB_STAGE_INDEX_ORDER to `stages` takes them
 * in index order instead. */
#define SPARLAB_STAGE_INDEX_ORDER 4
int sparlab_elements_evaluate_stage_device(const sparlab_material* material, const sparlab_rule* rule,
                                           int kinematics, int want_tangent,
                                           const sparlab_element_batch* batch, int stages, void* stream);

/* sparlab_elements_evaluate_device without a workspace and without any
 * device allocation. No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). batch->workspace is ignored; the
 * workspace is a file-scope device array of SPARLAB_NOALLOC_SCRATCH_BYTES,
 * one per GPU, held until the process ends. A batch that needs more
 * (sparlab_element_workspace_doubles x 8 bytes: about 21 000 elements of
 * 8 points) is refused with SPARLAB_E_ARGUMENT before any device call. Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, other host threads) race. The outputs equal those of
 * sparlab_elements_evaluate_device bit for bit (the same kernels). */
#define SPARLAB_NOALLOC_SCRATCH_BYTES (128ll << 20)
int sparlab_elements_evaluate_noalloc_device(const sparlab_material* material, const sparlab_rule* rule,
                                             int kinematics, int want_tangent,
                                             const sparlab_element_batch* batch, void* stream);

/* ------------------------------------------------------------------------
 * The nonlinear system: the drop-in for detail::NonlinearSystem
 * (src/fem/NonlinearSystem.hpp:85-334) with the Assembler's cached pattern
 * (src/fem/Assembler.cpp:86-173) on the device. The system owns device copies
 * of the mesh, the pattern, the committed and trial history; the caller owns
 * u and the outputs.
 * ------------------------------------------------------------------------ */
typedef struct sparlab_system sparlab_system;

/* The scalars of an evaluation (NonlinearSystem::Evaluation,
 * src/fem/NonlinearSystem.hpp:27-47). gross = |sum_e |f_e|| (no external load:
 * f_ext = 0 in the forming analysis). */
typedef struct sparlab_evaluation {
  double energy;
  double gross;
  int64_t yielding_points;
  int64_t failed_elements;
  int64_t first_failed_element;   /* -1 when none */
  int32_t first_failure_status;   /* SPARLAB_ELEMENT_* */
  int32_t pad;
  double first_failure_value;
} sparlab_evaluation;

/* Builds the system of a Hex8 mesh (host arrays: coords 3 num_nodes,
 * conn 8 num_elements, VTK order): uploads the mesh, builds the pattern
 * (Assembler::build_pattern) and its gather maps, and starts the history
 * virgin (NonlinearSystem::initialise_history, src/fem/NonlinearSystem.hpp:
 * 273-311). SPARLAB_E_MESH for a node index out of range or repeated. */
int sparlab_system_create(const double* coords, int64_t num_nodes, const int32_t* conn, int64_t num_elements,
                          const sparlab_material* material, const sparlab_rule* rule, int kinematics,
                          sparlab_system** out);
void sparlab_system_destroy(sparlab_system* system);

/* The pattern: nnz, and host copies of row_ptr (3 num_nodes + 1) and col_idx
 * (nnz). The matrix is structurally symmetric, so these are also the
 * reference's column-major outer and inner arrays. */
int64_t sparlab_system_nnz(const sparlab_system* system);
int sparlab_system_pattern(const sparlab_system* system, int32_t* row_ptr, int32_t* col_idx);
/* Per node: the index of the first value of its diagonal 3 x 3 block (the
 * entries add_node_block writes, src/fem/Forming.cpp:280-297); row k of the
 * block starts k * (row length) later. */
int sparlab_system_diagonal_blocks(const sparlab_system* system, int32_t* first_value);

/* NonlinearSystem::evaluate(u, 0, want_tangent) on device pointers: u (3 nn),
 * fint (3 nn), values (nnz; with a tangent). column_major = 0 writes the
 * row-major values of K; 1 the reference's column-major values (they differ
 * only for a non-symmetric K). Asynchronous on stream; the trial history is
 * kept for sparlab_system_commit. */
int sparlab_system_evaluate_device(sparlab_system* system, const double* u, int want_tangent, int column_major,
                                   double* fint, double* values, void* stream);
/* Layout flags of sparlab_system_assemble_device. */
#define SPARLAB_VALUES_COLUMN_MAJOR 1 /* write the reference's column-major values (else row-major) */
#define SPARLAB_KE_COLUMN_MAJOR 2     /* ke is column-major per element (Eigen's storage of the
                                         reference's Matrix<double, 24, 24>; else row-major) */
/* Assembler::assemble_elementwise (src/fem/Assembler.cpp:183-215) / the force
 * gather of NonlinearSystem for given element arrays (device pointers): fe
 * (24 per element) into fint (3 nn) and, when gross is non-null, gross[0] =
 * |sum_e |f_e|| (device); ke (576 per element) into values (nnz), in the
 * orders `layout` (SPARLAB_VALUES_COLUMN_MAJOR | SPARLAB_KE_COLUMN_MAJOR; 0:
 * row-major both; 1 is the column_major = 1 of the other calls) selects.
 * Either part may be skipped with null outputs. Every value is the sum of its
 * contributions in ascending element order, the order of the reference's
 * serial scatter. Values in the order of ke's storage read ke contiguously. */
int sparlab_system_assemble_device(sparlab_system* system, const double* fe, const double* ke, int layout,
                                   double* fint, double* values, double* gross, void* stream);
/* stiffness_gross (src/fem/NonlinearSystem.hpp:55-67), the round-off scale of
 * the stiffness times the displacement (src/fem/Forming.cpp:991-993):
 * out[0] = || (|K| |x|)_free ||_2 for the system's pattern with values (nnz;
 * column_major as in sparlab_system_evaluate), x (3 nn) and the free DOFs
 * free_dofs[0, num_free) (int32, each in [0, 3 nn); one outside gives a NaN
 * norm). product (3 nn, may be null) receives
 * |K| |x|; each entry is the reference's sum, bit for bit (ascending columns,
 * products rounded before they are added); the norm's sum is a fixed-order
 * tree. Device pointers; asynchronous on stream. Calls on one system must not
 * overlap (they share its scratch). */
int sparlab_system_stiffness_gross_device(sparlab_system* system, const double* values, int column_major,
                                         const double* x, const int32_t* free_dofs, int64_t num_free, double* out,
                                         double* product, void* stream);

/* The free-free block of the Newton system: PartitionedFactor::reset and
 * build (src/fem/Forming.cpp:75-85, 208-236) for the free DOFs
 * free_dofs[0, num_free) (device, int32, strictly ascending in [0, 3 nn)).
 * The block's pattern is built on the device and kept by the system until the
 * next partition; *block_nnz (host) receives its nonzero count. The block's
 * columns are the free columns in order, each keeping its free rows in the
 * full pattern's order renumbered to free positions, as the reference's
 * kff_. A list that does not strictly ascend or leaves [0, 3 nn) is refused
 * (SPARLAB_E_ARGUMENT, no partition). Synchronizes stream (it reads the
 * count). */
int sparlab_system_partition_device(sparlab_system* system, const int32_t* free_dofs, int64_t num_free,
                                    int64_t* block_nnz, void* stream);
/* The block's pattern (device buffers): block_ptr (num_free + 1) and
 * block_index (block_nnz), the reference's kff_ outer and inner index arrays.
 * The system's pattern is symmetric, so they are also the block's row
 * pointers and column indices. Asynchronous on stream. */
int sparlab_system_partition_pattern_device(sparlab_system* system, int32_t* block_ptr, int32_t* block_index,
                                            void* stream);
/* PartitionedFactor::assemble (src/fem/Forming.cpp:89-98): block_values
 * (block_nnz, device) from the assembled values (nnz, device), bit for bit.
 * The block keeps the values' layout: the reference's column-major values
 * give its column-major kff_, row-major values the block's rows. Asynchronous
 * on stream. */
int sparlab_system_free_block_device(sparlab_system* system, const double* values, double* block_values,
                                     void* stream);
/* Solves K_ff x = b on the partition's block (block_values from
 * sparlab_system_free_block_device; b, x: num_free, device) where the
 * reference factorises and solves the tangent (PartitionedFactor::factorize
 * and solve, src/fem/Forming.cpp:104-172, called at :1001-1016 with b = -r
 * and symmetric = the system's and the contact's symmetry). The method
 * differs: preconditioned with the block's diagonal, conjugate gradients for
 * symmetric != 0 (the reference: CHOLMOD LLT, else LDL^T) and BiCGStab for a
 * non-symmetric block (the reference: UMFPACK or SparseLU), from x = 0, until
 * the true residual satisfies ||b - K x|| <= tolerance ||b|| or
 * max_iterations pass (the iteration's own residual is read every 16
 * iterations, every 4 with the multigrid preconditioner; when it passes, the true one is computed, and the iteration
 * restarts from it when that does not). *iterations (host) receives the
 * iterations run and *relative_residual (host) the final true residual
 * ||b - K x|| / ||b||. SPARLAB_E_SOLVER when the iteration breaks down (a
 * symmetric block that is not positive definite: a non-positive diagonal or
 * curvature; the reference then falls back to LDL^T or LU) or does not
 * converge; x then holds the last iterate. The block is read by rows: its
 * row-major values give K_ff, the reference's column-major ones K_ff^T (the
 * same for a symmetric block). Every sum has a fixed order: a solve repeats
 * bit for bit. b and x must not overlap (b is read again for the true
 * residual). The solve's vectors are the system's (allocated with the
 * partition): solves on one system must not overlap. Synchronizes stream
 * between runs of iterations. */
int sparlab_system_solve_free_block_device(sparlab_system* system, const double* block_values, int symmetric,
                                           const double* b, double* x, double tolerance, int64_t max_iterations,
                                           int64_t* iterations, double* relative_residual, void* stream);
/* The preconditioner of sparlab_system_solve_free_block_device (no
 * counterpart in the reference, which factorises the block directly):
 * DIAGONAL (the default) or MULTIGRID, smoothed-aggregation algebraic
 * multigrid, one V-cycle per application. The hierarchy is built at the first
 * solve after a partition or new coordinates: aggregates of the block's node
 * graph (greedy, every node with its neighbours), each aggregate's six
 * rigid-body modes at the nodes' coordinates orthonormalised as the
 * tentative prolongator T, P = T - 4 / (3 rho) D^-1 A T, A_c = P^T A P, until
 * a level has at most 256 rows, which is solved by its dense inverse; every
 * solve sets the values up again from its block (the device: fixed-order
 * sums, a solve repeats bit for bit) and smooths with Chebyshev of degree 2
 * on D^-1 A over [1.2 rho / 30, 1.2 rho], rho a 25-step power estimate; a
 * BiCGStab solve (symmetric = 0) builds it on the block's symmetric part
 * (A + A^T) / 2, whose spectrum is real.
 * SPARLAB_E_SOLVER from the solve when no hierarchy can be built or its
 * coarsest level is singular. The segmented solve always preconditions with
 * the diagonal. */
#define SPARLAB_PRECONDITIONER_DIAGONAL 0
#define SPARLAB_PRECONDITIONER_MULTIGRID 1
#define SPARLAB_MULTIGRID_MAX_LEVELS 12
int sparlab_system_set_block_preconditioner(sparlab_system* system, int kind);
/* The hierarchy of the partition (after a multigrid solve): *count levels
 * (0 without one), and each level's rows and nonzeros (host arrays of
 * SPARLAB_MULTIGRID_MAX_LEVELS, either may be null). */
int sparlab_system_block_preconditioner_levels(sparlab_system* system, int32_t* count, int64_t* rows,
                                               int64_t* nonzeros);
/* The batched form's block solve (the reference's per-problem
 * PartitionedFactor::factorize and solve, src/fem/Forming.cpp:104-172, for
 * every problem of a batch at once): the partition's block is block-diagonal
 * over num_segments problems, rows [segment_rows[p], segment_rows[p + 1])
 * (host, num_segments + 1 ascending offsets from 0 to num_free; a block row
 * with a column outside its segment is refused, SPARLAB_E_ARGUMENT). Each
 * problem is solved as sparlab_system_solve_free_block_device solves its rows
 * alone (the same block cut, sums, scalars, reads and restarts: the same
 * iterates bit for bit), all in one set of launches per iteration; a problem
 * that has converged or failed is left as it is. Per problem (host arrays of
 * num_segments): iterations, relative_residual, and status (SPARLAB_OK, or
 * SPARLAB_E_SOLVER where that problem's plain solve would fail).
 * SPARLAB_E_SOLVER when any problem failed (every result written). The
 * solve's vectors are the system's: solves on one system must not overlap.
 * Synchronizes stream between runs of iterations. */
int sparlab_system_solve_free_block_segments_device(sparlab_system* system, const double* block_values,
                                                    int symmetric, int32_t num_segments,
                                                    const int64_t* segment_rows, const double* b, double* x,
                                                    double tolerance, int64_t max_iterations, int64_t* iterations,
                                                    double* relative_residual, int32_t* status, void* stream);

/* Per-element stress measures (the reference's ElastoplasticStress,
 * include/sparlab/fem/Elastoplastic.hpp:206-221), SPARLAB_STRESS_FIELDS
 * doubles per element at these offsets (stresses and strains in tensorial
 * Voigt order xx, yy, zz, xy, yz, zx). */
#define SPARLAB_STRESS_FIELDS 28
#define SPARLAB_STRESS_CAUCHY 0                          /* 6, averaged over the points [Pa] */
#define SPARLAB_STRESS_PIOLA_KIRCHHOFF 6                 /* 6, second Piola-Kirchhoff [Pa] */
#define SPARLAB_STRESS_KIRCHHOFF 12                      /* 6, logarithmic kinematics only [Pa] */
#define SPARLAB_STRESS_LOGARITHMIC_STRAIN 18             /* 6, logarithmic kinematics only [-] */
#define SPARLAB_STRESS_VON_MISES 24                      /* of the averaged Cauchy stress */
#define SPARLAB_STRESS_MAX_EQUIVALENT_PLASTIC_STRAIN 25
#define SPARLAB_STRESS_MAX_STRAIN 26                     /* largest tensor strain component */
#define SPARLAB_STRESS_MIN_JACOBIAN 27                   /* smallest det F (1 for small strain) */
/* elastoplastic_stress (src/fem/Elastoplastic.cpp:342-350 ->
 * src/fem/IncompatibleModes.cpp:602-664) of every element at displacement u
 * (device, 3 nn) from the system's committed history and modes, as the
 * end-of-step output of src/fem/Forming.cpp:1263-1281 calls it: stress
 * (SPARLAB_STRESS_FIELDS per element, device); status (int32 per element,
 * SPARLAB_ELEMENT_* codes; may be null). A failed element (the reference's
 * MeshError / SolverError) gets zeros and its status. Asynchronous on stream. */
int sparlab_system_stress_device(sparlab_system* system, const double* u, double* stress, int32_t* status,
                                 void* stream);
/* Waits for the last evaluation and reads its scalars. SPARLAB_E_SOLVER when
 * an element failed (the reference's SolverError after its element loop). */
int sparlab_system_result(sparlab_system* system, sparlab_evaluation* out, void* stream);
/* The same from host arrays (includes the PCIe copies). */
int sparlab_system_evaluate(sparlab_system* system, const double* u, int want_tangent, int column_major,
                            double* fint, double* values, sparlab_evaluation* out);
/* NonlinearSystem::commit: the last evaluation's trial history becomes the
 * committed one (a buffer swap). */
int sparlab_system_commit(sparlab_system* system);
/* committed_all / internal_all (trial = 0) or the last trial (trial = 1), in
 * the library's layout (sparlab_state_fields x points x elements, and 9 per
 * element); set_committed / set_internal. Host arrays; null skips. */
int sparlab_system_history(const sparlab_system* system, int trial, double* states, double* alpha);
int sparlab_system_set_history(sparlab_system* system, const double* states, const double* alpha);
/* What NonlinearSystem answers about itself (src/fem/NonlinearSystem.hpp:
 * 104-114, 273-311): symmetric() (no recovering backstress), plastic() (the
 * material yields), elastoplastic(e) and averaged(e), the same for every
 * element here (one material; the incompatible-mode Hex8 averages no
 * dilatation), and the layout of the history. */
typedef struct sparlab_system_info {
  int64_t num_nodes;
  int64_t num_elements;
  int64_t nnz;
  int32_t points;            /* integration points per element */
  int32_t state_fields;      /* doubles of history per point (sparlab_state_fields) */
  int32_t kinematic_fields;  /* 6-vectors of kinematic history per point */
  int32_t general;           /* 1: they are the general return's backstresses alpha_i; 0: the radial return's beta */
  int32_t symmetric;         /* NonlinearSystem::symmetric() */
  int32_t plastic;           /* NonlinearSystem::plastic() */
  int32_t elastoplastic;     /* NonlinearSystem::elastoplastic(e), every e */
  int32_t averaged;          /* NonlinearSystem::averaged(e), every e: 0 */
  int32_t modes;             /* internal-mode parameters per element: 9 */
  int32_t kinematics;        /* SPARLAB_SMALL_STRAIN, _FINITE or _FINITE_LOGARITHMIC */
} sparlab_system_info;
int sparlab_system_info_get(const sparlab_system* system, sparlab_system_info* out);

/* NonlinearSystem::max_plastic_strain (src/fem/NonlinearSystem.hpp:181-188):
 * the largest equivalent plastic strain of the committed history, reduced on
 * the device (0 for an elastic material). Synchronous on `stream`. */
int sparlab_system_max_plastic_strain(sparlab_system* system, double* out, void* stream);

/* committed(e) / internal(e) (trial = 0), or the last evaluation's states and
 * modes (Evaluation::states / modes, trial = 1), for elements first ..
 * first + count - 1 (src/fem/NonlinearSystem.hpp:39-46, 108-145): states per
 * element, point and field ([e - first][q][f], state_fields per point, the
 * library's fields), alpha 9 per element. Host arrays; null skips. */
int sparlab_system_history_range(const sparlab_system* system, int trial, int64_t first, int64_t count,
                                 double* states, double* alpha);

/* The last evaluation's element forces (24 per element) and tangents (576). */
int sparlab_system_element_outputs(const sparlab_system* system, double* fe, double* ke);
/* The same for elements first .. first + count - 1 (host arrays of 24 and
 * 576 per element; null skips), e.g. a sample of a large mesh. */
int sparlab_system_element_outputs_range(const sparlab_system* system, int64_t first, int64_t count, double* fe,
                                         double* ke);

/* Replace the nodal reference coordinates (host, 3 per node) or the material
 * of a system (the pattern and the history are kept; the material must keep
 * the history layout). For gradient-based shape compensation and
 * calibration loops. */
int sparlab_system_set_coords(sparlab_system* system, const double* coords);
int sparlab_system_set_material(sparlab_system* system, const sparlab_material* material);

/* The differentiation form (no reference counterpart): vector-Jacobian
 * products of the last evaluation's assembled internal force, at its
 * converged modes and the fixed committed history, for a cotangent v
 * (device, 3 per node). wrt selects (device outputs):
 *   SPARLAB_VJP_COORDS        grad_coords (3 nn) = v^T d f_int / d X (reference coordinates)
 *   SPARLAB_VJP_DISPLACEMENT  grad_disp (3 nn)   = v^T d f_int / d u (= v^T K_t)
 *   SPARLAB_VJP_PARAMETERS    grad_params (11)   = v^T d f_int / d (r0, r45, r90,
 *                             C_1, gamma_1, ..., C_4, gamma_4): Hill48 r-values (r-value
 *                             calibration) and the Chaboche backstresses; zero for
 *                             parameters the material does not have.
 * u must be the evaluation's displacement; call before commit and before new
 * coordinates, material or history. Asynchronous on stream. */
#define SPARLAB_VJP_COORDS 1
#define SPARLAB_VJP_DISPLACEMENT 2
#define SPARLAB_VJP_PARAMETERS 4
#define SPARLAB_VJP_NUM_PARAMETERS 11
int sparlab_system_vjp_device(sparlab_system* system, const double* u, const double* v, int wrt,
                              double* grad_coords, double* grad_disp, double* grad_params, void* stream);

/* The pattern of a Hex8 mesh on the host, without a device (the reference's
 * Assembler::build_pattern, src/fem/Assembler.cpp:86-173): returns nnz (or a
 * negative code); with row_ptr (3 num_nodes + 1) and col_idx (nnz) non-null,
 * fills them, and block_offset (64 per element, may be null) with the
 * reference's per-element block offsets. */
int64_t sparlab_pattern_build(int64_t num_nodes, int64_t num_elements, const int32_t* conn, int32_t* row_ptr,
                              int32_t* col_idx, int32_t* block_offset);

/* ------------------------------------------------------------------------
 * The batched form: many independent problems (forming meshes) advanced
 * together, each a NonlinearSystem (src/fem/NonlinearSystem.hpp:85-334) of
 * its own. Their union is one block-diagonal system evaluated by one launch
 * sequence, captured in a HIP graph and replayed every Newton iteration.
 * ------------------------------------------------------------------------ */
typedef struct sparlab_batch sparlab_batch;

/* Builds the batch from host arrays: problem p owns nodes
 * [node_offset[p], node_offset[p + 1]) of coords (3 per node) and elements
 * [element_offset[p], element_offset[p + 1]) of conn (8 per element, node
 * indices local to the problem, VTK order); one material, rule and kinematics
 * for all. The union's DOFs, CSR rows and element outputs follow problem
 * order, so problem p's DOFs start at 3 node_offset[p] and its CSR values
 * are one contiguous slice (sparlab_batch_value_offsets). */
int sparlab_batch_create(int32_t num_problems, const int64_t* node_offset, const int64_t* element_offset,
                         const double* coords, const int32_t* conn, const sparlab_material* material,
                         const sparlab_rule* rule, int kinematics, sparlab_batch** out);
void sparlab_batch_destroy(sparlab_batch* batch);
/* The union's system: pattern, history, commit and the plain entry points
 * (owned by the batch). */
sparlab_system* sparlab_batch_system(sparlab_batch* batch);
/* num_problems + 1 offsets of each problem's first CSR value. */
int sparlab_batch_value_offsets(const sparlab_batch* batch, int64_t* value_offset);
/* Every problem's NonlinearSystem::evaluate at once (device pointers u, fint
 * 3 per union node; values nnz with a tangent), and per-problem scalars.
 * The first call for a given (u, fint, values, want_tangent, column_major,
 * committed buffers) captures the launches into a HIP graph; later calls
 * with the same pointers replay it, so the buffers must stay allocated while
 * the batch lives (write new displacements into the same u). */
int sparlab_batch_evaluate_device(sparlab_batch* batch, const double* u, int want_tangent, int column_major,
                                  double* fint, double* values, void* stream);
/* Capture and replay (enabled = 1, the default) or direct launches (0). */
int sparlab_batch_set_graphs(sparlab_batch* batch, int enabled);
/* Waits for the last batched evaluation; per problem (host arrays of
 * num_problems, each may be null): energy, gross, yielding points, failed
 * elements, the lowest failed element (local index, or -1).
 * SPARLAB_E_SOLVER when any element failed. */
int sparlab_batch_result(sparlab_batch* batch, double* energy, double* gross, int64_t* yielding,
                         int64_t* failed, int64_t* first_failed, void* stream);
/* Times the local solve at 8, 16, ..., 256 elements per wavefront on the
 * batch at displacement u (device) and keeps the fastest (written to
 * *chosen when non-null). The timing runs write the trial buffers only. */
int sparlab_batch_autotune(sparlab_batch* batch, const double* u, int* chosen, void* stream);

/* ------------------------------------------------------------------------
 * Penalty contact of slave nodes with a rigid tool (ToolContact::evaluate,
 * src/fem/RigidTool.cpp:398-555, for one tool over a list of slave nodes).
 * ------------------------------------------------------------------------ */
typedef struct sparlab_tool {
  int32_t shape;               /* 0 plane, 1 sphere, 2 cylinder */
  int32_t pad;
  double centre[3];            /* c(t), the tool's position at the evaluation */
  double centre0[3];           /* c(t0), the increment's start (nodes new to contact) */
  double radius;
  double normal[3];            /* plane */
  double axis[3];              /* cylinder */
  double friction;             /* mu (0: frictionless) */
  double tangential_ratio;     /* kappa_T / kappa */
} sparlab_tool;

/* Device pointers. coords, u, start: 3 per node (start: u at the start of the
 * increment). Per slave: nodes (int32), kappa_area (kappa_j A_j,
 * src/fem/RigidTool.cpp:298-321), the committed friction history
 * (hist_present int32, hist_force 3, hist_relative 3; hist_present may be
 * null: no history). Outputs: residual (3 per node) -= the contact forces;
 * blocks (9 per slave, row-major, with want_tangent); active (int32 per
 * slave); the trial history (trial_present, trial_force 3, trial_relative 3);
 * work: 8 doubles and one int32 per slave; scalars (10 doubles): gross,
 * round_off, resultant[3], max_penetration, slipping, asymmetric blocks,
 * active, nodes through the tool (> 0: the reference's SolverError). */
int sparlab_contact_evaluate_device(const sparlab_tool* tool, int friction_symmetric, int want_tangent,
                                    const double* coords, const double* u, const double* start,
                                    int64_t num_slaves, const int32_t* nodes, const double* kappa_area,
                                    const int32_t* hist_present, const double* hist_force,
                                    const double* hist_relative, double* residual, double* blocks,
                                    int32_t* active, int32_t* trial_present, double* trial_force,
                                    double* trial_relative, double* work, int32_t* work_status,
                                    double* scalars, void* stream);

/* ToolContact::begin_increment (src/fem/RigidTool.cpp:345-393) for one tool:
 * the slaves whose node at u0 (3 per node) lies within radius + margin of the
 * tool's path from tool->centre0 = c(t0) to tool->centre = c(t1) (a cylinder:
 * in the plane normal to its axis; a plane: within margin of either
 * position), in slave order: candidates (int32 slave indices, num_slaves
 * capacity) and count (int32[1]). work: num_slaves + num_slaves / 256 + 1
 * int32 of scratch. Device pointers; asynchronous on stream. */
int sparlab_contact_candidates_device(const sparlab_tool* tool, const double* coords, const double* u0,
                                      int64_t num_slaves, const int32_t* nodes, double margin, int32_t* candidates,
                                      int32_t* count, int32_t* work, void* stream);

/* ToolContact::anticipate (src/fem/RigidTool.cpp:557-603) for one tool at
 * tool->centre over the listed slaves (nodes and kappa_area per listed slave,
 * e.g. the candidates): a slave whose node lies outside the tool at u and
 * that the Newton correction du (3 per node) drives into it (or through it)
 * gets residual += kappa_area g n (residual: 3 per node) and the block
 * kappa_area n n^T (blocks: 9 per listed slave, row-major, zero for the
 * others) with active = 1; added (int32[1]) counts them. The blocks go into
 * the matrix with sparlab_add_node_blocks_device. Device pointers;
 * asynchronous on stream. */
int sparlab_contact_anticipate_device(const sparlab_tool* tool, const double* coords, const double* u,
                                      const double* du, int64_t num_slaves, const int32_t* nodes,
                                      const double* kappa_area, double* residual, double* blocks, int32_t* active,
                                      int32_t* added, void* stream);

/* add_node_block (src/fem/Forming.cpp:280-297) for every active slave: its
 * block into the assembled values at the node's diagonal block
 * (first_value, row_len per slave: sparlab_system_diagonal_blocks and 3 times
 * the node's neighbour count); column_major as in sparlab_system_evaluate. */
int sparlab_add_node_blocks_device(const double* blocks, const int32_t* active, const int32_t* first_value,
                                   const int32_t* row_len, int64_t num_slaves, int column_major, double* values,
                                   void* stream);

/* The message of the calling thread's last failed call. */
const char* sparlab_last_error(void);

/* The library version string. */
const char* sparlab_version(void);

/* The released part's height map before its nearest fill: the reference's
 * formed_surface (python/precomp/fea/results.py:228-248) interpolates the
 * deformed tool-side surface with _interpolate_structured (:262-283) and
 * raycast_top (python/precomp/geometry/raycast.py:27-100). surface (device):
 * the deformed positions of the surface's structured node grid, (ny_s, nx_s,
 * 3), row y-major; cut into two triangles per cell ((a, b, c), (a, c, d),
 * counterclockwise from the cell's lowest corner); cast along vertical rays
 * through the grid nodes x0 + i h, y0 + j h (i < nx, j < ny). z (device,
 * ny nx): the highest hit per node, -inf where no triangle is hit; hit
 * (device, ny nx bytes): 1 where hit, bit for bit the reference's (numpy's
 * order of operations). The reference then gives each node outside the
 * surface the height of the nearest hit node (scipy's distance transform);
 * that fill stays with the caller. A surface with triangles of zero or
 * negative plan area is refused (SPARLAB_E_ARGUMENT, *folded (host) = their
 * count), as the reference's PrecompError. Synchronizes stream (it reads the
 * fold count first). */
int sparlab_surface_raycast_device(const double* surface, int64_t nx_s, int64_t ny_s, double x0, double y0,
                                   int64_t nx, int64_t ny, double h, double* z, uint8_t* hit, int64_t* folded,
                                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SPARLAB_GPU_H */
