/*
 * onsas_hip.h -- C ABI of libonsas_hip.so, the MI355X (gfx950) engine for the
 * nonlinear static Newton path of ONSAS.jl.
 *
 * Every function is extern "C", takes plain pointers and sizes, returns an int
 * status (ONSAS_OK = 0, negative on error, message in onsas_last_error) and
 * never throws or aborts across the boundary.  Host pointers are borrowed for
 * the duration of the call only; results are copied into caller buffers.  A
 * context owns its device memory and one HIP stream, is single-threaded, and
 * every call is synchronous at return.
 *
 * Each declaration cites the ONSAS.jl v0.4.6 interface it replaces
 * (file:line, relative to the ONSAS.jl source tree).
 */
#ifndef ONSAS_HIP_H
#define ONSAS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct onsas_ctx onsas_ctx;

/* ---- status codes ------------------------------------------------------- */
#define ONSAS_OK 0
#define ONSAS_ERR_ARG (-1)             /* invalid argument or size                  */
#define ONSAS_ERR_HIP (-2)             /* HIP runtime failure                        */
#define ONSAS_ERR_NEGATIVE_VOLUME (-3) /* Tetrahedrons.jl:136 assertion              */
#define ONSAS_ERR_STATE (-4)           /* call out of order (e.g. before finalize)   */
#define ONSAS_ERR_NOT_FINITE (-5)      /* NaN/Inf in a residual or update norm       */
#define ONSAS_ERR_LINEAR_SOLVE (-6)    /* PCG breakdown with no fallback / no conv.  */
#define ONSAS_ERR_ZERO_NORM (-7)       /* StructuralSolvers.jl:152-153 assertions    */
#define ONSAS_ERR_ALLOC (-8)           /* device or host allocation failed           */
#define ONSAS_ERR_INJECTED (-9)        /* forced by onsas_inject_error (tests only)  */
#define ONSAS_ERR_UNSUPPORTED (-10)    /* combination not provided by the engine     */

/* ---- enums --------------------------------------------------------------- */
#define ONSAS_PRECISION_MIXED 0 /* FP32 tangent + FP32 PCG, FP64 f_int/residual/norms/U */
#define ONSAS_PRECISION_FP64 1  /* everything FP64, the multigrid's stored values too   */

#define ONSAS_MAT_SVK 1        /* src/Materials/SVKMaterial.jl          params (lambda, G) */
#define ONSAS_MAT_NEOHOOKEAN 2 /* src/Materials/NeoHookeanMaterial.jl   params (K, G)      */
#define ONSAS_MAT_LINEAR 3     /* src/Materials/IsotropicLinearElasticMaterial.jl (E, nu)  */

#define ONSAS_STRAIN_ROTATED 0 /* RotatedEngineeringStrain (src/Entities/Trusses.jl:37) */
#define ONSAS_STRAIN_GREEN 1   /* GreenStrain (src/Entities/Trusses.jl:44)              */

#define ONSAS_LINSOLVE_AUTO 0  /* PCG; GMRES when PCG breaks down (p'Kp <= 0)        */
#define ONSAS_LINSOLVE_PCG 1   /* PCG only (preconditioner: ONSAS_PRECOND_*)          */
#define ONSAS_LINSOLVE_GMRES 2 /* restarted GMRES only (block-Jacobi, right precond.) */

#define ONSAS_PRECOND_BLOCK_JACOBI 0 /* 3x3 nodal block-Jacobi                          */
#define ONSAS_PRECOND_AMG 1          /* smoothed-aggregation multigrid V-cycle (rigid-body
                                        modes, Chebyshev smoothing)                        */

#define ONSAS_AMG_REFRESH_STEP 0     /* multigrid coarse levels rebuilt at the first solve of
                                        each load step (after onsas_set_loads)             */
#define ONSAS_AMG_REFRESH_ASSEMBLY 1 /* ... rebuilt after every assembly                   */
#define ONSAS_AMG_REFRESH_ONCE 2     /* ... built from the first solve's tangent only (with
                                        the smoothing bounds; the finest level is still the
                                        current tangent)                                    */

/* Linear-solver settings; replace the tolerances of
 * src/StructuralSolvers/StructuralSolvers.jl:229-234 (abstol 0, reltol sqrt(eps),
 * maxiter = number of free dofs) and the LinearSolve algorithm argument of
 * solve! (StructuralSolvers.jl:217-222). */
typedef struct {
    int32_t kind;          /* ONSAS_LINSOLVE_*                                         */
    int32_t maxiter;       /* <= 0: number of free dofs (the reference default)        */
    double rtol;           /* stop at ||b - A x|| <= rtol ||b||; <= 0: sqrt(eps) in FP64
                              mode, 1e-6 in mixed mode                                  */
    int32_t gmres_restart; /* Krylov subspace dimension (<= 0: 40)                      */
    int32_t check_every;   /* PCG iterations per device-resident chunk (<= 0: 16)       */
    int32_t forcing;       /* 1: inexact-Newton forcing terms (Eisenstat-Walker choice 2
                              with a "just enough" floor), 0: fixed rtol                 */
    double newton_tol;     /* forcing: the Newton rel_res_force_tol                      */
    double eta_max;        /* forcing: loosest inner tolerance (<= 0: 0.1)               */
    double eta_first;      /* forcing: inner tolerance of the first Newton iteration of a
                              load step (<= 0: 1e-3)                                      */
    int32_t precond;       /* PCG preconditioner: ONSAS_PRECOND_* (GMRES: block-Jacobi)  */
    int32_t amg_refresh;   /* ONSAS_AMG_REFRESH_*: when the multigrid coarse levels follow
                              the tangent (its finest level, the tangent itself, always
                              does)                                                      */
} onsas_linear_settings;

/* Per-Newton-iteration record: the four norms of
 * src/StructuralAnalyses/NonLinearStaticAnalyses.jl:138-141 plus solver data. */
typedef struct {
    double dU_norm;          /* ||dU||                                               */
    double dU_rel;           /* ||dU|| / ||U before the update||                     */
    double r_norm;           /* ||f_ext[free] - f_int[free]||  (FP64)                */
    double r_rel;            /* r_norm / ||f_ext|| (all dofs)                        */
    double fext_norm;        /* ||f_ext||                                            */
    double U_norm;           /* ||U|| before the update                              */
    double lin_rel_residual; /* final ||b - A x|| / ||b|| of the inner solve (recurrence) */
    int32_t lin_iters;       /* inner iterations (PCG + GMRES)                       */
    int32_t lin_solver;      /* 1 PCG, 2 GMRES, 3 PCG then GMRES                     */
    int32_t lin_converged;   /* 1 if the inner tolerance was met                     */
    int32_t reserved;
    double assemble_ms;      /* device time of the preceding onsas_assemble          */
    double solve_ms;         /* device time of this step's residual+solve+update     */
    double lin_rtol;         /* inner relative tolerance used by this step            */
} onsas_step_stats;

/* Cumulative device-time counters (hipEvents), reset by onsas_reset_timings. */
typedef struct {
    double assemble_ms;      /* whole assembly: f_int, K, block-Jacobi inverse       */
    double scatter_ms;       /* element + block kernels (gather: state + gather pass; */
                             /* scatter: the colour launches)                        */
    double solve_ms;         /* residual + linear solve + update                     */
    double pcg_ms;           /* PCG iterations only                                  */
    int64_t n_assemble;
    int64_t n_steps;
    int64_t pcg_iters;
    int64_t gmres_iters;
    double spmv_ms;          /* event-timed SpMV kernel launches (second iteration of each
                                PCG)                                                  */
    int64_t spmv_launches;   /* number of launches in spmv_ms                          */
    double elem_ms;          /* gather assembly: the per-element state pass alone     */
    int64_t assembly_mode;   /* 1 gather (default), 2 staged gather (one tet group,  */
                             /* ONSAS_GATHER=staged), 0 coloured scatter             */
    double cycle_spmv_ms;    /* multigrid: event-timed exact fine-level products of the
                                V-cycle (r = b - K x, both, second iteration of each
                                PCG)                                                  */
    int64_t cycle_spmv_launches;
} onsas_timings;

/* ---- lifetime ------------------------------------------------------------ */

/* Create a context on the current HIP device.  Replaces FullStaticState(s)
 * allocation (src/StructuralAnalyses/StaticStates.jl:79-102). */
int onsas_create(int32_t precision, onsas_ctx** out);
void onsas_destroy(onsas_ctx* ctx);
/* Message of the last failing call on ctx (NULL ctx: last creation error). */
const char* onsas_last_error(const onsas_ctx* ctx);
/* Library version string. */
const char* onsas_version(void);

/* ---- model description (replaces Mesh / Structure construction) ----------- */

/* Node coordinates, SoA.  Node i owns dofs 3i, 3i+1, 3i+2 (set_dofs!,
 * src/Meshes/Meshes.jl:85-98, 0-based here). */
int onsas_set_nodes(onsas_ctx* ctx, int64_t nn, const double* x, const double* y,
                    const double* z);

/* A group of Tetrahedron elements with one material
 * (src/Entities/Tetrahedrons.jl:24-48, StructuralMaterial pairs of
 * src/StructuralModel/StructuralMaterials.jl:25-35).  conn: 4*ne node ids,
 * element-major, reference node order.  params: 2 doubles (see ONSAS_MAT_*).
 * shear_per_elem: optional ne values replacing params[1] (G) per element.
 * Returns the group id in *group. */
int onsas_add_tet_group(onsas_ctx* ctx, int64_t ne, const int32_t* conn, int32_t material,
                        const double* params, const double* shear_per_elem, int32_t* group);

/* A group of Truss elements (src/Entities/Trusses.jl:54-99) with an
 * AbstractHyperElasticMaterial of elasticity modulus E
 * (elasticity_modulus, SVKMaterial.jl:75-78).  conn: 2*ne; area: ne values
 * (area(cross_section), e.g. src/CrossSections/Circles.jl:20). */
int onsas_add_truss_group(onsas_ctx* ctx, int64_t ne, const int32_t* conn,
                          int32_t strain_model, double E, const double* area, int32_t* group);

/* Triangular faces carrying one Pressure boundary condition
 * (src/BoundaryConditions/LocalLoadBoundaryConditions.jl:21-56 applied to
 * TriangularFace, src/Entities/TriangularFaces.jl).  conn: 3*nf.  Returns the
 * pressure-group id whose value is set per load step by onsas_set_loads. */
int onsas_add_pressure_faces(onsas_ctx* ctx, int64_t nf, const int32_t* conn, int32_t* group);

/* Fixed dofs (0-based, any order, duplicates allowed): the union of FixedField
 * conditions (src/BoundaryConditions/FixedFieldBoundaryConditions.jl:39-48,
 * src/StructuralModel/Structures.jl:129-142). */
int onsas_set_fixed_dofs(onsas_ctx* ctx, int64_t nfixed, const int64_t* dofs);

/* Build the 3x3 block-CSR pattern, the element colouring and the
 * element->block slot maps; upload everything.  Fails with
 * ONSAS_ERR_NEGATIVE_VOLUME on an inverted tetrahedron (Tetrahedrons.jl:136). */
int onsas_finalize(onsas_ctx* ctx);

/* Replaces the linear_solve argument and default tolerances of solve!
 * (StructuralSolvers.jl:217-234).  With ONSAS_PRECOND_AMG on a finalized
 * context it also builds the multigrid hierarchy's pattern (host work, once). */
int onsas_set_linear_solver(onsas_ctx* ctx, const onsas_linear_settings* s);

/* ---- per load step ------------------------------------------------------ */

/* External forces at the current load factor: apply!(sa, load_bcs)
 * (src/StructuralAnalyses/StructuralAnalyses.jl:228-241).  nodal: n values
 * (GlobalLoad contributions already summed per dof, NULL = 0); pressures: one
 * value p(t) per pressure group (NULL if none).  Both arrays are read before
 * the call returns; with pressures only, the device work stays queued. */
int onsas_set_loads(onsas_ctx* ctx, const double* nodal, int32_t npressure,
                    const double* pressures);

/* ---- per Newton iteration ------------------------------------------------ */

/* assemble!(s, sa) (src/StructuralAnalyses/StaticAnalyses.jl:99-122): f_int
 * (FP64) and the tangent at the device-resident U.  fint_out: optional n
 * values copied back (the call then waits for the assembly); without it the
 * call returns once the work is queued, and every later call sees its result
 * (they run on the context's one stream). */
int onsas_assemble(onsas_ctx* ctx, double* fint_out);

/* step!(sa, ::NewtonRaphson, linear_solver)
 * (src/StructuralAnalyses/NonLinearStaticAnalyses.jl:107-148): r = f_ext - f_int
 * on free dofs, solve K dU = r, norms, U[free] += dU.  Requires a preceding
 * onsas_assemble. */
int onsas_newton_step(onsas_ctx* ctx, onsas_step_stats* out);

/* ---- state access -------------------------------------------------------- */
int onsas_get_U(onsas_ctx* ctx, double* U);       /* displacements(state), n values */
/* The same copy in two halves, so that a load step's store! (src/StructuralSolvers/
 * Solutions.jl, the per-step copy of U) overlaps the next step's work:
 * onsas_snapshot_U enqueues the copy of U, as of the work enqueued so far,
 * into pinned snapshot buffer `slot` (0 or 1) and returns at once;
 * onsas_read_snapshot waits for that copy and writes the n values to U.  The
 * read uses no stream and touches nothing but the slot, so it may run on
 * another thread while this thread keeps driving the context; a slot must be
 * read before it is snapshot again. */
int onsas_snapshot_U(onsas_ctx* ctx, int32_t slot);
int onsas_read_snapshot(onsas_ctx* ctx, int32_t slot, double* U);
int onsas_set_U(onsas_ctx* ctx, const double* U); /* initial_state / resume         */
int onsas_get_fext(onsas_ctx* ctx, double* fext); /* external_forces(state)         */
int onsas_get_fint(onsas_ctx* ctx, double* fint); /* internal_forces(state)         */
int onsas_get_sizes(onsas_ctx* ctx, int64_t* nn, int64_t* nnzb, int64_t* ncolors);
/* Layout of the gather assembly.  No counterpart in the reference (its
 * assembly is a COO scatter); for the byte model of benchmarks and tests.
 * out holds 6 values: [0] rows per workgroup of the staged gather (0: the
 * list-driven gather or the scatter), [1] its workgroups, [2] staged elements
 * over all workgroups, [3] contribution entries stored (with padding), [4] LDS
 * bytes per workgroup, [5] steps. */
int onsas_gather_layout(onsas_ctx* ctx, int64_t* out);

/* Block-CSR pattern of the tangent (3x3 blocks, node indices): row_ptr nn+1,
 * col nnzb.  Replaces the SparseMatrixCSC pattern of tangent_matrix(state)
 * (StaticStates.jl:88, 118-119). */
int onsas_get_pattern(onsas_ctx* ctx, int32_t* row_ptr, int32_t* col);
/* Tangent values, 9 per block (row-major 3x3), as double.  Rows and columns of
 * fixed dofs hold the identity (the free-free block equals K[free, free] of
 * NonLinearStaticAnalyses.jl:120). */
int onsas_get_tangent(onsas_ctx* ctx, double* vals);

/* Multigrid hierarchy inspection.  No counterpart in the reference (its CG is
 * unpreconditioned); for tests and diagnostics.  Available once a PCG solve
 * with ONSAS_PRECOND_AMG has built the hierarchy (0 levels before).  Level 0
 * is the tangent (onsas_get_tangent); level l >= 1 has 6 dofs per node. */
int onsas_amg_levels(onsas_ctx* ctx, int32_t* nlevels);
/* Level l: nodes, dofs per node, operator blocks, blocks of the prolongator
 * from level l+1 (0 on the coarsest), and the lambda_max(D^-1 A) estimate. */
int onsas_amg_level_info(onsas_ctx* ctx, int32_t level, int64_t* n, int32_t* bs, int64_t* nnzb,
                         int64_t* nnz_p, double* lam);
/* Level l >= 1 operator, 6x6 block CSR: row_ptr (n+1), col (nnzb), vals (36 per
 * block, row-major; FP64 copy of the stored values).  NULL outputs are skipped. */
int onsas_amg_get_operator(onsas_ctx* ctx, int32_t level, int32_t* row_ptr, int32_t* col,
                           double* vals);
/* Prolongator from level l+1 to level l, block CSR over level-l nodes with
 * bs x 6 blocks: p_ptr (n+1), p_col (nnz_p), vals (bs*6 per block, row-major;
 * the stored FP16 values widened (mixed) or the stored FP64 values (FP64 mode)). */
int onsas_amg_get_prolongator(onsas_ctx* ctx, int32_t level, int32_t* p_ptr, int32_t* p_col,
                              double* vals);
/* Measurement (no counterpart in the reference): `launches` (2..10000)
 * back-to-back launches of the multigrid cycle's two fine-level tangent
 * products, alternating, on the last solve's vectors, between two events on
 * the context's stream; *ms_per_launch = elapsed / launches.  Spreading the
 * events' own cost over the batch gives the kernel's duration as a kernel
 * trace reports it.  Needs the hierarchy (onsas_amg_levels > 0); scribbles
 * on the cycle's work vectors only. */
int onsas_time_cycle_products(onsas_ctx* ctx, int32_t launches, double* ms_per_launch);
/* The same measurement for any smoothed level (no counterpart in the
 * reference): level 0 as onsas_time_cycle_products; a level >= 1 stored in
 * sliced ELL cycles through its products of the V-cycle on that level's work
 * vectors: the four of the degree-2 smoothing (x = d0 + d1, r = b - A x, r and
 * d0 = D^-1 r / theta, x += d0 + d1), or the two of the degree-1 step.  Levels without a sliced-ELL operator and the coarsest level are
 * rejected (ONSAS_ERR_ARG). */
int onsas_time_level_products(onsas_ctx* ctx, int32_t level, int32_t launches,
                              double* ms_per_launch);
/* y = K x with the assembled tangent (n values each). */
int onsas_tangent_matvec(onsas_ctx* ctx, const double* x, double* y);
/* The same product without allocating anything.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  x and
 * y pass through a fixed 32 MiB device array that each GPU the library runs on
 * holds until the process ends, so n is limited to 2,097,152 dofs (699,050
 * nodes; larger models get ONSAS_ERR_ARG).  Calls that overlap on one GPU
 * share that array and race; the result is otherwise that of
 * onsas_tangent_matvec. */
int onsas_tangent_matvec_noalloc(onsas_ctx* ctx, const double* x, double* y);

/* Batch internal_forces(mat, e, u_e) over one element group at the current U
 * (src/Entities/Entities.jl:173-218; Tetrahedrons.jl:186-266;
 * Trusses.jl:126-184).  Outputs in user element order, FP64, any may be NULL:
 * f (ne*nd), K (ne*nd*nd row-major), stress (ne*9), strain (ne*9) where nd =
 * 12 (tet) or 6 (truss).  Stress/strain are (P = F S, C = F'F) for
 * hyperelastic tets, (sigma, eps) for linear tets, and 3x3 with only [0,0] set
 * (Piola sigma11, strain) for trusses. */
int onsas_element_forces(onsas_ctx* ctx, int32_t group, double* f, double* K, double* stress,
                         double* strain);

/* Per-element stress and strain (the sigma/eps stored by store!,
 * src/StructuralAnalyses/StaticAnalyses.jl:157-174).  at_last_assembly != 0
 * evaluates them at the U of the last onsas_assemble, which is what the
 * reference stores (its last assemble! precedes the final U += dU,
 * NonLinearStaticAnalyses.jl:90-98); 0 evaluates them at the current U. */
int onsas_get_stress_strain(onsas_ctx* ctx, int32_t group, int32_t at_last_assembly,
                            double* stress, double* strain);

/* Solve K x = b with the assembled tangent and the configured linear solver
 * (free dofs; b at fixed dofs is ignored, x there is 0). */
int onsas_linear_solve(onsas_ctx* ctx, const double* b, double* x, int32_t* iters,
                       double* rel_residual);

/* Adjoint sensitivity of J(U) to the per-element shear moduli of a tet group
 * (no counterpart in the reference; the pullback of the solve, for a
 * ChainRulesCore rrule of solve! on the NeoHookean / SVK shear modulus of
 * src/Materials/NeoHookeanMaterial.jl:94-108 and SVKMaterial.jl:89-100).
 * Assembles K at the current U (the converged state), solves K lambda = dJ/dU
 * on the free dofs to ||dJ/dU - K lambda|| <= 1e-11 ||dJ/dU|| by iterative
 * refinement (FP64 element-wise residual, corrections with the configured
 * solver: FP32 PCG in mixed mode), then dJdG[e] = -lambda_e . d f_int,e / d G_e.
 * dJdU: n values (fixed dofs ignored); dJdG: ne values in user order; iters:
 * total inner iterations. */
int onsas_adjoint_shear(onsas_ctx* ctx, int32_t group, const double* dJdU, double* dJdG,
                        int32_t* iters);
/* Replace the per-element shear moduli G of an SVK / NeoHookean tet group (ne
 * values, user order) without rebuilding the model: the parameter update of a
 * gradient-based inverse problem (no counterpart in the reference).  The next
 * onsas_assemble uses the new moduli. */
int onsas_set_shear_moduli(onsas_ctx* ctx, int32_t group, const double* G);

/* Point location in a tet group, reference configuration (replaces
 * PointEvalHandler(mesh, points) with the Serial search of
 * src/Meshes/Handlers.jl:67-87 and src/Meshes/Searches.jl:64-82, and the
 * weights of src/Entities/Tetrahedrons.jl:311-313): for each point, the first
 * element in user order that contains it (barycentric coordinates >= -1e-12)
 * in elem[i] (-1 outside the group), its four barycentric weights in
 * weights[4i..4i+3] (0 outside).  Interpolating a nodal field at the point is
 * sum_a weights[a] * field(node a).  A bin grid is built on the first call. */
int onsas_locate_points(onsas_ctx* ctx, int32_t group, int64_t npts, const double* x,
                        const double* y, const double* z, int64_t* elem, double* weights);

/* Timers. */
int onsas_get_timings(onsas_ctx* ctx, onsas_timings* t);
int onsas_reset_timings(onsas_ctx* ctx);

/* Test hook: the next fallible call on ctx returns `code` (error-path tests). */
int onsas_inject_error(onsas_ctx* ctx, int32_t code);

#ifdef __cplusplus
}
#endif
#endif /* ONSAS_HIP_H */
