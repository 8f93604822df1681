/*
 * onsas_oracle.h -- FP64 CPU restatement of the ONSAS.jl nonlinear static path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the parity tests, the smoke
 * check and the CPU-baseline leg of the benchmark compare the HIP engine
 * against.  No product code links or calls it.
 *
 * Every function restates a reference function line for line; the cited file
 * paths are relative to the ONSAS.jl v0.4.6 source tree.  Third-party pieces
 * that are not vendored in the reference (IterativeSolvers 0.9.4 `cg!`,
 * ForwardDiff 0.10.38 `gradient`, SparseArrays 1.11 `setindex!`) are restated
 * from their published algorithms and named at the function that replaces
 * them.
 *
 * Conventions (all arrays are row-major C arrays):
 *   X, u of a tet      : 12 doubles, node-major  [n1x n1y n1z n2x ... n4z]
 *   3x3 tensors        : 9 doubles, row-major
 *   element matrices   : n x n doubles, row-major (symmetric in every case)
 *   Voigt order        : 11 22 33 23 13 12  (src/Utils.jl:42)
 *   dof numbering      : node i (0-based) owns dofs 3i, 3i+1, 3i+2
 *                        (src/Meshes/Meshes.jl:85-98, shifted to 0-based)
 */
#ifndef ONSAS_ORACLE_H
#define ONSAS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_MAT_SVK = 1, OR_MAT_NEOHOOKEAN = 2, OR_MAT_LINEAR = 3 };
enum { OR_STRAIN_ROTATED = 0, OR_STRAIN_GREEN = 1 };
enum { OR_OK = 0, OR_ERR_NEGATIVE_VOLUME = -2, OR_ERR_ARG = -1, OR_ERR_ZERO_NORM = -3,
       OR_ERR_ALLOC = -4 };

/* ---- materials ---------------------------------------------------------- */
/* SVK: S = lam tr(E) I + 2G E, dSdE Voigt blocks (src/Materials/SVKMaterial.jl:89-100) */
void or_svk_stress(double lam, double G, const double E[9], double S[9], double dSdE[36]);
/* SVK energy (src/Materials/SVKMaterial.jl:57-60) */
double or_svk_energy(double lam, double G, const double E[9]);
/* NeoHookean S = G(I - C^-1) + K J (J-1) C^-1 (src/Materials/NeoHookeanMaterial.jl:94-108);
 * dSdE by forward-mode differentiation of the same function with respect to the
 * upper triangle of E, Voigt factor 0.5 (NeoHookeanMaterial.jl:115-129). */
void or_nh_stress(double K, double G, const double E[9], double S[9], double dSdE[36]);
/* NeoHookean energy (src/Materials/NeoHookeanMaterial.jl:59-66) */
double or_nh_energy(double K, double G, const double E[9]);
/* Isotropic linear elastic sigma and dsigma/deps (IsotropicLinearElasticMaterial.jl:70-92) */
void or_linear_stress(double Emod, double nu, const double eps[9], double sig[9],
                      double dsde[36]);
/* Lame parameters from (E, nu) (IsotropicLinearElasticMaterial.jl:70-76, SVKMaterial.jl:50-54) */
void or_lame_from_E_nu(double Emod, double nu, double* lam, double* G);

/* ---- elements ----------------------------------------------------------- */
/* reference volume (src/Entities/Tetrahedrons.jl:114-138); OR_ERR_NEGATIVE_VOLUME if <= 0 */
int or_tet_volume(const double X[12], double* V);
/* shape-function gradients dN (4 nodes x 3, node-major) and volume (Tetrahedrons.jl:193-197) */
int or_tet_gradients(const double X[12], double dN[12], double* V);
/* hyperelastic tet: f, K, P = F S, C = F^T F (Tetrahedrons.jl:186-224).
 * mat = OR_MAT_SVK (p0 = lam, p1 = G) or OR_MAT_NEOHOOKEAN (p0 = K, p1 = G). */
int or_tet_hyperelastic(int mat, double p0, double p1, const double X[12], const double u[12],
                        double f[12], double K[144], double P[9], double C[9]);
/* linear tet: K = B^T C B V, f = K u, sigma, eps (Tetrahedrons.jl:240-266) */
int or_tet_linear(double Emod, double nu, const double X[12], const double u[12],
                  double f[12], double K[144], double sig[9], double eps[9]);
/* truss (src/Entities/Trusses.jl:126-184); Emod = elasticity_modulus(material) */
int or_truss(int strain_model, double Emod, double A, const double X[6], const double u[6],
             double f[6], double K[36], double* sigma11, double* eps11);
/* pressure on a triangular face: -p n A / 3 on each node
 * (src/BoundaryConditions/LocalLoadBoundaryConditions.jl:36-56,
 *  src/Entities/TriangularFaces.jl:44-65) */
void or_face_pressure(const double X[9], double p, double f[9]);

/* ---- global model ------------------------------------------------------- */
typedef struct {
    int64_t nn;              /* number of nodes                                  */
    const double* X;         /* 3*nn reference coordinates, node-major           */
    int64_t ntet;            /* tetrahedra                                       */
    const int32_t* tet;      /* 4*ntet node ids (0-based)                        */
    int tet_mat;             /* OR_MAT_*                                         */
    double tet_p0, tet_p1;   /* SVK: lam,G  NH: K,G  linear: E,nu                */
    const double* tet_G;     /* optional per-element shear modulus (NH/SVK G)    */
    int64_t ntruss;          /* trusses                                          */
    const int32_t* truss;    /* 2*ntruss node ids                                */
    int truss_strain;        /* OR_STRAIN_*                                      */
    double truss_E;          /* elasticity modulus                               */
    const double* truss_A;   /* ntruss areas                                     */
    int64_t nfree;           /* number of free dofs                              */
    const int64_t* free_dofs;/* ascending free dofs (src/StructuralModel/Structures.jl:129-142) */
} or_model;

/* CSR matrix owned by the oracle (free with or_csr_free). */
typedef struct {
    int64_t n;
    int64_t nnz;
    int64_t* rowptr;
    int32_t* col;
    double* val;
} or_csr;

void or_csr_free(or_csr* A);
void or_csr_matvec(const or_csr* A, const double* x, double* y);
/* `K[I[k], J[k]] += V[k]` for k in triplet order into an n x n CSR matrix
 * (end_assemble!, src/StructuralSolvers/Assemblers.jl:84-88); 0-based indices. */
int or_coo_to_csr(int64_t nt, const int32_t* I, const int32_t* J, const double* V, int64_t n,
                  or_csr* A);

/* assemble!(s, sa) (src/StructuralAnalyses/StaticAnalyses.jl:99-122): element
 * loop, f_int += f_e, COO triplets in the reference's column-major order
 * (src/StructuralSolvers/Assemblers.jl:46-67), then K[i,j] += v in triplet
 * order (Assemblers.jl:84-88).  Returns the full n x n tangent in CSR (Kfull,
 * may be NULL) and the free-free block (Kff, may be NULL).  Per-element
 * stress/strain (9 doubles each, tets first then trusses) when non-NULL. */
int or_assemble(const or_model* m, const double* U, double* fint, or_csr* Kfull, or_csr* Kff,
                double* stress, double* strain);

/* IterativeSolvers 0.9.4 cg! restated (called through LinearSolve 2.39.1 at
 * src/StructuralAnalyses/NonLinearStaticAnalyses.jl:129-134 with the tolerances of
 * src/StructuralSolvers/StructuralSolvers.jl:229-234): x0 = 0, unpreconditioned,
 * stop when ||r|| <= max(reltol ||b||, abstol) or after maxiter iterations. */
int64_t or_cg(const or_csr* A, const double* b, double* x, double abstol, double reltol,
              int64_t maxiter);

/* Newton record of one load step (src/StructuralSolvers/StructuralSolvers.jl:85-97) */
typedef struct {
    double dU_norm, dU_rel, r_norm, r_rel;
    int32_t iters;
    int32_t criterion;       /* 1 residual, 2 dU, 3 max-iter                     */
    int64_t cg_iters;        /* total CG iterations over the step                */
} or_step_record;

/* _solve!(NonLinearStaticAnalysis, NewtonRaphson)
 * (src/StructuralAnalyses/NonLinearStaticAnalyses.jl:70-148 with isconverged!
 * of src/StructuralSolvers/StructuralSolvers.jl:148-171).  fext_steps holds the
 * full external-force vector of every load step (nsteps x n).  U (n) is the
 * initial state on entry and the final state on exit; U_steps (nsteps x n,
 * may be NULL) receives the converged displacements of every step.
 * cg_reltol < 0 selects the reference default sqrt(eps). */
int or_newton_static(const or_model* m, int nsteps, const double* fext_steps,
                     double rel_U_tol, double rel_res_tol, int max_iter, double cg_reltol,
                     double* U, double* U_steps, or_step_record* records);

#ifdef __cplusplus
}
#endif
#endif
