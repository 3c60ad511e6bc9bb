/*
 * dfx.h -- C ABI of the MI355X-native Kohn-Sham eigensolver hot path
 * (gfx950 / CDNA4).  Plain pointers and sizes only; no exceptions cross this
 * boundary.  Every entry point returns an int status:
 *     0            success
 *   < 0            library error (DFX_E_*), message in dfx_ctx_last_error()
 *   > 0            the reference's own error code for that solver
 *                  (EigenSolverErrorCode, OrthonormalizationErrorCode or
 *                  LinearSolverErrorCode, src/linearAlgebra/LinearAlgebraTypes.h:84-125)
 *
 * Layout conventions (same as the reference):
 *  - multivectors are N x B row-major, vector index fastest: entry (dof i,
 *    vector j) at data[i*B + j]  (src/linearAlgebra/MultiVector.h:50-53);
 *  - cell matrices are concatenated per cell, cell e's n_e x n_e block
 *    column-major at offset sum_{e'<e} n_{e'}^2
 *    (src/ksdft/KohnShamOperatorContextFE.t.cpp:1194-1195);
 *  - cell->dof maps are concatenated per cell, classical dofs first, then
 *    enriched ids (+ nClassical)  (src/basis/FEBasisManager.h:116-145);
 *  - constraints are the flattened arrays of
 *    src/basis/CFEConstraintsLocalDealii.h:160-170 (rows, rowSizes,
 *    rowOffsets, cols, weights, inhomogeneities), rows ascending.
 *
 * Threading: calls on one dfx_ctx are serialized on that context's HIP
 * stream; different contexts may be driven from different host threads.
 */
#ifndef DFX_H
#define DFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dfx_ctx         dfx_ctx;
typedef struct dfx_mv          dfx_mv;
typedef struct dfx_constraints dfx_constraints;
typedef struct dfx_op          dfx_op;
typedef struct dfx_chfsi       dfx_chfsi;
typedef struct dfx_problem     dfx_problem;
typedef struct dfx_server      dfx_server;
typedef struct dfx_atom        dfx_atom;

/* ---- status codes ---------------------------------------------------- */
#define DFX_SUCCESS          0
#define DFX_E_ARG           (-1)
#define DFX_E_HIP           (-2)
#define DFX_E_OOM           (-3)
#define DFX_E_STATE         (-4)
#define DFX_E_UNSUPPORTED   (-5)
#define DFX_E_NONFINITE     (-6)

/* EigenSolverErrorCode (LinearAlgebraTypes.h:93-106) */
#define DFX_EIG_LAPACK_ERROR                   1
#define DFX_EIG_LANCZOS_BETA_ZERO              2
#define DFX_EIG_LANCZOS_SUBSPACE_INSUFFICIENT  3
#define DFX_EIG_CHFSI_ORTHONORMALIZATION_ERROR 4
#define DFX_EIG_CHFSI_RAYLEIGH_RITZ_ERROR      5
#define DFX_EIG_KS_MAX_PASS_ERROR              6
#define DFX_EIG_KS_CHFSI_ERROR                 7
#define DFX_EIG_KS_LANCZOS_ERROR               8
/* OrthonormalizationErrorCode (LinearAlgebraTypes.h:108-115) */
#define DFX_ORTHO_LAPACK_ERROR                      1
#define DFX_ORTHO_NON_ORTHONORMALIZABLE_MULTIVECTOR 2
/* LinearSolverErrorCode (LinearAlgebraTypes.h:84-91) */
#define DFX_LS_FAILED_TO_CONVERGE  1
#define DFX_LS_RESIDUAL_DIVERGENCE 2
#define DFX_LS_DIVISON_BY_ZERO     3
#define DFX_LS_OTHER_ERROR         4

/* ---- value types ----------------------------------------------------- */
#define DFX_F64 0
#define DFX_F32 1

/* ======================================================================
 * Context: one GPU, one HIP stream.
 * Replaces LinAlgOpContext<DEVICE> (src/linearAlgebra/LinAlgOpContext.h)
 * and the BlasQueue it carries (src/linearAlgebra/BlasLapackTypedef.h:188-211).
 * ==================================================================== */
int         dfx_ctx_create(int device, dfx_ctx **ctx);
int         dfx_ctx_destroy(dfx_ctx *ctx);
const char *dfx_ctx_last_error(const dfx_ctx *ctx);
int         dfx_ctx_synchronize(dfx_ctx *ctx);
/* the hipStream_t all work of this context is launched on */
int         dfx_ctx_stream(dfx_ctx *ctx, void **stream);
/* captured degree-loop graphs of this context: how many captures and replays
 * have run, and how many graphs are cached now (a graph is captured again
 * when an operator it launches on has reallocated a buffer, and released
 * with that operator).  No reference counterpart: the reference has no graph
 * capture; for tests and callers that want to see replays happen. */
int         dfx_ctx_graph_stats(const dfx_ctx *ctx, uint64_t *captures, uint64_t *replays,
                                uint32_t *cached);

/* ======================================================================
 * Device multivector: N x B, row-major, vector index fastest.
 * Replaces linearAlgebra::MultiVector<T, DEVICE> storage
 * (src/linearAlgebra/MultiVector.h:133-507).
 * ==================================================================== */
int dfx_mv_create(dfx_ctx *ctx, uint32_t N, uint32_t B, int dtype, dfx_mv **mv);
/* wrap caller-owned device memory (not freed by dfx_mv_destroy) */
int dfx_mv_wrap(dfx_ctx *ctx, void *device_ptr, uint32_t N, uint32_t B,
                int dtype, dfx_mv **mv);
int dfx_mv_destroy(dfx_mv *mv);
int dfx_mv_upload(dfx_mv *mv, const void *host);
int dfx_mv_download(const dfx_mv *mv, void *host);
int dfx_mv_device_ptr(dfx_mv *mv, void **ptr);
int dfx_mv_set(dfx_mv *mv, double value);
/* entry t = i*B + j set to a + (b - a) u_t, u_t the t-th SplitMix64 output
 * of `seed` mapped to [0,1) by its top 53 bits (replaces the reference's
 * time-seeded random initial subspace, src/utils/RandNumGen.cpp:40-43) */
int dfx_mv_fill_random(dfx_mv *mv, uint64_t seed, double a, double b);
/* dst = src with value-type conversion; same N and B */
int dfx_mv_copy(dfx_mv *dst, const dfx_mv *src);
/* Z = a X + b Y  (blasLapack::axpby, src/linearAlgebra/BlasLapack.t.cpp:228-237) */
int dfx_mv_axpby(dfx_mv *Z, double a, const dfx_mv *X, double b, const dfx_mv *Y);
/* Z[i,j] = a1 a[j] X[i,j] + b1 b[j] Y[i,j]  (blasLapack::axpbyBlocked,
 * src/linearAlgebra/BlasLapack.t.cpp:242-257); a, b host arrays of length B */
int dfx_mv_axpby_blocked(dfx_mv *Z, double a1, const double *a, const dfx_mv *X,
                         double b1, const double *b, const dfx_mv *Y);
/* column 2-norms and column dots, FP64 results on the host
 * (MultiVector::l2Norms / linearAlgebra::dot, src/linearAlgebra/MultiVector.h:620) */
int dfx_mv_l2norms(const dfx_mv *X, double *norms);
int dfx_mv_dots(const dfx_mv *X, const dfx_mv *Y, double *dots);

/* ======================================================================
 * Constraints (hanging-node + Dirichlet), device resident.
 * Replaces basis::ConstraintsLocal (src/basis/ConstraintsLocal.h:77-85) with
 * the flattened arrays of CFEConstraintsLocalDealii
 * (src/basis/CFEConstraintsLocalDealii.t.cpp:290-460).
 * distribute_child_to_parent is deterministic: each parent gathers its
 * children's contributions in row order (no atomics), which reproduces the
 * reference's sequential loop (src/basis/ConstraintsInternal.cpp:113-170).
 * ==================================================================== */
int dfx_constraints_create(dfx_ctx *ctx, uint32_t nDofs, uint32_t nRows,
                           const uint32_t *rows, const uint32_t *rowSizes,
                           const uint32_t *rowOffsets, const uint32_t *cols,
                           const double *weights, const double *inhom,
                           dfx_constraints **out);
int dfx_constraints_destroy(dfx_constraints *c);
/* x[row,:] = inhom + sum_j w_j x[col_j,:]  (ConstraintsInternal.cpp:38-107) */
int dfx_constraints_distribute_parent_to_child(dfx_constraints *c, dfx_mv *X);
/* x[col_j,:] += w_j x[row,:]; x[row,:] = 0  (ConstraintsInternal.cpp:113-170) */
int dfx_constraints_distribute_child_to_parent(dfx_constraints *c, dfx_mv *Y);
/* x[row,:] = value  (ConstraintsInternal.cpp:176-217) */
int dfx_constraints_set_constrained_nodes(dfx_constraints *c, dfx_mv *X, double value);
/* Download the device-resident index sets: rows (nRows), rowOffsets
 * (nRows + 1), cols (nEntries), and the parent-major transpose used by
 * child->parent: parents (nParents), parentOffsets (nParents + 1),
 * parentRows (nEntries, index into rows), parentEntries (nEntries, index into
 * cols/weights).  Pass NULL for arrays not wanted; sizes via dfx_constraints_sizes. */
int dfx_constraints_sizes(const dfx_constraints *c, uint32_t *nRows,
                          uint32_t *nEntries, uint32_t *nParents);
int dfx_constraints_export(const dfx_constraints *c, uint32_t *rows,
                           uint32_t *rowOffsets, uint32_t *cols,
                           uint32_t *parents, uint32_t *parentOffsets,
                           uint32_t *parentRows, uint32_t *parentEntries);

/* ======================================================================
 * Operators: OperatorContext<ValueTypeOperator, ValueTypeOperand, DEVICE>::apply
 * (src/linearAlgebra/OperatorContext.h:104-108).  apply(X, Y) overwrites Y
 * and may modify X (constrained rows of X are set by parent->child), as the
 * reference allows (OperatorContext.h:79-82).
 * ==================================================================== */
#define DFX_OP_KEEP_FP32 0x1u /* also keep an FP32 copy of the cell matrices */

/* Cell-wise operator Y = C^T (sum_e Q_e^T A_e Q_e) C X: KohnShamOperatorContextFE
 * (src/ksdft/KohnShamOperatorContextFE.h:102-127, apply .t.cpp:1326-1443),
 * CFEOverlapOperatorContext (src/basis/CFEOverlapOperatorContext.t.cpp:539-640)
 * and LaplaceOperatorContextFE (src/electrostatics/LaplaceOperatorContextFE.t.cpp:397-467).
 * cellMatrices: sum_e n_e^2 doubles (host).  nClassical: dofs >= nClassical are
 * enrichment dofs (shared by many cells; reduced deterministically). */
int dfx_op_create_cellwise(dfx_ctx *ctx, uint32_t nCells,
                           const uint32_t *cellNdofs, const uint32_t *cellDofMap,
                           const double *cellMatrices, uint32_t nDofs,
                           uint32_t nClassical, dfx_constraints *cons,
                           uint32_t flags, dfx_op **op);
/* replace the cell matrices (KohnShamOperatorContextFE::reinit, .t.cpp:1240-1310) */
int dfx_op_update_cells(dfx_op *op, const double *cellMatrices);

/* Diagonal (+ dense enriched block) operator
 *   Y = C2P( diag o P2C(X) + [0 ; E X_enr] )
 * as CFEOverlapInverseOpContextGLL::apply (src/basis/CFEOverlapInverseOpContextGLL.t.cpp:525-556)
 * and OrthoEFEOverlapInverseOpContextGLL::apply
 * (src/basis/OrthoEFEOverlapInverseOpContextGLL.t.cpp:1188-1273).  With
 * cons == NULL it is PreconditionerJacobi::apply (src/linearAlgebra/PreconditionerJacobi.t.cpp:56-80).
 * enrBlock: nEnriched x nEnriched column-major (may be NULL if nEnriched == 0);
 * enriched dofs are the last nEnriched of nDofs. */
int dfx_op_create_diagonal(dfx_ctx *ctx, uint32_t nDofs, const double *diag,
                           uint32_t nEnriched, const double *enrBlock,
                           dfx_constraints *cons, dfx_op **op);
int dfx_op_apply(dfx_op *op, dfx_mv *X, dfx_mv *Y);
/* dfx_op_apply with no device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No reference counterpart.  The enriched partial rows of a
 * cell-wise apply (nEnrichedSlots x B values) go to a per-GPU scratch array
 * of DFX_NOALLOC_SCRATCH_BYTES owned by the library (set up when a context
 * is created on the device, held until the process ends); larger applies are
 * refused with DFX_E_ARG.  Calls that overlap on one GPU share the array and
 * race: serialize them.  Results equal dfx_op_apply's bit for bit. */
#define DFX_NOALLOC_SCRATCH_BYTES (64u << 20)
int dfx_op_apply_noalloc(dfx_op *op, dfx_mv *X, dfx_mv *Y);
/* Y = A (BInv X): the residual Chebyshev filter's pair of applies
 * BInv.apply(ResidualNew, scratch1); A.apply(scratch1, scratch2)
 * (src/linearAlgebra/ChebyshevFilter.t.cpp:355-356).  When BInv is diagonal
 * and shares A's homogeneous Dirichlet constraints, BInv is folded into A's
 * cell gather (the same values bit for bit); otherwise the two applies run
 * in turn.  X receives BInv's parent->child update, as BInv.apply does. */
int dfx_op_apply_after_inverse(dfx_op *A, dfx_op *BInv, dfx_mv *X, dfx_mv *Y);
int dfx_op_destroy(dfx_op *op);

typedef struct dfx_op_info
{
  uint32_t nCells, nDofs, nClassical, nColors;
  uint32_t maxCellDofs;
  uint64_t sumNe;       /* sum_e n_e */
  uint64_t sumNe2;      /* sum_e n_e^2 */
  uint64_t matBytesF64; /* device bytes of the FP64 cell-matrix store */
  uint64_t matBytesF32; /* device bytes of the FP32 cell-matrix store (0 if none) */
} dfx_op_info;
int dfx_op_get_info(const dfx_op *op, dfx_op_info *info);
/* Download the device-resident cell->dof map in the caller's cell order
 * (bit-exact round trip of what dfx_op_create_cellwise uploaded). */
int dfx_op_export_map(const dfx_op *op, uint32_t *cellNdofs, uint32_t *cellDofMap);
/* Timing of this operator's fused cell-apply kernels per colour step (the
 * size-class launches of one colour, which run side by side on the
 * context's two streams): with profiling on, every colour step is bracketed
 * by HIP events on the context's stream.  get_profile synchronizes, and
 * reports the colour steps ("launches"), their summed time, and the
 * algorithmic flops (2 B sum n_e^2) and bytes (matrices + map + gathered X +
 * scattered Y) of their cells. */
typedef struct dfx_op_profile
{
  uint64_t launches;
  double   kernelMs;
  double   flops;
  double   bytes;
} dfx_op_profile;
/* enable bits 0-1: 0 off; 1 FP64 launches; 2 FP32 launches; 3 both.
 * bits 8-23: k, time every k-th apply of that kind only (0 or 1: every apply);
 * consecutive colour steps of an apply share one event between them. */
int dfx_op_set_profiling(dfx_op *op, int enable);
int dfx_op_get_profile(dfx_op *op, dfx_op_profile *out, int reset);
/* Column-chunk autotuner.  No counterpart in the reference, whose cell-block
 * sizes are fixed defaults (KohnShamOperatorContextFE.t.cpp:1370-1371).
 * Times the candidate column-chunk widths of the fused cell-apply kernel
 * (16 NT columns per chunk, NT in {1, 2, 4, 5, 8}) for every size class of a
 * cell-wise operator, for dtype and exactly B columns, on scratch device
 * multivectors (reps passes per candidate), and keeps the fastest width per
 * class for later applies of that dtype with B columns.  Results of an apply
 * do not depend on the width (columns are independent, same K order).
 * classes / widths (room for maxOut entries; may be NULL when maxOut is 0):
 * the class (M tiles per wave) and the chosen NT, in launch order; *nOut
 * (nullable) = number of classes.  B = 0 forgets the widths tuned for dtype.
 * DFX_CELL_NT_F32 / DFX_CELL_NT_F64 in the environment take precedence. */
int dfx_op_autotune(dfx_op *op, int dtype, uint32_t B, uint32_t reps, uint32_t maxOut,
                    uint32_t *classes, uint32_t *widths, uint32_t *nOut);
/* Stand-alone cell gather: Xcell[slot, :] = X[map[slot], :], Xcell is
 * (sum_e n_e) x B  (FECellWiseDataOperations::copyFieldToCellWiseData,
 * src/basis/FECellWiseDataOperations.t.cpp:31-84). */
int dfx_op_gather(dfx_op *op, dfx_mv *X, dfx_mv *Xcell);
/* Stand-alone cell scatter-add: Y[d, :] += sum of Ycell over the slots of d in
 * cell order (addCellWiseDataToFieldData, FECellWiseDataOperations.t.cpp:89-151);
 * bit-identical to the reference's sequential loop. */
int dfx_op_scatter_add(dfx_op *op, dfx_mv *Ycell, dfx_mv *Y);

/* ======================================================================
 * Chebyshev-filtered subspace iteration.
 * Replaces linearAlgebra::ChebyshevFilteredEigenSolver
 * (src/linearAlgebra/ChebyshevFilteredEigenSolver.h:90-122, solve .t.cpp:194-437).
 * ==================================================================== */
#define DFX_PREC_FP64  0 /* everything FP64 */
#define DFX_PREC_MIXED 1 /* filter bulk FP32 (FP32 MFMA, FP32 accumulate);
                            residual seed, CholGS, RR, residual norms FP64 */

typedef struct dfx_chfsi_params
{
  double   wantedSpectrumLowerBound;
  double   wantedSpectrumUpperBound;
  double   unWantedSpectrumUpperBound;
  uint32_t polynomialDegree;
  double   illConditionTolerance;
  int      isResidualChebyshevFilter; /* reference default: 1 */
  uint32_t eigenVectorBatchSize;      /* 0 -> all vectors in one block */
  int      precision;                 /* DFX_PREC_* */
  int      useGraph;                  /* capture the degree loop in a hipGraph */
  /* OrthogonalizationType (ChebyshevFilteredEigenSolver.h:45-50): only
   * DFX_ORTHO_CHOLESKY_GRAMSCHMIDT is implemented; the others are refused at
   * creation (DFX_E_UNSUPPORTED) */
  int      orthogonalizationType;
  /* keep copies of the filtered subspace and of its orthonormalized form
   * (storeIntermediateSubspaces, ChebyshevFilteredEigenSolver.t.cpp:70-78,
   * 336-381), read back with dfx_chfsi_get_filtered_subspace */
  int      storeIntermediateSubspaces;
  /* 0 (default): the solver keeps H X and M X of its eigenvectors (two
   * rotations of products Cholesky-Gram-Schmidt and Rayleigh-Ritz form
   * anyway; two N x n FP64 arrays).  A later solve whose guess has the same
   * content takes them for the residual filter's seed, and
   * dfx_chfsi_residual_norms for the norms, instead of applying H and M
   * again (equal up to FP64 rounding).  Needs constraints without
   * inhomogeneities; 1: apply H and M every time, as the reference does. */
  int      noProductReuse;
} dfx_chfsi_params;
#define DFX_ORTHO_CHOLESKY_GRAMSCHMIDT 0
#define DFX_ORTHO_MULTIPASS_LOWDIN     1
#define DFX_ORTHO_MULTIPASS_CGS        2

int dfx_chfsi_create(dfx_ctx *ctx, const dfx_chfsi_params *params,
                     dfx_mv *eigenSubspaceGuess, dfx_chfsi **out);
int dfx_chfsi_reinit(dfx_chfsi *s, double wantedLower, double wantedUpper,
                     double unwantedUpper, uint32_t degree,
                     double illConditionTolerance, dfx_mv *eigenSubspaceGuess);
/* eigenValues: host, length nVectors (in: filter seeds for the residual
 * filter; out: Ritz values).  eigenVectors: in/out device multivector. */
int dfx_chfsi_solve(dfx_chfsi *s, dfx_op *A, double *eigenValues,
                    dfx_mv *eigenVectors, int computeEigenVectors, dfx_op *B,
                    dfx_op *BInv);
/* getFilteredSubspace (orthonormalized = 0) / getOrthogonalizedFilteredSubspace
 * (orthonormalized = 1) of the last solve (ChebyshevFilteredEigenSolver.t.cpp:
 * 441-480), copied into `out` (FP64, the eigenvectors' shape); DFX_E_STATE
 * when the solver was created without storeIntermediateSubspaces */
int dfx_chfsi_get_filtered_subspace(const dfx_chfsi *s, int orthonormalized, dfx_mv *out);
/* wall times (ms) of the last solve: filter, orthonormalization, RR */
int dfx_chfsi_timings(const dfx_chfsi *s, double *ms3);
int dfx_chfsi_destroy(dfx_chfsi *s);

/* Cholesky-Gram-Schmidt (OrthonormalizationFunctions.t.cpp:153-431):
 * S = X^T B X, S = L L^T, Xo = X L^-T.  Returns OrthonormalizationErrorCode. */
int dfx_cholesky_gram_schmidt(dfx_ctx *ctx, dfx_mv *X, dfx_mv *Xo, dfx_op *B);
/* Rayleigh-Ritz (RayleighRitzEigenSolver.t.cpp:67-302): Hp = X^T A X,
 * eig(Hp) = Q diag(w) Q^T, Xo = X Q.  Returns EigenSolverErrorCode. */
int dfx_rayleigh_ritz(dfx_ctx *ctx, dfx_op *A, dfx_mv *X, double *eigenValues,
                      dfx_mv *Xo);

/* LanczosExtremeEigenSolver::solve, non-adaptive
 * (src/linearAlgebra/LanczosExtremeEigenSolver.t.cpp:224-540). */
int dfx_lanczos_extreme(dfx_ctx *ctx, dfx_op *A, dfx_op *B, dfx_op *BInv,
                        dfx_mv *guess, uint32_t maxKrylov, double *lowest,
                        double *highest, double *betaLast);

/* KohnShamEigenSolver::solve (src/ksdft/KohnShamEigenSolver.t.cpp:214-536):
 * Lanczos bounds, degree lookup, ChFSI passes, Fermi level by Newton-Raphson
 * on Fermi-Dirac occupancies (FractionalOccupancyFunction.cpp:9-71), residual
 * norms; converged when every state with occupancy > fracOccupancyTolerance
 * has residual <= eigenSolveResidualTolerance.  numElectrons == 0 counts every
 * state as occupied. */
typedef struct dfx_ks_params
{
  uint32_t numWantedEigenvalues;
  uint32_t maxChebyshevFilterPass;
  uint32_t polynomialDegree;          /* 0 -> CHEBY_ORDER_LOOKUP */
  double   eigenSolveResidualTolerance;
  uint32_t waveFunctionBatchSize;     /* 0 -> all */
  int      isResidualChebyshevFilter;
  int      precision;                 /* DFX_PREC_* */
  uint32_t lanczosMaxKrylov;          /* 0 -> 20 */
  double   numElectrons;
  double   smearingTemperature;       /* K */
  double   fermiEnergyTolerance;
  double   fracOccupancyTolerance;
  int      useGraph;                  /* capture the filter's degree loop in a hipGraph */
} dfx_ks_params;
typedef struct dfx_ks_result
{
  uint32_t passes;
  double   lanczosLow, lanczosHigh, lanczosBeta;
  double   a, b, c;                   /* last filter bounds */
  uint32_t degree;
  double   filterMs, orthoMs, rrMs, residualMs, lanczosMs;
  double   fermiEnergy;
  uint32_t numOccupied;               /* states with occupancy > tolerance */
} dfx_ks_result;
/* ||H x_j - lambda_j M x_j||_2 per column of X (FP64, host results):
 * KohnShamEigenSolver::getLinearEigenSolveResidual (src/ksdft/
 * KohnShamEigenSolver.t.cpp:574-684), the convergence test of every pass */
int dfx_eigen_residual_norms(dfx_ctx *ctx, dfx_op *H, dfx_op *M, dfx_mv *X,
                             const double *eigenValues, double *residualNorms);
/* The same norms for the eigenvectors X of the solver's last solve: from the
 * kept H X and M X when X still holds that solve's output (content hash) and
 * H, M are the operators it used; otherwise as dfx_eigen_residual_norms. */
int dfx_chfsi_residual_norms(dfx_chfsi *s, dfx_op *H, dfx_op *M, dfx_mv *X,
                             const double *eigenValues, double *residualNorms);
/* residualNorms: host, length numWantedEigenvalues (may be NULL) */
int dfx_ks_eigensolve(dfx_ctx *ctx, const dfx_ks_params *p, dfx_op *H,
                      dfx_op *M, dfx_op *MInv, dfx_op *MLanczos,
                      dfx_mv *waveFunctionGuess, dfx_mv *lanczosGuess,
                      double *energies, dfx_mv *waveFunctions,
                      double *residualNorms, dfx_ks_result *result);

/* CGLinearSolver::solve with a Jacobi (or any) preconditioner operator
 * (src/linearAlgebra/CGLinearSolver.t.cpp:68-313).  b, x: N x nRhs; x holds
 * the initial guess and receives the per-column converged solution.
 * Returns LinearSolverErrorCode. */
int dfx_cg_solve(dfx_ctx *ctx, dfx_op *A, dfx_op *pc, dfx_mv *b, dfx_mv *x,
                 double absTol, double relTol, double divTol, uint32_t maxIter,
                 uint32_t *iterations, double *residualNorms);

/* ======================================================================
 * Atom data files (the reference's enrichment XML format) and their
 * numerical radial functions.  Replaces AtomSphericalDataEnrichment (src/atoms/
 * AtomSphericalDataEnrichment.h:41-89, .cpp:511-640: libxml2 XPath reads of
 * /dft-efe:atom/dft-efe:<name>) and SphericalDataNumerical (src/atoms/
 * SphericalDataNumerical.h:40-140, .cpp:40-140, 323-420) for l <= 2.
 * Host-side reads; only dfx_atom_evaluate touches the GPU.
 * ==================================================================== */

/* Parse `path` and keep the functions of every listed field (e.g. "orbital",
 * "vnuclear"): per field element its qNumbers (n l m), cutoffInfo (cutoff,
 * smoothness) and NR radial values, each interpolated by the C^2 cubic spline
 * with zero end slopes (utils::Spline defaults, src/utils/Spline.h:87-95).
 * Errors (DFX_E_ARG, message in dfx_ctx_last_error(NULL)) as the reference's
 * checks: unparsable file, missing or repeated single elements, NR not
 * positive or not matching <r>, value counts, |m| > l, repeated quantum
 * numbers, incomplete m = -l..l shells; also a radial grid that is not
 * strictly increasing or has fewer than two points (the spline needs it). */
int dfx_atom_read(const char *path, const char *const *fieldNames, uint32_t nFields,
                  dfx_atom **out);
/* Pseudopotential file, UPF v2 (AtomSphericalDataPSP, src/atoms/
 * AtomSphericalDataPSP.h:47-140, .cpp:291-830): PP_HEADER attributes become
 * the metadata (plus "dij": D_ij in Ha when "beta" was read), PP_MESH/PP_R the
 * radial grid; fields "vlocal", "beta", "pswfc", "nlcc", "rhoatom" with the
 * reference's unit changes, divisions by r / 4 pi r^2 and Y_00 scaling;
 * quantum numbers {0,0,0} for vlocal / nlcc / rhoatom, {p, l, m} per
 * projector (p counts projectors of equal l in a run), {n, l, m} per
 * pseudo-wavefunction (n from its label).  vlocal is mixed: a spline of the
 * points within 10.0001 bohr (second-derivative condition on the left, slope
 * -v/r on the right) and -|z_valence| / r beyond.  Errors as the reference's
 * checks (header values, counts, repeated quantum numbers, l_max vs the
 * projectors, unknown field names). */
int dfx_atom_read_upf(const char *path, const char *const *fieldNames, uint32_t nFields,
                      dfx_atom **out);
/* The spline knots of function k (the radial grid except for a mixed
 * function), whether it is mixed, its tail charge (value tail / r beyond the
 * last knot, before the Y_00 factor) and its left quadratic coefficient. */
int dfx_atom_function_knots(const dfx_atom *a, const char *field, uint32_t k, uint32_t *n,
                            double *knots, int *mixed, double *tail, double *cLeft);
/* Trimmed text of the single child element `name` of <atom> (symbol, Z,
 * charge, NR, ...): AtomSphericalDataEnrichment::getMetadata (.h:74-75); for
 * a UPF file the PP_HEADER attribute `name` (AtomSphericalDataPSP::getMetadata).
 * Writes at most len bytes including the terminating 0; *needed = length + 1. */
int dfx_atom_metadata(const dfx_atom *a, const char *name, char *buf, uint64_t len,
                      uint64_t *needed);
/* The radial grid (NR points; r may be NULL to query NR). */
int dfx_atom_radial_grid(const dfx_atom *a, uint32_t *NR, double *r);
/* Functions of a field, in file order (getSphericalData, .h:56-62). */
int dfx_atom_num_functions(const dfx_atom *a, const char *field, uint32_t *count);
/* Function k of a field: quantum numbers, cutoff info, radial values and the
 * spline coefficients (value y_i + b_i h + c_i h^2 + d_i h^3 at h = r - r_i on
 * [r_i, r_i+1); left of r_0 y_0 + b_0 h; right of r_{NR-1} quadratic).  Any
 * output may be NULL; arrays hold one value per knot (NR for atom XML files;
 * dfx_atom_function_knots gives the count). */
int dfx_atom_function(const dfx_atom *a, const char *field, uint32_t k, int *qNumbers,
                      double *cutoff, double *smoothness, double *values, double *splineB,
                      double *splineC, double *splineD);
/* On the device: f(x) = R(r) Y_lm(theta, phi) f_c(r) of function k at every
 * point (points: FP64 nPts x 3), relative to origin (host, 3 values), and its
 * gradient (FP64 nPts x 3, may be NULL): SphericalDataNumerical::getValue /
 * getGradientValue (.cpp:330-420); a mixed function (UPF vlocal) as
 * SphericalDataMixed::getValue (src/atoms/SphericalDataMixed.cpp:43-69).
 * Zero past cutoff (1 + 1/smoothness); at the origin itself the value uses
 * theta = phi = 0 as the reference does and the gradient is 0 (the reference
 * asserts r > 0 there).  l > 2: DFX_E_UNSUPPORTED. */
int dfx_atom_evaluate(dfx_ctx *ctx, const dfx_atom *a, const char *field, uint32_t k,
                      const double *origin, dfx_mv *points, dfx_mv *values, dfx_mv *gradients);
int dfx_atom_destroy(dfx_atom *a);

/* ======================================================================
 * Problem builder: structured hex mesh, FE_Q(p) on GLL points, classical +
 * Slater-enriched basis, Dirichlet constraints and cell matrices
 * (H = 1/2 K + V, M, K), built on this GPU.  Stands in for the reference's
 * deal.II-backed setup (FEBasisManager::reinit, src/basis/FEBasisManager.t.cpp:
 * 118-260; FEBasisOperations::computeFEMatrices, src/basis/FEBasisOperations.t.cpp:
 * 41-433; EFE enrichment, src/basis/EnrichmentClassicalInterfaceSpherical.h:76-127).
 * Numbering: cells lexicographic (x fastest); nodes numbered by first
 * appearance over cells, cell-local order tensor-lexicographic; enriched dof
 * of atom a, function k is nClassical + (functions of atoms 0..a-1) + k, with
 * the analytic set 1s, 2s, 2p_y, 2p_z, 2p_x (five per atom) or the atom data's
 * functions in file order.
 * ==================================================================== */
typedef struct dfx_problem_spec
{
  uint32_t      p;
  uint32_t      n[3];
  double        lo[3], hi[3];
  uint32_t      nq;           /* Gauss points per direction, classical cells (0 -> p+1) */
  uint32_t      nqEnr;        /* Gauss points per direction, enriched cells (0 -> 2(p+1)) */
  uint32_t      nAtoms;
  const double *atomXYZ;      /* 3 nAtoms */
  const int    *atomZ;        /* nAtoms */
  double        vSmear;       /* V = -sum Z erf(r / vSmear) / r */
  int           enrich;       /* 1s, 2s, 2p Slater functions per atom */
  double        rcut, smooth; /* smooth cutoff, support ends at rcut (1 + 1/smooth) */
  int           orthogonalize;/* L2-orthogonalize enrichment against the classical space */
  /* numerical enrichment (enrich != 0): atom a carries the functions of field
   * enrichField ("orbital" when NULL) of atomData[a] (AtomSphericalDataContainer,
   * src/atoms/AtomSphericalDataContainer.h:41-98), each with its own cutoff;
   * NULL: the analytic Slater set above.  The handles must outlive the build. */
  const dfx_atom *const *atomData;
  const char   *enrichField;
  /* local potential: per atom a UPF file read with the field "vlocal" (its
   * local part, ElectrostaticLocalFE with pseudopotentials, src/ksdft/
   * ElectrostaticLocalFE.t.cpp) or NULL for the smeared Coulomb above; the
   * array itself may be NULL.  The handles must outlive the build. */
  const dfx_atom *const *localPotential;
} dfx_problem_spec;

#define DFX_PROB_CELL_NDOFS 0  /* uint32 [nCells] */
#define DFX_PROB_CELL_MAP   1  /* uint32 [sum n_e] */
#define DFX_PROB_ROWS       2  /* uint32 [nRows] */
#define DFX_PROB_ROW_SIZES  3  /* uint32 [nRows] */
#define DFX_PROB_ROW_OFFSETS 4 /* uint32 [nRows] */
#define DFX_PROB_COLS       5  /* uint32 [nEntries] */
#define DFX_PROB_WEIGHTS    6  /* double [nEntries] */
#define DFX_PROB_INHOM      7  /* double [nRows] */
#define DFX_PROB_H          8  /* double [sum n_e^2] */
#define DFX_PROB_M          9
#define DFX_PROB_K          10
#define DFX_PROB_DINV       11 /* double [nDofs] GLL overlap inverse diagonal */
#define DFX_PROB_MGLL       12 /* double [nDofs] GLL lumped mass */
#define DFX_PROB_ENR_MINV   13 /* double [nEnriched^2] (M^ee)^-1, column-major */
#define DFX_PROB_ENR_M      14 /* double [nEnriched^2] M^ee */
#define DFX_PROB_ALPHA      15 /* double [nEnriched * nClassical] orthogonalization coefficients */
#define DFX_PROB_NODE_XYZ   16 /* double [3 nClassical] node coordinates (x, y, z per node) */
#define DFX_PROB_QUAD_OFFSETS 17 /* uint64 [nCells + 1] cell c's quadrature points are
                                   [off[c], off[c + 1]), x fastest within a cell */
#define DFX_PROB_QUAD_W     18 /* double [off[nCells]] JxW of every quadrature point */
#define DFX_PROB_QUAD_XYZ   19 /* double [3 off[nCells]] coordinates of every quadrature
                                  point (x, y, z per point) */

int dfx_problem_build(dfx_ctx *ctx, const dfx_problem_spec *spec, dfx_problem **out);
/* sizes[8]: nCells, nDofs, nClassical, nEnriched, sum n_e, sum n_e^2, nRows, nEntries */
int dfx_problem_sizes(const dfx_problem *prob, uint64_t *sizes);
int dfx_problem_get(const dfx_problem *prob, int which, void *host);
/* Operators on the problem's device-resident cell matrices: H (KS), M
 * (consistent overlap), K (Laplace, for the Poisson CG), MInv (GLL overlap
 * inverse + enriched block), MLump (Lanczos B).  Any output may be NULL. */
int dfx_problem_create_operators(dfx_problem *prob, uint32_t flags, dfx_op **H,
                                 dfx_op **M, dfx_op **K, dfx_op **MInv,
                                 dfx_op **MLump, dfx_constraints **cons);
/* Electron density at every quadrature point of the problem's cells,
 * rho[q] = sum_i 2 occupation[i] psi_i(x_q)^2, psi_i(x_q) = sum_k N_k(x_q)
 * X[map_e[k], i], with the problem's own basis (orthogonalized enrichment
 * included) re-tabulated on the device.  Replaces ksdft::DensityCalculator::
 * computeRho (src/ksdft/DensityCalculator.h:41-122, .t.cpp:288-420) with the
 * interpolation of FEBasisOperations::interpolate (src/basis/
 * FEBasisOperations.t.cpp:996-1277); all states in one batch; X is used as
 * stored (no constraint distribution, as the reference).  X: FP64 nDofs x
 * nStates; occupation: host, nStates; rho: FP64, (quadrature points) x 1,
 * laid out by DFX_PROB_QUAD_OFFSETS. */
int dfx_problem_density(dfx_problem *prob, const double *occupation, dfx_mv *X,
                        dfx_mv *rho);
/* out = C2P(sum_e Q_e^T N_e^T (JxW o f_e)): the integral of quadrature values
 * f against every basis function of the problem (orthogonalized enrichment
 * included), added over cells in cell order, then the constraints' child ->
 * parent distribution (cons NULL: none).  Replaces
 * FEBasisOperations::integrateWithBasisValues (src/basis/
 * FEBasisOperations.t.cpp:1846-2190).  f: FP64, (quadrature points) x nComp,
 * laid out by DFX_PROB_QUAD_OFFSETS; out: FP64 nDofs x nComp. */
int dfx_problem_integrate(dfx_problem *prob, dfx_constraints *cons, dfx_mv *f,
                          dfx_mv *out);
int dfx_problem_destroy(dfx_problem *prob);

/* ======================================================================
 * Poisson linear-solver function (electrostatics, one GPU).
 * Replaces electrostatics::PoissonLinearSolverFunctionFE
 * (src/electrostatics/PoissonLinearSolverFunctionFE.h, .t.cpp:38-660) and
 * its LinearSolverFunction interface (src/linearAlgebra/
 * LinearSolverFunction.h:79-105): getAxContext / getPCContext / getRhs /
 * getInitialGuess / setSolution / getSolution.
 *  - K: cell-wise stiffness operator on the homogeneous constraints (the
 *    Ax context, LaplaceOperatorContextFE on the homogeneous basis manager);
 *  - field: the field's constraints, Dirichlet values in the
 *    inhomogeneities (feBasisManagerField);
 *  - rhsQuad: FP64 (quadrature points of prob) x nComponents (inpRhs);
 *  - pcType: DFX_PC_JACOBI (assembled, condensed diagonal; constrained rows
 *    1) or DFX_PC_NONE.
 * b = integrate(rhsQuad) - C2P_hom(K P2C_field(0)).  prob and K must outlive
 * the object.
 * ==================================================================== */
typedef struct dfx_poisson dfx_poisson;
#define DFX_PC_NONE   0
#define DFX_PC_JACOBI 1
int dfx_poisson_create(dfx_problem *prob, dfx_op *K, dfx_constraints *field,
                       dfx_mv *rhsQuad, int pcType, dfx_poisson **out);
/* new right-hand side (and optionally new field constraints): reinit */
int dfx_poisson_reinit(dfx_poisson *p, dfx_constraints *field, dfx_mv *rhsQuad);
/* borrowed handles, valid while p lives */
int dfx_poisson_get_ax(dfx_poisson *p, dfx_op **A);
int dfx_poisson_get_pc(dfx_poisson *p, dfx_op **pc);
/* copies into caller multivectors (FP64, nDofs x nComponents) */
int dfx_poisson_get_rhs(dfx_poisson *p, dfx_mv *b);
int dfx_poisson_get_initial_guess(dfx_poisson *p, dfx_mv *x);
int dfx_poisson_set_solution(dfx_poisson *p, const dfx_mv *x);
/* solution = P2C_field(x + u_D); it becomes the next initial guess */
int dfx_poisson_get_solution(dfx_poisson *p, dfx_mv *solution);
/* CGLinearSolver::solve(function) (src/linearAlgebra/CGLinearSolver.t.cpp:
 * 68-313) followed by getSolution: returns the LinearSolverErrorCode */
int dfx_poisson_solve(dfx_poisson *p, uint32_t maxIter, double absTol, double relTol,
                      double divTol, dfx_mv *solution, uint32_t *iterations,
                      double *residualNorms);
int dfx_poisson_destroy(dfx_poisson *p);

/* ======================================================================
 * Serving: a stream of independent Kohn-Sham eigenproblems on one GPU with
 * continuous batching (no counterpart in the reference; each request is the
 * problem build plus KohnShamEigenSolver::solve,
 * src/ksdft/KohnShamEigenSolver.t.cpp:214-536).  Requests are admitted in
 * arrival order when they have arrived, a slot is free and their estimated
 * device footprint fits in memoryBudgetBytes; each admitted request runs on
 * its own context / HIP stream, so the work of all admitted requests
 * overlaps on the device.  Results do not depend on the mix.
 * ==================================================================== */
typedef struct dfx_request
{
  dfx_problem_spec spec;       /* mesh, atoms, enrichment (caller-owned arrays) */
  uint32_t numStates;
  double   tolerance;          /* residual tolerance of the occupied states */
  double   numElectrons;
  uint32_t maxPasses;
  uint64_t guessSeed;          /* SplitMix64 seeds of the subspace (seed) and Lanczos (seed+1) */
  double   arrivalSeconds;     /* from the start of the stream */
} dfx_request;
typedef struct dfx_request_result
{
  int      status;             /* 0, an EigenSolverErrorCode (> 0) or DFX_E_* */
  uint32_t passes;
  double   admitSeconds;       /* from the start of the stream */
  double   doneSeconds;
  uint64_t footprintBytes;
  uint32_t slot;
} dfx_request_result;
typedef struct dfx_serve_params
{
  int      device;
  uint32_t maxSlots;           /* requests in flight (one host thread + stream each) */
  uint64_t memoryBudgetBytes;  /* admission limit on the summed footprints */
  int      precision;          /* DFX_PREC_* */
  int      useGraph;           /* replay each request's captured degree loop (dfx_ks_params) */
} dfx_serve_params;
/* estimated peak device bytes of one request (what admission counts) */
int dfx_request_footprint(const dfx_request *r, uint64_t *bytes);
/* runs the whole stream; energies: host, sum of numStates doubles in request
 * order; results: one per request.  Returns the first library error (< 0) of
 * any request, else 0 (per-request solver codes are in results). */
int dfx_serve(const dfx_serve_params *p, uint32_t nRequests, const dfx_request *reqs,
              double *energies, dfx_request_result *results);

#ifdef __cplusplus
}
#endif
#endif /* DFX_H */
