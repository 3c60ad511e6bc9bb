/* Drop-in classes for dft-efe's eigensolver hot path, backed by libdfx.
 *
 * A header for dft-efe's own build: it includes the reference's interface
 * headers (OperatorContext, MultiVector, HermitianIterativeEigenSolver,
 * LinearSolverFunction, LinearSolverImpl, ...) from its source tree and
 * derives from them, so the reference's callers (ChebyshevFilter*,
 * RayleighRitzEigenSolver, KohnShamEigenSolver, CGLinearSolver users) take
 * these objects unchanged.  Only the C ABI of include/dfx.h is called.
 *
 * Classes (namespace dftefe::linearAlgebra):
 *   DfxCellOperatorContext        OperatorContext::apply (OperatorContext.h:
 *                                 104-108) for KohnShamOperatorContextFE /
 *                                 CFEOverlapOperatorContext /
 *                                 LaplaceOperatorContextFE, reinit
 *                                 (KohnShamOperatorContextFE.h:102-127)
 *   DfxDiagonalOperatorContext    CFEOverlapInverseOpContextGLL,
 *                                 OrthoEFEOverlapInverseOpContextGLL,
 *                                 PreconditionerJacobi
 *   DfxChebyshevFilteredEigenSolver  ChebyshevFilteredEigenSolver, same
 *                                 constructor (ChebyshevFilteredEigenSolver.h:
 *                                 90-102), reinit, solve, getFilteredSubspace,
 *                                 getOrthogonalizedFilteredSubspace
 *   DfxPoissonLinearSolverFunction  LinearSolverFunction (LinearSolverFunction.h:
 *                                 79-105) of PoissonLinearSolverFunctionFE
 *   DfxCGLinearSolver             LinearSolverImpl::solve (LinearSolverImpl.h:
 *                                 79-82) with CGLinearSolver's constructor
 *                                 (CGLinearSolver.h:90-94)
 *
 * Multivectors stay where dft-efe keeps them: device MultiVectors are
 * borrowed through dfx_mv_wrap (no copies).  Operators built here run on one
 * process-wide libdfx context per device (dfxContext()); the
 * ElpaScalapackManager argument is accepted and unused on one GPU.  Library
 * errors (< 0) are raised with utils::throwException; solver outcomes (> 0)
 * come back in the reference's error structs with its own numbering.
 */
#pragma once

#include <dfx.h>

#include <linearAlgebra/BlasLapackTypedef.h>
#include <linearAlgebra/ChebyshevFilteredEigenSolver.h>
#include <linearAlgebra/ElpaScalapackManager.h>
#include <linearAlgebra/HermitianIterativeEigenSolver.h>
#include <linearAlgebra/IdentityOperatorContext.h>
#include <linearAlgebra/LinearAlgebraProfiler.h>
#include <linearAlgebra/LinearAlgebraTypes.h>
#include <linearAlgebra/LinearSolverFunction.h>
#include <linearAlgebra/LinearSolverImpl.h>
#include <linearAlgebra/MultiVector.h>
#include <linearAlgebra/OperatorContext.h>
#include <utils/Exceptions.h>
#include <utils/MPITypes.h>
#include <utils/MemorySpaceType.h>
#include <utils/TypeConfig.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace dftefe
{
  namespace linearAlgebra
  {
    constexpr utils::MemorySpace dfxMem = utils::MemorySpace::DEVICE;

    // ------------------------------------------------------------ context
    // One libdfx context (HIP stream, rocBLAS handle) per device for the
    // process, created on first use and kept until exit.
    inline dfx_ctx *
    dfxContext(int device = 0)
    {
      static std::mutex             m;
      static std::vector<dfx_ctx *> ctxs;
      std::lock_guard<std::mutex>   lk(m);
      if ((size_t)device >= ctxs.size())
        ctxs.resize(device + 1, nullptr);
      if (!ctxs[device])
        {
          dfx_ctx *c  = nullptr;
          const int rc = dfx_ctx_create(device, &c);
          utils::throwException(rc == 0, std::string("libdfx: no context on device ") +
                                            std::to_string(device) + ": " +
                                            dfx_ctx_last_error(nullptr));
          ctxs[device] = c;
        }
      return ctxs[device];
    }

    inline void
    dfxCheck(dfx_ctx *ctx, int rc)
    {
      if (rc < 0)
        utils::throwException(false, std::string("libdfx error ") + std::to_string(rc) +
                                       ": " + dfx_ctx_last_error(ctx));
    }

    template <typename T>
    constexpr int
    dfxType()
    {
      static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                    "libdfx multivectors are real FP64 or FP32");
      return std::is_same<T, double>::value ? DFX_F64 : DFX_F32;
    }

    // A device MultiVector borrowed as a dfx_mv for one call (no copy):
    // locally owned rows, vector index fastest (MultiVector.h:50-53).
    template <typename T>
    class DfxView
    {
    public:
      DfxView(dfx_ctx *ctx, MultiVector<T, dfxMem> &X)
      {
        dfxCheck(ctx, dfx_mv_wrap(ctx, X.data(), (uint32_t)X.locallyOwnedSize(),
                                  (uint32_t)X.getNumberComponents(), dfxType<T>(), &d_mv));
      }
      DfxView(dfx_ctx *ctx, const MultiVector<T, dfxMem> &X)
        : DfxView(ctx, const_cast<MultiVector<T, dfxMem> &>(X))
      {}
      ~DfxView() { dfx_mv_destroy(d_mv); }
      DfxView(const DfxView &) = delete;
      DfxView &operator=(const DfxView &) = delete;
      dfx_mv *get() const { return d_mv; }

    private:
      dfx_mv *d_mv = nullptr;
    };

    // Common base of the libdfx-backed operators: the handle the solvers
    // below hand to the library.
    class DfxOperatorHandle
    {
    public:
      virtual ~DfxOperatorHandle() = default;
      virtual dfx_op *handle() const = 0;
    };

    template <typename ValueTypeOperator, typename ValueTypeOperand>
    const DfxOperatorHandle &
    dfxOperator(const OperatorContext<ValueTypeOperator, ValueTypeOperand, dfxMem> &o)
    {
      auto *d = dynamic_cast<const DfxOperatorHandle *>(&o);
      utils::throwException(d != nullptr,
                            "libdfx solvers need operator contexts built on libdfx");
      return *d;
    }

    // ----------------------------------------------------- operators
    // Cell-wise operator Y = C^T (sum_e Q_e^T A_e Q_e) C X: the Kohn-Sham
    // Hamiltonian, the overlap M or the Laplace K.  cellNdofs / cellDofMap
    // as FEBasisManager::nLocallyOwnedCellDofs / locallyOwnedCellLocalDofIdsBegin
    // (FEBasisManager.h:116-145); cellMatrices the Hamiltonian terms'
    // getLocal() storage summed per cell (KohnShamOperatorContextFE.t.cpp:
    // 1264-1282), column-major, concatenated; constraints from
    // dfx_constraints_create (the arrays of CFEConstraintsLocalDealii.h:160-170).
    template <typename ValueTypeOperator, typename ValueTypeOperand>
    class DfxCellOperatorContext
      : public OperatorContext<ValueTypeOperator, ValueTypeOperand, dfxMem>,
        public DfxOperatorHandle
    {
    public:
      using ValueTypeUnion = blasLapack::scalar_type<ValueTypeOperator, ValueTypeOperand>;

      DfxCellOperatorContext(const std::vector<size_type> &cellNdofs,
                             const std::vector<size_type> &cellDofMap,
                             const std::vector<double> &cellMatrices, size_type nLocal,
                             size_type nClassical, dfx_constraints *constraints,
                             bool keepFp32 = false, int device = 0)
        : d_ctx(dfxContext(device))
      {
        static_assert(sizeof(size_type) == sizeof(uint32_t), "size_type is 32-bit");
        dfxCheck(d_ctx,
                 dfx_op_create_cellwise(d_ctx, (uint32_t)cellNdofs.size(),
                                        reinterpret_cast<const uint32_t *>(cellNdofs.data()),
                                        reinterpret_cast<const uint32_t *>(cellDofMap.data()),
                                        cellMatrices.data(), nLocal, nClassical, constraints,
                                        keepFp32 ? DFX_OP_KEEP_FP32 : 0, &d_op));
      }
      // an operator handle owned elsewhere (not destroyed here)
      DfxCellOperatorContext(dfx_ctx *ctx, dfx_op *borrowed)
        : d_ctx(ctx), d_op(borrowed), d_owns(false)
      {}
      ~DfxCellOperatorContext()
      {
        if (d_owns)
          dfx_op_destroy(d_op);
      }
      DfxCellOperatorContext(const DfxCellOperatorContext &) = delete;
      DfxCellOperatorContext &operator=(const DfxCellOperatorContext &) = delete;

      // KohnShamOperatorContextFE::reinit: new cell matrices (an SCF step's
      // local potential), same mesh and maps
      void
      reinit(const std::vector<double> &cellMatrices)
      {
        dfxCheck(d_ctx, dfx_op_update_cells(d_op, cellMatrices.data()));
      }

      // Y overwritten; the constrained rows of X are set in place
      // (parent -> child), as OperatorContext::apply allows
      void
      apply(MultiVector<ValueTypeOperand, dfxMem> &X,
            MultiVector<ValueTypeUnion, dfxMem> &  Y,
            bool                                   updateGhostX = false,
            bool                                   updateGhostY = false) const override
      {
        (void)updateGhostX;
        (void)updateGhostY; // one process: no ghost rows
        DfxView<ValueTypeOperand> x(d_ctx, X);
        DfxView<ValueTypeUnion>   y(d_ctx, Y);
        dfxCheck(d_ctx, dfx_op_apply(d_op, x.get(), y.get()));
      }

      dfx_op *handle() const override { return d_op; }

    private:
      dfx_ctx *d_ctx;
      dfx_op  *d_op   = nullptr;
      bool     d_owns = true;
    };

    // Y = C2P(d o P2C(X) + [0; E X_enr]): the GLL overlap inverse (classical
    // diagonal plus the dense enriched block, row-major nEnriched^2 as
    // OrthoEFEOverlapInverseOpContextGLL.t.cpp:1245-1260 applies it) or, with
    // no constraints, the Jacobi preconditioner's inverse diagonal.
    template <typename ValueTypeOperator, typename ValueTypeOperand>
    class DfxDiagonalOperatorContext
      : public OperatorContext<ValueTypeOperator, ValueTypeOperand, dfxMem>,
        public DfxOperatorHandle
    {
    public:
      using ValueTypeUnion = blasLapack::scalar_type<ValueTypeOperator, ValueTypeOperand>;

      DfxDiagonalOperatorContext(const std::vector<double> &diagonal, size_type nEnriched,
                                 const std::vector<double> &enrichedBlock,
                                 dfx_constraints *constraints, int device = 0)
        : d_ctx(dfxContext(device))
      {
        dfxCheck(d_ctx, dfx_op_create_diagonal(d_ctx, (uint32_t)diagonal.size(),
                                               diagonal.data(), nEnriched,
                                               nEnriched ? enrichedBlock.data() : nullptr,
                                               constraints, &d_op));
      }
      DfxDiagonalOperatorContext(dfx_ctx *ctx, dfx_op *borrowed)
        : d_ctx(ctx), d_op(borrowed), d_owns(false)
      {}
      ~DfxDiagonalOperatorContext()
      {
        if (d_owns)
          dfx_op_destroy(d_op);
      }
      DfxDiagonalOperatorContext(const DfxDiagonalOperatorContext &) = delete;
      DfxDiagonalOperatorContext &operator=(const DfxDiagonalOperatorContext &) = delete;

      void
      apply(MultiVector<ValueTypeOperand, dfxMem> &X,
            MultiVector<ValueTypeUnion, dfxMem> &  Y,
            bool                                   updateGhostX = false,
            bool                                   updateGhostY = false) const override
      {
        (void)updateGhostX;
        (void)updateGhostY;
        DfxView<ValueTypeOperand> x(d_ctx, X);
        DfxView<ValueTypeUnion>   y(d_ctx, Y);
        dfxCheck(d_ctx, dfx_op_apply(d_op, x.get(), y.get()));
      }

      dfx_op *handle() const override { return d_op; }

    private:
      dfx_ctx *d_ctx;
      dfx_op  *d_op   = nullptr;
      bool     d_owns = true;
    };

    // ------------------------------------------------------- eigensolver
    template <typename ValueTypeOperator, typename ValueTypeOperand>
    class DfxChebyshevFilteredEigenSolver
      : public HermitianIterativeEigenSolver<ValueTypeOperator, ValueTypeOperand, dfxMem>
    {
    public:
      using Base      = HermitianIterativeEigenSolver<ValueTypeOperator, ValueTypeOperand, dfxMem>;
      using ValueType = blasLapack::scalar_type<ValueTypeOperator, ValueTypeOperand>;
      using RealType  = blasLapack::real_type<ValueType>;
      using OpContext = typename Base::OpContext;
      static_assert(std::is_same<ValueType, double>::value &&
                      std::is_same<ValueTypeOperand, double>::value,
                    "the libdfx eigensolver works on real FP64 subspaces");

      // ChebyshevFilteredEigenSolver.h:90-102.  orthoType: CHOLESKY_GRAMSCHMIDT
      // (MULTIPASS_LOWDIN is refused as by the reference, "Orthogonalization
      // type not present"; MULTIPASS_CGS is not implemented here and refused).
      DfxChebyshevFilteredEigenSolver(
        const double wantedSpectrumLowerBound, const double wantedSpectrumUpperBound,
        const double unWantedSpectrumUpperBound, const double polynomialDegree,
        const double illConditionTolerance,
        MultiVector<ValueTypeOperand, dfxMem> &eigenSubspaceGuess,
        const ElpaScalapackManager &elpaScala, bool isResidualChebyshevFilter = true,
        const size_type       eigenVectorBatchSize = 0,
        OrthogonalizationType orthoType = OrthogonalizationType::CHOLESKY_GRAMSCHMIDT,
        bool storeIntermediateSubspaces = false, int device = 0)
        : d_ctx(dfxContext(device)), d_store(storeIntermediateSubspaces)
      {
        (void)elpaScala; // dense factorizations run with rocSOLVER on the GPU
        d_p                            = dfx_chfsi_params{};
        d_p.wantedSpectrumLowerBound   = wantedSpectrumLowerBound;
        d_p.wantedSpectrumUpperBound   = wantedSpectrumUpperBound;
        d_p.unWantedSpectrumUpperBound = unWantedSpectrumUpperBound;
        d_p.polynomialDegree           = (uint32_t)polynomialDegree;
        d_p.illConditionTolerance      = illConditionTolerance;
        d_p.isResidualChebyshevFilter  = isResidualChebyshevFilter ? 1 : 0;
        d_p.eigenVectorBatchSize       = eigenVectorBatchSize;
        d_p.precision                  = DFX_PREC_FP64;
        d_p.orthogonalizationType =
          orthoType == OrthogonalizationType::CHOLESKY_GRAMSCHMIDT ?
            DFX_ORTHO_CHOLESKY_GRAMSCHMIDT :
            (orthoType == OrthogonalizationType::MULTIPASS_LOWDIN ?
               DFX_ORTHO_MULTIPASS_LOWDIN :
               DFX_ORTHO_MULTIPASS_CGS);
        d_p.storeIntermediateSubspaces = storeIntermediateSubspaces ? 1 : 0;
        create(eigenSubspaceGuess);
      }

      ~DfxChebyshevFilteredEigenSolver() { release(); }
      DfxChebyshevFilteredEigenSolver(const DfxChebyshevFilteredEigenSolver &) = delete;
      DfxChebyshevFilteredEigenSolver &
      operator=(const DfxChebyshevFilteredEigenSolver &) = delete;

      // FP32 filter bulk with FP64 seed / CholGS / RR / residuals (no
      // reference counterpart: the reference filters in the operand type)
      void
      setMixedPrecision(bool mixed)
      {
        d_p.precision = mixed ? DFX_PREC_MIXED : DFX_PREC_FP64;
        release();
        create(*d_guessMv);
      }

      // ChebyshevFilteredEigenSolver.h:108-114
      void
      reinit(const double wantedSpectrumLowerBound, const double wantedSpectrumUpperBound,
             const double unWantedSpectrumUpperBound, const double polynomialDegree,
             const double                           illConditionTolerance,
             MultiVector<ValueTypeOperand, dfxMem> &eigenSubspaceGuess)
      {
        d_p.wantedSpectrumLowerBound   = wantedSpectrumLowerBound;
        d_p.wantedSpectrumUpperBound   = wantedSpectrumUpperBound;
        d_p.unWantedSpectrumUpperBound = unWantedSpectrumUpperBound;
        d_p.polynomialDegree           = (uint32_t)polynomialDegree;
        d_p.illConditionTolerance      = illConditionTolerance;
        if (&eigenSubspaceGuess != d_guessMv)
          {
            auto g = std::make_unique<DfxView<ValueTypeOperand>>(d_ctx, eigenSubspaceGuess);
            dfxCheck(d_ctx, dfx_chfsi_reinit(d_s, wantedSpectrumLowerBound,
                                             wantedSpectrumUpperBound,
                                             unWantedSpectrumUpperBound,
                                             d_p.polynomialDegree, illConditionTolerance,
                                             g->get()));
            d_guess   = std::move(g);
            d_guessMv = &eigenSubspaceGuess;
          }
        else
          dfxCheck(d_ctx, dfx_chfsi_reinit(d_s, wantedSpectrumLowerBound,
                                           wantedSpectrumUpperBound,
                                           unWantedSpectrumUpperBound, d_p.polynomialDegree,
                                           illConditionTolerance, nullptr));
      }

      // HermitianIterativeEigenSolver::solve (HermitianIterativeEigenSolver.h:
      // 76-86): A, B and BInv must be libdfx operator contexts
      EigenSolverError
      solve(const OpContext &A, std::vector<RealType> &eigenValues,
            MultiVector<ValueType, dfxMem> &eigenVectors, bool computeEigenVectors = false,
            const OpContext &B =
              IdentityOperatorContext<ValueTypeOperator, ValueTypeOperand, dfxMem>(),
            const OpContext &BInv =
              IdentityOperatorContext<ValueTypeOperator, ValueTypeOperand, dfxMem>()) override
      {
        utils::throwException(eigenValues.size() == eigenVectors.getNumberComponents(),
                              "eigenValues must hold one value per eigenvector");
        DfxView<ValueType> X(d_ctx, eigenVectors);
        const int rc = dfx_chfsi_solve(d_s, dfxOperator(A).handle(), eigenValues.data(),
                                       X.get(), computeEigenVectors ? 1 : 0,
                                       dfxOperator(B).handle(), dfxOperator(BInv).handle());
        dfxCheck(d_ctx, rc);
        if (d_store && rc == 0)
          {
            keep(d_filtered, eigenVectors, 0);
            keep(d_filteredOrtho, eigenVectors, 1);
          }
        return EigenSolverErrorMsg::isSuccessAndMsg(static_cast<EigenSolverErrorCode>(rc));
      }

      // ChebyshevFilteredEigenSolver.t.cpp:441-480
      MultiVector<ValueType, dfxMem> &
      getFilteredSubspace()
      {
        return stored(d_filtered);
      }
      MultiVector<ValueType, dfxMem> &
      getOrthogonalizedFilteredSubspace()
      {
        return stored(d_filteredOrtho);
      }

    private:
      void
      create(MultiVector<ValueTypeOperand, dfxMem> &guess)
      {
        auto g = std::make_unique<DfxView<ValueTypeOperand>>(d_ctx, guess);
        dfxCheck(d_ctx, dfx_chfsi_create(d_ctx, &d_p, g->get(), &d_s));
        d_guess   = std::move(g);
        d_guessMv = &guess;
      }
      void
      release()
      {
        if (d_s)
          dfx_chfsi_destroy(d_s);
        d_s = nullptr;
      }
      void
      keep(std::shared_ptr<MultiVector<ValueType, dfxMem>> &dst,
           const MultiVector<ValueType, dfxMem> &like, int ortho)
      {
        if (!dst)
          dst = std::make_shared<MultiVector<ValueType, dfxMem>>(like, (ValueType)0);
        DfxView<ValueType> v(d_ctx, *dst);
        dfxCheck(d_ctx, dfx_chfsi_get_filtered_subspace(d_s, ortho, v.get()));
      }
      MultiVector<ValueType, dfxMem> &
      stored(std::shared_ptr<MultiVector<ValueType, dfxMem>> &v)
      {
        utils::throwException(d_store && v != nullptr,
                              "storeIntermediateSubspaces is false in CHFSI class. Cannot "
                              "return the filtered Subspace.");
        return *v;
      }

      dfx_ctx                                   *d_ctx;
      dfx_chfsi_params                           d_p{};
      dfx_chfsi                                 *d_s = nullptr;
      std::unique_ptr<DfxView<ValueTypeOperand>> d_guess;
      MultiVector<ValueTypeOperand, dfxMem>     *d_guessMv = nullptr;
      bool                                       d_store;
      std::shared_ptr<MultiVector<ValueType, dfxMem>> d_filtered, d_filteredOrtho;
    };

    // ------------------------------------------------------ Poisson / CG
    // LinearSolverFunction of PoissonLinearSolverFunctionFE (LinearSolverFunction.h:
    // 79-105) over a dfx_poisson: the right-hand side (integral of the
    // quadrature values minus the stiffness on the Dirichlet lift), the
    // Jacobi preconditioner, the initial guess and the solution live in the
    // library; rhs / initialGuess are caller MultiVectors of nDofs x
    // nComponents that receive copies for getRhs / getInitialGuess.
    class DfxPoissonLinearSolverFunction : public LinearSolverFunction<double, double, dfxMem>
    {
    public:
      DfxPoissonLinearSolverFunction(dfx_problem *problem, dfx_op *stiffness,
                                     dfx_constraints *fieldConstraints, dfx_mv *rhsQuad,
                                     int pcType, MultiVector<double, dfxMem> &rhs,
                                     MultiVector<double, dfxMem> &initialGuess,
                                     const utils::mpi::MPIComm &comm, int device = 0)
        : d_ctx(dfxContext(device)), d_rhs(&rhs), d_init(&initialGuess), d_comm(comm)
      {
        dfxCheck(d_ctx, dfx_poisson_create(problem, stiffness, fieldConstraints, rhsQuad,
                                           pcType, &d_fn));
        dfx_op *a = nullptr, *pc = nullptr;
        dfx_poisson_get_ax(d_fn, &a);
        dfx_poisson_get_pc(d_fn, &pc);
        d_ax = std::make_unique<DfxCellOperatorContext<double, double>>(d_ctx, a);
        d_pc = std::make_unique<DfxDiagonalOperatorContext<double, double>>(d_ctx, pc);
        refresh();
      }
      ~DfxPoissonLinearSolverFunction() { dfx_poisson_destroy(d_fn); }
      DfxPoissonLinearSolverFunction(const DfxPoissonLinearSolverFunction &) = delete;
      DfxPoissonLinearSolverFunction &
      operator=(const DfxPoissonLinearSolverFunction &) = delete;

      // PoissonLinearSolverFunctionFE::reinit (.t.cpp:303-487)
      void
      reinit(dfx_constraints *fieldConstraints, dfx_mv *rhsQuad)
      {
        dfxCheck(d_ctx, dfx_poisson_reinit(d_fn, fieldConstraints, rhsQuad));
        refresh();
      }

      const OperatorContext<double, double, dfxMem> &
      getAxContext() const override
      {
        return *d_ax;
      }
      const OperatorContext<double, double, dfxMem> &
      getPCContext() const override
      {
        return *d_pc;
      }
      void
      setSolution(const MultiVector<double, dfxMem> &x) override
      {
        DfxView<double> v(d_ctx, x);
        dfxCheck(d_ctx, dfx_poisson_set_solution(d_fn, v.get()));
      }
      void
      getSolution(MultiVector<double, dfxMem> &solution) override
      {
        DfxView<double> v(d_ctx, solution);
        dfxCheck(d_ctx, dfx_poisson_get_solution(d_fn, v.get()));
        DfxView<double> i(d_ctx, *d_init);
        dfxCheck(d_ctx, dfx_poisson_get_initial_guess(d_fn, i.get()));
      }
      const MultiVector<double, dfxMem> &
      getRhs() const override
      {
        return *d_rhs;
      }
      const MultiVector<double, dfxMem> &
      getInitialGuess() const override
      {
        return *d_init;
      }
      const utils::mpi::MPIComm &
      getMPIComm() const override
      {
        return d_comm;
      }

    private:
      void
      refresh()
      {
        DfxView<double> r(d_ctx, *d_rhs), i(d_ctx, *d_init);
        dfxCheck(d_ctx, dfx_poisson_get_rhs(d_fn, r.get()));
        dfxCheck(d_ctx, dfx_poisson_get_initial_guess(d_fn, i.get()));
      }

      dfx_ctx                                                  *d_ctx;
      dfx_poisson                                              *d_fn = nullptr;
      MultiVector<double, dfxMem>                              *d_rhs, *d_init;
      utils::mpi::MPIComm                                       d_comm;
      std::unique_ptr<DfxCellOperatorContext<double, double>>     d_ax;
      std::unique_ptr<DfxDiagonalOperatorContext<double, double>> d_pc;
    };

    // CGLinearSolver on libdfx (CGLinearSolver.h:90-94, .t.cpp:68-313):
    // multi-RHS Jacobi PCG with per-column convergence on the function's
    // libdfx Ax and PC contexts, x starting from the initial guess; the
    // converged x goes to setSolution
    class DfxCGLinearSolver : public LinearSolverImpl<double, double, dfxMem>
    {
    public:
      DfxCGLinearSolver(const size_type maxIter, const double absoluteTol,
                        const double relativeTol, const double divergenceTol,
                        LinearAlgebraProfiler profiler = LinearAlgebraProfiler(),
                        int device = 0)
        : d_ctx(dfxContext(device)), d_maxIter(maxIter), d_absTol(absoluteTol),
          d_relTol(relativeTol), d_divTol(divergenceTol)
      {
        (void)profiler;
      }

      LinearSolverError
      solve(LinearSolverFunction<double, double, dfxMem> &fn) override
      {
        const MultiVector<double, dfxMem> &b = fn.getRhs();
        MultiVector<double, dfxMem>        x(fn.getInitialGuess());
        DfxView<double>                    bv(d_ctx, b), xv(d_ctx, x);
        d_residualNorms.assign(b.getNumberComponents(), 0.0);
        uint32_t  iters = 0;
        const int rc    = dfx_cg_solve(d_ctx, dfxOperator(fn.getAxContext()).handle(),
                                    dfxOperator(fn.getPCContext()).handle(), bv.get(),
                                    xv.get(), d_absTol, d_relTol, d_divTol, d_maxIter, &iters,
                                    d_residualNorms.data());
        dfxCheck(d_ctx, rc);
        d_iterations = iters;
        fn.setSolution(x);
        return LinearSolverErrorMsg::isSuccessAndMsg(static_cast<LinearSolverErrorCode>(rc));
      }

      size_type iterations() const { return d_iterations; }
      const std::vector<double> &residualNorms() const { return d_residualNorms; }

    private:
      dfx_ctx            *d_ctx;
      size_type           d_maxIter;
      double              d_absTol, d_relTol, d_divTol;
      size_type           d_iterations = 0;
      std::vector<double> d_residualNorms;
    };
  } // namespace linearAlgebra
} // namespace dftefe
