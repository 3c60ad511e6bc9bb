// Stand-in of linearAlgebra::HermitianIterativeEigenSolver
// (src/linearAlgebra/HermitianIterativeEigenSolver.h:64-88).
#pragma once
#include <linearAlgebra/IdentityOperatorContext.h>
#include <linearAlgebra/LinearAlgebraTypes.h>
#include <linearAlgebra/OperatorContext.h>
#include <vector>
namespace dftefe
{
  namespace linearAlgebra
  {
    template <typename ValueTypeOperator, typename ValueTypeOperand,
              utils::MemorySpace memorySpace>
    class HermitianIterativeEigenSolver
    {
    public:
      using OpContext = OperatorContext<ValueTypeOperator, ValueTypeOperand, memorySpace>;
      using ValueType = blasLapack::scalar_type<ValueTypeOperator, ValueTypeOperand>;
      using RealType  = blasLapack::real_type<ValueType>;
      virtual ~HermitianIterativeEigenSolver() = default;
      virtual EigenSolverError
      solve(const OpContext &A, std::vector<RealType> &eigenValues,
            MultiVector<ValueType, memorySpace> &eigenVectors,
            bool computeEigenVectors = false,
            const OpContext &B = IdentityOperatorContext<ValueTypeOperator, ValueTypeOperand,
                                                         memorySpace>(),
            const OpContext &BInv = IdentityOperatorContext<ValueTypeOperator,
                                                            ValueTypeOperand, memorySpace>()) = 0;
    };
  } // namespace linearAlgebra
} // namespace dftefe
