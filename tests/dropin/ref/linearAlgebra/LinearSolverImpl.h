// Stand-in of linearAlgebra::LinearSolverImpl (src/linearAlgebra/LinearSolverImpl.h:59-83).
#pragma once
#include <linearAlgebra/LinearAlgebraTypes.h>
#include <linearAlgebra/LinearSolverFunction.h>
namespace dftefe
{
  namespace linearAlgebra
  {
    template <typename ValueTypeOperator, typename ValueTypeOperand,
              utils::MemorySpace memorySpace>
    class LinearSolverImpl
    {
    public:
      virtual ~LinearSolverImpl() = default;
      virtual LinearSolverError
      solve(LinearSolverFunction<ValueTypeOperator, ValueTypeOperand, memorySpace>
              &linearSolverFunction) = 0;
    };
  } // namespace linearAlgebra
} // namespace dftefe
