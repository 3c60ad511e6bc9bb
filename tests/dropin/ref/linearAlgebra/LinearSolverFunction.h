// Stand-in of linearAlgebra::LinearSolverFunction
// (src/linearAlgebra/LinearSolverFunction.h:59-105).
#pragma once
#include <linearAlgebra/OperatorContext.h>
#include <utils/MPITypes.h>
namespace dftefe
{
  namespace linearAlgebra
  {
    template <typename ValueTypeOperator, typename ValueTypeOperand,
              utils::MemorySpace memorySpace>
    class LinearSolverFunction
    {
    public:
      using ValueType = blasLapack::scalar_type<ValueTypeOperator, ValueTypeOperand>;
      virtual ~LinearSolverFunction() = default;
      virtual const OperatorContext<ValueTypeOperator, ValueTypeOperand, memorySpace> &
      getAxContext() const = 0;
      virtual const OperatorContext<ValueTypeOperator, ValueTypeOperand, memorySpace> &
      getPCContext() const = 0;
      virtual void setSolution(const MultiVector<ValueTypeOperand, memorySpace> &x) = 0;
      virtual void getSolution(MultiVector<ValueType, memorySpace> &solution) = 0;
      virtual const MultiVector<ValueTypeOperand, memorySpace> &getRhs() const = 0;
      virtual const MultiVector<ValueType, memorySpace> &getInitialGuess() const = 0;
      virtual const utils::mpi::MPIComm &getMPIComm() const = 0;
    };
  } // namespace linearAlgebra
} // namespace dftefe
