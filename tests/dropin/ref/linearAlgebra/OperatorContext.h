// Stand-in of linearAlgebra::OperatorContext (src/linearAlgebra/OperatorContext.h:47-109).
#pragma once
#include <linearAlgebra/BlasLapackTypedef.h>
#include <linearAlgebra/MultiVector.h>
namespace dftefe
{
  namespace linearAlgebra
  {
    template <typename ValueTypeOperator, typename ValueTypeOperand,
              utils::MemorySpace memorySpace>
    class OperatorContext
    {
    public:
      using ValueTypeUnion = blasLapack::scalar_type<ValueTypeOperator, ValueTypeOperand>;
      virtual ~OperatorContext() = default;
      virtual void apply(MultiVector<ValueTypeOperand, memorySpace> &X,
                         MultiVector<ValueTypeUnion, memorySpace> &  Y,
                         bool updateGhostX = false, bool updateGhostY = false) const = 0;
    };
  } // namespace linearAlgebra
} // namespace dftefe
