// Stand-in of linearAlgebra::IdentityOperatorContext (Y = X).
#pragma once
#include <linearAlgebra/OperatorContext.h>
namespace dftefe
{
  namespace linearAlgebra
  {
    template <typename ValueTypeOperator, typename ValueTypeOperand,
              utils::MemorySpace memorySpace>
    class IdentityOperatorContext
      : public OperatorContext<ValueTypeOperator, ValueTypeOperand, memorySpace>
    {
    public:
      using ValueTypeUnion = blasLapack::scalar_type<ValueTypeOperator, ValueTypeOperand>;
      void apply(MultiVector<ValueTypeOperand, memorySpace> &X,
                 MultiVector<ValueTypeUnion, memorySpace> &Y, bool = false,
                 bool = false) const override
      {
        Y = MultiVector<ValueTypeUnion, memorySpace>(X);
      }
    };
  } // namespace linearAlgebra
} // namespace dftefe
