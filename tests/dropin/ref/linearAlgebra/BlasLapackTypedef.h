// Stand-in of blasLapack::scalar_type / real_type
// (src/linearAlgebra/BlasLapackTypedef.h), real types only.
#pragma once
#include <type_traits>
namespace dftefe
{
  namespace linearAlgebra
  {
    namespace blasLapack
    {
      template <typename... T>
      using scalar_type = typename std::common_type<T...>::type;
      template <typename T>
      using real_type = T;
    } // namespace blasLapack
  }   // namespace linearAlgebra
} // namespace dftefe
