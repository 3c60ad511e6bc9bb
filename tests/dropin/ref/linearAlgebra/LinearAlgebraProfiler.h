// Stand-in of linearAlgebra::LinearAlgebraProfiler (src/linearAlgebra/
// LinearAlgebraProfiler.h:41-103), default-constructible.
#pragma once
namespace dftefe
{
  namespace linearAlgebra
  {
    class LinearAlgebraProfiler
    {};
  } // namespace linearAlgebra
} // namespace dftefe
