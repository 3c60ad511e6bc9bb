// Stand-in of linearAlgebra::OrthogonalizationType
// (src/linearAlgebra/ChebyshevFilteredEigenSolver.h:45-50).
#pragma once
namespace dftefe
{
  namespace linearAlgebra
  {
    enum class OrthogonalizationType
    {
      CHOLESKY_GRAMSCHMIDT,
      MULTIPASS_LOWDIN,
      MULTIPASS_CGS,
    };
  } // namespace linearAlgebra
} // namespace dftefe
