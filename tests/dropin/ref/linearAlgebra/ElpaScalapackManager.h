// Stand-in of linearAlgebra::ElpaScalapackManager (an opaque ELPA / ScaLAPACK
// grid in the reference; src/linearAlgebra/ElpaScalapackManager.h).
#pragma once
namespace dftefe
{
  namespace linearAlgebra
  {
    class ElpaScalapackManager
    {};
  } // namespace linearAlgebra
} // namespace dftefe
