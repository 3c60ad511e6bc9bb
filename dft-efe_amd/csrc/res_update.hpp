// One element of the residual Chebyshev degree step
// (src/linearAlgebra/ChebyshevFilter.t.cpp:360-397):
//   s1  = alpha1 s2 + (-c alpha1) rn
//   res = s1 + alpha2 res
//   res = res + alpha1 mu r0
// shared by the stand-alone step kernels and the cell kernel's fused
// epilogue, so that a value gets the same bits on every path.
#pragma once
#include <hip/hip_runtime.h>

namespace dfx
{
  namespace k
  {
    template <class T>
    __device__ inline T resUpdate(T s2, T rn, T res, T r0, double mu, T alpha1, T alpha2, T c)
    {
      const T s1 = alpha1 * s2 + (-c * alpha1) * rn;
      const T v  = (T)1 * s1 + alpha2 * res;
      return (T)1 * (T)1 * v + alpha1 * (T)mu * r0;
    }
  } // namespace k
} // namespace dfx
