// Stand-in of the error types of src/linearAlgebra/LinearAlgebraTypes.h:83-165
// and the isSuccessAndMsg helpers (:199-217).
#pragma once
#include <string>
namespace dftefe
{
  namespace linearAlgebra
  {
    enum class LinearSolverErrorCode
    {
      SUCCESS,
      FAILED_TO_CONVERGE,
      RESIDUAL_DIVERGENCE,
      DIVISON_BY_ZERO,
      OTHER_ERROR
    };
    enum class EigenSolverErrorCode
    {
      SUCCESS,
      LAPACK_ERROR,
      LANCZOS_BETA_ZERO,
      LANCZOS_SUBSPACE_INSUFFICIENT,
      CHFSI_ORTHONORMALIZATION_ERROR,
      CHFSI_RAYLEIGH_RITZ_ERROR,
      KS_MAX_PASS_ERROR,
      KS_CHFSI_ERROR,
      KS_LANCZOS_ERROR,
      KS_NEWTON_RAPHSON_ERROR,
      ELPASCALAPACK_ERROR,
      OTHER_ERROR
    };
    struct LinearSolverError
    {
      bool                  isSuccess;
      LinearSolverErrorCode err;
      std::string           msg;
    };
    struct EigenSolverError
    {
      bool                 isSuccess;
      EigenSolverErrorCode err;
      std::string          msg;
    };
    class LinearSolverErrorMsg
    {
    public:
      static LinearSolverError isSuccessAndMsg(const LinearSolverErrorCode &e)
      {
        return {e == LinearSolverErrorCode::SUCCESS, e, ""};
      }
    };
    class EigenSolverErrorMsg
    {
    public:
      static EigenSolverError isSuccessAndMsg(const EigenSolverErrorCode &e)
      {
        return {e == EigenSolverErrorCode::SUCCESS, e, ""};
      }
    };
  } // namespace linearAlgebra
} // namespace dftefe
