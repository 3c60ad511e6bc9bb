// Stand-in of the non-MPI branch of src/utils/MPITypes.h:84.
#pragma once
namespace dftefe
{
  namespace utils
  {
    namespace mpi
    {
      using MPIComm = int;
    }
  } // namespace utils
} // namespace dftefe
