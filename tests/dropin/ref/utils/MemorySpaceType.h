// Stand-in of src/utils/MemorySpaceType.h:36-41.
#pragma once
namespace dftefe
{
  namespace utils
  {
    enum class MemorySpace
    {
      HOST,
      HOST_PINNED,
      DEVICE
    };
  } // namespace utils
} // namespace dftefe
