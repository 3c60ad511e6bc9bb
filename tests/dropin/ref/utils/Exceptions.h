// Stand-in of utils::throwException (src/utils/Exceptions.h:118-132).
#pragma once
#include <stdexcept>
#include <string>
namespace dftefe
{
  namespace utils
  {
    inline void
    throwException(bool condition, std::string msg = "")
    {
      if (!condition)
        throw std::logic_error(msg);
    }
  } // namespace utils
} // namespace dftefe
