// Stand-in declaration of the reference's type aliases
// (src/utils/TypeConfig.h:8-9), for compiling include/dfx_dropin.hpp without
// the reference tree.
#pragma once
namespace dftefe
{
  using size_type        = unsigned int;
  using global_size_type = unsigned long int;
} // namespace dftefe
