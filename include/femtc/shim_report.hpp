// Out-of-line helper of the operators:: shims (header-only, inline).
#pragma once

#include <iostream>
#include <stdexcept>
#include <string>

#include "femtc/linalg.hpp"

namespace femtc_shim
{
// Reference behaviour for unsupported ladder arguments: print the message
// and return with W untouched (src/mass.hpp:134-136, 165-167, 190-192).
// Any other failure throws, since the reference has no such path.
inline void report(int status, const char* where)
{
  switch (status)
  {
  case FEMTC_OK:
    return;
  case FEMTC_ERR_DEGREE_NOT_SUPPORTED:
  case FEMTC_ERR_ORDER_NOT_SUPPORTED:
  case FEMTC_ERR_BLOCK_NOT_SUPPORTED:
    std::cout << femtc_status_string(status) << std::endl;
    return;
  default:
    throw std::runtime_error(std::string(where) + ": " +
                             femtc_status_string(status) + " (" +
                             femtc_last_error() + ")");
  }
}
} // namespace femtc_shim
