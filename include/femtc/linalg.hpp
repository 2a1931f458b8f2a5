// Drop-in replacement for the reference's linalg vocabulary
// (src/linalg.hpp:11-37) on top of the femtc C ABI.  The loop order is a CPU
// blocking knob of the reference; the GPU kernels accept and validate it.
#pragma once

#include <stdexcept>
#include <string>
#include <type_traits>

#include "femtc.h"

namespace linalg
{
// Same enumerators in the same order as src/linalg.hpp:11-19.
enum class Order
{
  ijk,
  ikj,
  jik,
  jki,
  kij,
  kji
};

// src/linalg.hpp:21-37: throws std::runtime_error("Invalid loop order").
inline Order string2order(std::string order_str)
{
  int o = 0;
  if (femtc_string2order(order_str.c_str(), &o) != FEMTC_OK)
    throw std::runtime_error("Invalid loop order");
  return static_cast<Order>(o);
}
} // namespace linalg

namespace femtc_shim
{
// Mode used by the operators:: shims.  Default: reproduce the reference's
// own arithmetic (FEMTC_MODE_REFERENCE); define FEMTC_SHIM_MODE to
// FEMTC_MODE_CORRECT for the mathematically intended operator.
#ifndef FEMTC_SHIM_MODE
#define FEMTC_SHIM_MODE FEMTC_MODE_REFERENCE
#endif

template <typename T>
constexpr int dtype_of()
{
  static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>,
                "femtc operators support float and double");
  return std::is_same_v<T, double> ? FEMTC_F64 : FEMTC_F32;
}
} // namespace femtc_shim
