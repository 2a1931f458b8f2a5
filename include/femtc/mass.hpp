// Drop-in for the reference's src/mass.hpp entry point
// operators::mass_operator<T> (src/mass.hpp:196-199): same signature, same
// cell-local layouts (U, W, detJ: [num_cells][(P+1)^3]; phi: [(P+1)^2]),
// Q = P+1.  The spans are host memory; the call copies them to the current
// HIP device, applies the operator there and copies W back.
#pragma once

#include <cstddef>
#include <span>
#include <type_traits>

#include "femtc/linalg.hpp"
#include "femtc/shim_report.hpp"

namespace operators
{
template <typename T>
void mass_operator(std::span<T> phi, std::span<T> U, std::span<T> W,
                   std::span<T> detJ, std::size_t num_cells, int degree,
                   linalg::Order order, int Mb, int Nb)
{
  const int st = femtc_mass_operator(
      femtc_shim::dtype_of<T>(), phi.data(), phi.size(), U.data(), W.data(),
      detJ.data(), num_cells, degree, static_cast<int>(order), Mb, Nb,
      FEMTC_SHIM_MODE);
  femtc_shim::report(st, "mass_operator");
}
} // namespace operators
