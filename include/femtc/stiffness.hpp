// Drop-in for the reference's src/stiffness.hpp entry point
// operators::stiffness_operator<T> (src/stiffness.hpp:245-248): same
// signature (note `int num_cells`), same layouts (U, W: [num_cells][(P+1)^3];
// G: [num_cells][(P+1)^3][6]), Q = P+1.  phi must hold [phi | dphi], i.e.
// 2 (P+1)^2 values: the reference kernel reads the derivative table at offset
// (P+1)^2 (src/stiffness.hpp:45), which bench_stiffness.cpp:74 does not
// allocate; here a short table is rejected instead of read out of bounds.
#pragma once

#include <cstddef>
#include <span>
#include <type_traits>

#include "femtc/linalg.hpp"
#include "femtc/shim_report.hpp"

namespace operators
{
template <typename T>
void stiffness_operator(std::span<T> phi, std::span<T> U, std::span<T> W,
                        std::span<T> G, int num_cells, int degree,
                        linalg::Order order, int Mb, int Nb)
{
  const int st = femtc_stiffness_operator(
      femtc_shim::dtype_of<T>(), phi.data(), phi.size(), U.data(), W.data(),
      G.data(), num_cells, degree, static_cast<int>(order), Mb, Nb,
      FEMTC_SHIM_MODE);
  femtc_shim::report(st, "stiffness_operator");
}
} // namespace operators
