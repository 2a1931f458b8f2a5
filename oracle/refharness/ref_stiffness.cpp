// Harness exposing the reference's stiffness operator (src/stiffness.hpp)
// through a C ABI.  Separate translation unit from the mass harness because
// src/linalg.hpp has no include guard and a non-inline string2order.
// Test infrastructure only.
#include <iostream>
#include "stiffness.hpp"

extern "C" {
// operators::stiffness_operator<double> (src/stiffness.hpp:245-248); the
// kernel reads [phi | dphi] (2 N^2 values, src/stiffness.hpp:44-45).
void ref_stiffness_operator_f64(double* phi2, double* U, double* W, double* G,
                                int num_cells, int degree, int order, int Mb,
                                int Nb)
{
  const std::size_t N = degree + 1, nd = N * N * N;
  operators::stiffness_operator<double>(
      std::span<double>(phi2, 2 * N * N),
      std::span<double>(U, (std::size_t)num_cells * nd),
      std::span<double>(W, (std::size_t)num_cells * nd),
      std::span<double>(G, (std::size_t)num_cells * nd * 6), num_cells,
      degree, (linalg::Order)order, Mb, Nb);
}
}
