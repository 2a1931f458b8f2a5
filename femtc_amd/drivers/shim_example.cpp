// Source-level drop-in check: code written against the reference's operator
// headers (operators::mass_operator<T>, operators::stiffness_operator<T>,
// linalg::string2order) compiles and runs unchanged against the femtc shims.
// Prints the relative difference between two calls and exits non-zero on
// failure.  Usage: shim_example [degree]
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "femtc/mass.hpp"
#include "femtc/stiffness.hpp"

int main(int argc, char** argv)
{
  const int degree = argc > 1 ? std::atoi(argv[1]) : 3;
  const int N = degree + 1, nd = N * N * N;
  const std::size_t num_cells = 100;
  std::vector<double> phi(2 * N * N), U(num_cells * nd), W(num_cells * nd),
      detJ(num_cells * nd), G(num_cells * nd * 6);
  for (auto* v : {&phi, &U, &detJ, &G})
    for (auto& x : *v)
      x = std::rand() / (double)RAND_MAX;
  linalg::Order order = linalg::string2order("kij");
  std::span<double> phi_m(phi.data(), N * N);
  operators::mass_operator<double>(phi_m, U, W, detJ, num_cells, degree, order,
                                   8, 16);
  double s = 0;
  for (double w : W)
    s += w;
  operators::stiffness_operator<double>(phi, U, W, G, (int)num_cells, degree,
                                        order, 0, 0);
  double t = 0;
  for (double w : W)
    t += w;
  // unsupported degree: prints "degree not supported", returns
  operators::mass_operator<double>(phi_m, U, W, detJ, num_cells, 30, order, 0,
                                   0);
  try
  {
    linalg::string2order("xyz");
    return 2;
  }
  catch (const std::runtime_error&)
  {
  }
  std::cout << "mass_sum " << s << " stiffness_sum " << t << std::endl;
  return (std::isfinite(s) && std::isfinite(t)) ? 0 : 1;
}
