// Harness exposing the reference's mass operator (src/mass.hpp) through a C
// ABI so the oracle tests can call the compiled reference.  The reference
// headers are included from where they lie (-I <reference>/src); nothing of
// them is copied here.  Test infrastructure only.
#include <iostream> // src/mass.hpp:135 uses std::cout without including it
#include "mass.hpp"

extern "C" {
// operators::mass_operator<double> (src/mass.hpp:196-199); Q = P+1.
void ref_mass_operator_f64(double* phi, double* U, double* W, double* detJ,
                           long num_cells, int degree, int order, int Mb,
                           int Nb)
{
  const std::size_t N = degree + 1, nd = N * N * N;
  operators::mass_operator<double>(
      std::span<double>(phi, N * N), std::span<double>(U, num_cells * nd),
      std::span<double>(W, num_cells * nd),
      std::span<double>(detJ, num_cells * nd), num_cells, degree,
      (linalg::Order)order, Mb, Nb);
}
// linalg::string2order (src/linalg.hpp:21-37); returns -1 on the throw.
int ref_string2order(const char* s)
{
  try
  {
    return (int)linalg::string2order(s);
  }
  catch (const std::runtime_error&)
  {
    return -1;
  }
}
}
