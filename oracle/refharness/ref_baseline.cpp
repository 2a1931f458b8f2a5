// CPU baseline harness: the reference's per-cell operator templates
// (src/mass.hpp:10-49, src/stiffness.hpp:34-99) instantiated for the
// degrees the benchmarks use, every loop order, and the block sizes
// {0,4,8,16} (the sweep of run.sh:13-23), built with -Ofast and an explicit
// -march.  Called on disjoint cell ranges from several host threads, which is
// the reference's one-rank-per-core model without MPI.
// Test / benchmark infrastructure only.
#include <iostream>
#include "stiffness.hpp"

using linalg::Order;

namespace
{
template <int P, Order O, int MB, int NB>
void st(double* phi2, double* U, double* W, double* G, long n)
{
  constexpr std::size_t N = P + 1, nd = N * N * N;
  operators::stiffness_operator<double, P, O, MB, NB>(
      std::span<const double>(phi2, 2 * N * N), std::span<double>(U, n * nd),
      std::span<double>(W, n * nd), std::span<const double>(G, n * nd * 6),
      n);
}

template <int P, Order O>
int st_blocks(int MB, int NB, double* phi2, double* U, double* W, double* G,
              long n)
{
#define FEMTC_ST(mb, nb)                                                       \
  if (MB == mb && NB == nb)                                                    \
  {                                                                            \
    st<P, O, mb, nb>(phi2, U, W, G, n);                                        \
    return 0;                                                                  \
  }
  FEMTC_ST(0, 0)
  FEMTC_ST(0, 4)
  FEMTC_ST(0, 8)
  FEMTC_ST(0, 16)
  FEMTC_ST(4, 0)
  FEMTC_ST(4, 4)
  FEMTC_ST(4, 8)
  FEMTC_ST(4, 16)
  FEMTC_ST(8, 0)
  FEMTC_ST(8, 8)
  FEMTC_ST(8, 16)
  FEMTC_ST(16, 16)
#undef FEMTC_ST
  return -1;
}

template <int P>
int st_order(int order, int MB, int NB, double* phi2, double* U, double* W,
             double* G, long n)
{
  switch (order)
  {
  case 0: return st_blocks<P, Order::ijk>(MB, NB, phi2, U, W, G, n);
  case 1: return st_blocks<P, Order::ikj>(MB, NB, phi2, U, W, G, n);
  case 2: return st_blocks<P, Order::jik>(MB, NB, phi2, U, W, G, n);
  case 3: return st_blocks<P, Order::jki>(MB, NB, phi2, U, W, G, n);
  case 4: return st_blocks<P, Order::kij>(MB, NB, phi2, U, W, G, n);
  case 5: return st_blocks<P, Order::kji>(MB, NB, phi2, U, W, G, n);
  }
  return -1;
}
} // namespace

extern "C" {
// Stiffness per-cell operator for P in {4, 6} (config degrees).
int ref_baseline_stiffness(int P, int order, int MB, int NB, double* phi2,
                           double* U, double* W, double* G, long num_cells)
{
  switch (P)
  {
  case 4: return st_order<4>(order, MB, NB, phi2, U, W, G, num_cells);
  case 6: return st_order<6>(order, MB, NB, phi2, U, W, G, num_cells);
  }
  return -1;
}
}
