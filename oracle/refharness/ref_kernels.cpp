// Harness exposing the reference's gemm_micro / gemm_blocked
// (src/linalg.hpp) and apply_contractions (src/tmm.hpp) for a fixed set of
// compile-time shapes.  Test infrastructure only.
#include <array> // src/tmm.hpp:49 uses std::array without including it
#include "linalg.hpp"
#include "tmm.hpp"

using linalg::Order;

namespace
{
template <int P, Order O, int MB, int NB>
void gb(const double* a, const double* b, double* c)
{
  constexpr int N = P + 1;
  linalg::gemm_blocked<double, N, N, N * N, O, MB, NB>(a, b, c);
}

template <int P, Order O>
int gb_blocks(int MB, int NB, const double* a, const double* b, double* c)
{
#define FEMTC_GB(mb, nb)                                                       \
  if (MB == mb && NB == nb)                                                    \
  {                                                                            \
    gb<P, O, mb, nb>(a, b, c);                                                 \
    return 0;                                                                  \
  }
  FEMTC_GB(0, 0)
  FEMTC_GB(4, 4)
  FEMTC_GB(8, 16)
  FEMTC_GB(4, 16)
  FEMTC_GB(16, 8)
  FEMTC_GB(16, 16)
  FEMTC_GB(8, 4)
#undef FEMTC_GB
  return -1;
}

template <int P>
int gb_order(int order, int MB, int NB, const double* a, const double* b,
             double* c)
{
  switch (order)
  {
  case 0: return gb_blocks<P, Order::ijk>(MB, NB, a, b, c);
  case 1: return gb_blocks<P, Order::ikj>(MB, NB, a, b, c);
  case 2: return gb_blocks<P, Order::jik>(MB, NB, a, b, c);
  case 3: return gb_blocks<P, Order::jki>(MB, NB, a, b, c);
  case 4: return gb_blocks<P, Order::kij>(MB, NB, a, b, c);
  case 5: return gb_blocks<P, Order::kji>(MB, NB, a, b, c);
  }
  return -1;
}

template <int Na, int Nb, bool Tr>
void tmm(const double* I0, const double* I1, const double* I2, double* B,
         double* A)
{
  Buffer<double, Na, Nb> buf;
  apply_contractions<double, Na, Nb, Tr>(I0, I1, I2, B, A, buf);
}
} // namespace

extern "C" {
// gemm_blocked<double, k=P+1, m=P+1, n=(P+1)^2, order, MB, NB>
int ref_gemm_blocked(int P, int order, int MB, int NB, const double* a,
                     const double* b, double* c)
{
  switch (P)
  {
  case 1: return gb_order<1>(order, MB, NB, a, b, c);
  case 2: return gb_order<2>(order, MB, NB, a, b, c);
  case 3: return gb_order<3>(order, MB, NB, a, b, c);
  case 4: return gb_order<4>(order, MB, NB, a, b, c);
  case 6: return gb_order<6>(order, MB, NB, a, b, c);
  case 8: return gb_order<8>(order, MB, NB, a, b, c);
  case 12: return gb_order<12>(order, MB, NB, a, b, c);
  case 25: return gb_order<25>(order, MB, NB, a, b, c);
  }
  return -1;
}

// gemm_micro<double, 26, 4, 16, 4, 16, 16, order> (tests/test_micro.cpp shape)
int ref_gemm_micro_26_4_16(int order, const double* a, const double* b,
                           double* c)
{
  switch (order)
  {
  case 0: linalg::gemm_micro<double, 26, 4, 16, 4, 16, 16, Order::ijk>(a, b, c); return 0;
  case 1: linalg::gemm_micro<double, 26, 4, 16, 4, 16, 16, Order::ikj>(a, b, c); return 0;
  case 2: linalg::gemm_micro<double, 26, 4, 16, 4, 16, 16, Order::jik>(a, b, c); return 0;
  case 3: linalg::gemm_micro<double, 26, 4, 16, 4, 16, 16, Order::jki>(a, b, c); return 0;
  case 4: linalg::gemm_micro<double, 26, 4, 16, 4, 16, 16, Order::kij>(a, b, c); return 0;
  case 5: linalg::gemm_micro<double, 26, 4, 16, 4, 16, 16, Order::kji>(a, b, c); return 0;
  }
  return -1;
}

// apply_contractions<double, Na, Nb, Tr> for the shapes below.
int ref_tmm_apply(int Na, int Nb, int Tr, const double* I0, const double* I1,
                  const double* I2, double* B, double* A)
{
#define FEMTC_TMM(na, nb)                                                      \
  if (Na == na && Nb == nb)                                                    \
  {                                                                            \
    if (Tr)                                                                    \
      tmm<na, nb, true>(I0, I1, I2, B, A);                                     \
    else                                                                       \
      tmm<na, nb, false>(I0, I1, I2, B, A);                                    \
    return 0;                                                                  \
  }
  FEMTC_TMM(2, 2)
  FEMTC_TMM(3, 2)
  FEMTC_TMM(2, 3)
  FEMTC_TMM(4, 4)
  FEMTC_TMM(5, 4)
  FEMTC_TMM(4, 5)
  FEMTC_TMM(6, 5)
  FEMTC_TMM(5, 6)
  FEMTC_TMM(7, 7)
#undef FEMTC_TMM
  return -1;
}
}
