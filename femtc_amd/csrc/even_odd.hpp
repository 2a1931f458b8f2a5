// Even-odd form of the 1-D basis tables.
//
// The operators take their tables as input (src/mass.hpp:19-21,
// src/stiffness.hpp:45-49).  Tables of a nodal basis on nodes symmetric about
// 0, evaluated at symmetric quadrature points (femtc_basis_tables: the GLL
// Lagrange basis at Gauss points) satisfy
//   phi [N-1-b][Q-1-q] =  phi [b][q]   (S = +1)
//   phi'[N-1-b][Q-1-q] = -phi'[b][q]   (S = -1).
// A contraction out[m] = sum_k A(k, m) v[k] with such a table (A(k, m) =
// t[k][m] forward, t[m][k] backward) then splits into two half-size ones.
// Arbitrary tables fail eo_symmetric and keep the general form.
// With e[k] = v[k] + v[K-1-k], o[k] = v[k] - v[K-1-k] (k < K/2) and
// e[K/2] = v[K/2] for odd K:
//   Ce[m][k] = (A(k,m) + A(K-1-k,m)) / 2   (A(K/2, m) at the middle k)
//   Co[m][k] = (A(k,m) - A(K-1-k,m)) / 2
//   S = +1: EV = Ce e, OD = Co o;   S = -1: EV = Co o, OD = Ce e
//   out[m] = EV[m] + OD[m], out[M-1-m] = EV[m] - OD[m] (m < M/2),
//   out[M/2] = EV[M/2] for odd M.
// For K = M = 7 one contraction costs 25 multiply-adds and 12 additions
// instead of 49 multiply-adds, and its table rows hold 4 values instead of 7.
//
// Table set layout (elements of T): four terms, each a block A of rowsA rows
// of RA values (the columns of Ce or Co that feed EV) and a block B of rowsB
// rows of RB values (those that feed OD), rows padded to a multiple of 4;
//   term 0: phi  forward  (K = N, M = Q, S = +1)
//   term 1: phi' forward  (K = N, M = Q, S = -1)
//   term 2: phi  backward (K = Q, M = N, S = +1)
//   term 3: phi' backward (K = Q, M = N, S = -1)
// then the Q quadrature weights (the per-cell geometry operators read them).
#pragma once

#include <cfloat>
#include <cmath>
#include <cstring>

namespace femtc
{

struct EoDims
{
  int Kp, Hk, Mp, Hm; // ceil/floor halves of the input and output lengths
  int RA, RB;         // padded row lengths of blocks A and B
  int rowsA, rowsB;
  int size;           // elements of the term
};

constexpr EoDims eo_dims(int K, int M, int S)
{
  EoDims d{};
  d.Kp = (K + 1) / 2;
  d.Hk = K / 2;
  d.Mp = (M + 1) / 2;
  d.Hm = M / 2;
  d.RA = (d.Mp + 3) & ~3;
  d.RB = (d.Hm + 3) & ~3;
  d.rowsA = S > 0 ? d.Kp : d.Hk;
  d.rowsB = S > 0 ? d.Hk : d.Kp;
  d.size = d.rowsA * d.RA + d.rowsB * d.RB;
  return d;
}

constexpr int EO_PHI_F = 0, EO_DPHI_F = 1, EO_PHI_B = 2, EO_DPHI_B = 3,
              EO_WEIGHTS = 4;

constexpr EoDims eo_term_dims(int N, int Q, int t)
{
  return eo_dims(t < 2 ? N : Q, t < 2 ? Q : N, t % 2 ? -1 : 1);
}

// element offset of term t (EO_WEIGHTS: of the weights)
constexpr int eo_offset(int N, int Q, int t)
{
  int off = 0;
  for (int i = 0; i < t; ++i)
    off += eo_term_dims(N, Q, i).size;
  return off;
}

constexpr int eo_table_size(int N, int Q) { return eo_offset(N, Q, EO_WEIGHTS) + Q; }

// True when phi ([N][Q], tab) is symmetric and phi' ([N][Q], tab + N*Q)
// antisymmetric under (b, q) -> (N-1-b, Q-1-q) to 64 units of `eps` (the
// tables' own precision) of the largest entry: the even-odd form then
// equals the general one to rounding.
inline bool eo_symmetric(const double* tab, int N, int Q, double eps = DBL_EPSILON)
{
  const int NQ = N * Q;
  double mx = 0.0;
  for (int i = 0; i < 2 * NQ; ++i)
    mx = std::fmax(mx, std::fabs(tab[i]));
  const double tol = 64.0 * eps * mx;
  for (int t = 0; t < 2; ++t)
  {
    const double s = t ? -1.0 : 1.0;
    const double* T = tab + t * NQ;
    for (int b = 0; b < N; ++b)
      for (int q = 0; q < Q; ++q)
        if (!(std::fabs(T[(N - 1 - b) * Q + (Q - 1 - q)] - s * T[b * Q + q]) <=
              tol))
          return false;
  }
  return true;
}

// The table set of [phi | phi'] (tab, FP64) and the weights w (may be null:
// zeros) into out[eo_table_size(N, Q)].
inline void eo_build(const double* tab, const double* w, int N, int Q,
                     double* out)
{
  std::memset(out, 0, sizeof(double) * eo_table_size(N, Q));
  for (int t = 0; t < 4; ++t)
  {
    const bool fwd = t < 2;
    const int K = fwd ? N : Q, S = t % 2 ? -1 : 1;
    const double* T = tab + (t % 2) * N * Q;
    auto A = [&](int k, int m) { return fwd ? T[k * Q + m] : T[m * Q + k]; };
    auto Ce = [&](int m, int k) {
      return 2 * k + 1 == K ? A(k, m) : 0.5 * (A(k, m) + A(K - 1 - k, m));
    };
    auto Co = [&](int m, int k) { return 0.5 * (A(k, m) - A(K - 1 - k, m)); };
    const EoDims d = eo_term_dims(N, Q, t);
    double* a = out + eo_offset(N, Q, t);
    double* b = a + d.rowsA * d.RA;
    for (int k = 0; k < d.rowsA; ++k)
      for (int m = 0; m < d.Mp; ++m)
        a[k * d.RA + m] = S > 0 ? Ce(m, k) : Co(m, k);
    for (int k = 0; k < d.rowsB; ++k)
      for (int m = 0; m < d.Hm; ++m)
        b[k * d.RB + m] = S > 0 ? Co(m, k) : Ce(m, k);
  }
  double* wo = out + eo_offset(N, Q, EO_WEIGHTS);
  for (int q = 0; q < Q; ++q)
    wo[q] = w ? w[q] : 0.0;
}

} // namespace femtc
