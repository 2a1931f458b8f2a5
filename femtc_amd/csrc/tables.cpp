// 1-D tables on [0,1]: Gauss-Legendre rule, Gauss-Lobatto-Legendre nodes and
// the Lagrange basis on the GLL nodes evaluated at the Gauss points.
// The reference takes its table as input (bench_mass.cpp:74-76 fills it with
// rand(); main.cpp:43 notes it should come from a basis library); these are
// the tables a real Q_P discretisation uses.
#include <cmath>
#include <vector>

#include "even_odd.hpp"
#include "femtc.h"

namespace
{
// Legendre P_n(x) and P_{n-1}(x) by the three-term recurrence.
void legendre(int n, double x, double& pn, double& pn1)
{
  double p0 = 1.0, p1 = x;
  if (n == 0)
  {
    pn = 1.0;
    pn1 = 0.0;
    return;
  }
  for (int k = 2; k <= n; ++k)
  {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  pn = p1;
  pn1 = p0;
}

// Gauss-Legendre on [-1,1]: Newton on P_n from the asymptotic guess.
void gauss_m11(int n, std::vector<double>& x, std::vector<double>& w)
{
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < (n + 1) / 2; ++i)
  {
    double z = std::cos(pi * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it)
    {
      double pn, pn1;
      legendre(n, z, pn, pn1);
      dp = n * (z * pn - pn1) / (z * z - 1.0);
      const double dz = pn / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16)
        break;
    }
    double pn, pn1;
    legendre(n, z, pn, pn1);
    dp = n * (z * pn - pn1) / (z * z - 1.0);
    const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    // z decreases with i: store symmetric pairs in increasing order
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  if (n % 2 == 1)
    x[n / 2] = 0.0;
}

// GLL nodes on [-1,1]: the roots of f(x) = x P_P(x) - P_{P-1}(x), which
// vanishes at +-1 and at the roots of P'_P; f'(x) = (P+1) P_P(x), so the
// update below is Newton's method, started from the Chebyshev-Lobatto nodes.
void gll_m11(int P, std::vector<double>& x)
{
  const int n = P + 1;
  x.assign(n, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < n; ++i)
  {
    double z = -std::cos(pi * i / P);
    if (i == 0 || i == P)
    {
      x[i] = z;
      continue;
    }
    for (int it = 0; it < 100; ++it)
    {
      double pn, pn1;
      legendre(P, z, pn, pn1);
      const double dz = (z * pn - pn1) / ((P + 1.0) * pn);
      z -= dz;
      if (std::fabs(dz) < 1e-16)
        break;
    }
    x[i] = z;
  }
  // exact symmetry about 0
  for (int i = 0; i < n / 2; ++i)
  {
    const double s = 0.5 * (x[n - 1 - i] - x[i]);
    x[i] = -s;
    x[n - 1 - i] = s;
  }
  if (n % 2 == 1)
    x[n / 2] = 0.0;
  x[0] = -1.0;
  x[n - 1] = 1.0;
}
} // namespace

extern "C" {

int femtc_gauss_rule(int nq, double* x, double* w)
{
  if (nq < 1 || nq > 64 || !x || !w)
    return FEMTC_ERR_INVALID_ARGUMENT;
  std::vector<double> xx, ww;
  gauss_m11(nq, xx, ww);
  for (int i = 0; i < nq; ++i)
  {
    x[i] = 0.5 * (xx[i] + 1.0);
    w[i] = 0.5 * ww[i];
  }
  return FEMTC_OK;
}

int femtc_gll_points(int degree, double* x)
{
  if (degree < 1 || degree > 64 || !x)
    return FEMTC_ERR_INVALID_ARGUMENT;
  std::vector<double> xx;
  gll_m11(degree, xx);
  for (int i = 0; i <= degree; ++i)
    x[i] = 0.5 * (xx[i] + 1.0);
  x[0] = 0.0;
  x[degree] = 1.0;
  return FEMTC_OK;
}

int femtc_basis_tables(int degree, int nq, double* tables)
{
  if (degree < 1 || degree > 64 || nq < 1 || nq > 64 || !tables)
    return FEMTC_ERR_INVALID_ARGUMENT;
  const int n = degree + 1;
  std::vector<double> xn(n), xq(nq), wq(nq);
  femtc_gll_points(degree, xn.data());
  femtc_gauss_rule(nq, xq.data(), wq.data());
  double* phi = tables;
  double* dphi = tables + n * nq;
  for (int d = 0; d < n; ++d)
  {
    double den = 1.0;
    for (int m = 0; m < n; ++m)
      if (m != d)
        den *= xn[d] - xn[m];
    for (int q = 0; q < nq; ++q)
    {
      const double t = xq[q];
      double v = 1.0;
      for (int m = 0; m < n; ++m)
        if (m != d)
          v *= t - xn[m];
      // derivative of prod_{m != d} (t - x_m): sum over the skipped factor
      double dv = 0.0;
      for (int k = 0; k < n; ++k)
      {
        if (k == d)
          continue;
        double term = 1.0;
        for (int m = 0; m < n; ++m)
          if (m != d && m != k)
            term *= t - xn[m];
        dv += term;
      }
      phi[d * nq + q] = v / den;
      dphi[d * nq + q] = dv / den;
    }
  }
  return FEMTC_OK;
}

int femtc_even_odd_tables(int degree, int nq, const double* tables,
                          const double* weights, double* out, int64_t* size,
                          int* symmetric)
{
  if (degree < 1 || degree > FEMTC_MAX_FUSED_DEGREE || nq < 1 || nq > 64 ||
      !tables || !size || !symmetric)
    return FEMTC_ERR_INVALID_ARGUMENT;
  const int n = degree + 1;
  *size = femtc::eo_table_size(n, nq);
  *symmetric = femtc::eo_symmetric(tables, n, nq) ? 1 : 0;
  if (out)
    femtc::eo_build(tables, weights, n, nq, out);
  return FEMTC_OK;
}

} // extern "C"
