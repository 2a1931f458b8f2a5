/*
 * femtc CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the arithmetic of the reference
 * (IgorBaratta/femtc) for the sum-factorized mass / stiffness operator path,
 * plus the pieces the reference leaves out (dofmap gather/scatter, affine
 * geometry factors, Jacobi-PCG).  It is the checker for the HIP library
 * (femtc_amd/libfemtc.so): only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Conventions (shared with the HIP kernels):
 *   N = P+1 dofs per direction, Q quadrature points per direction.
 *   Cell tensors are row-major [a][b][c] with a the slowest index.
 *   A 1-D table is stored [dof][q] (row-major N x Q); the stiffness tables
 *   are [phi | dphi], dphi at offset N*Q (src/stiffness.hpp:44-45).
 *   Geometry G is [cell][q][6] packed as (S22,S12,S02,S11,S01,S00) where S is
 *   the symmetric 3x3 tensor acting on (d/da, d/db, d/dc)
 *   (src/stiffness.hpp:25-27).
 *
 * Two modes:
 *   literal  -- what the reference computes, defects included
 *               (phi_t aliases phi: src/mass.hpp:20-21; the stiffness
 *               backward chains overwrite W: src/stiffness.hpp:85,91,97;
 *               "Wz" is not differentiated: src/stiffness.hpp:76-77).
 *   correct  -- the operator the reference intends: true transposes, summed
 *               chains, a gradient in every direction, Q independent of N.
 *
 * Parity pins (what checks each function) are listed in the test files
 * tests/test_oracle_*.py; every function below has at least one.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Counter-based random numbers (SplitMix64 finaliser).                */
/* value(seed, i) depends only on (seed, i), so CPU and GPU harnesses  */
/* produce identical bits.                                             */
/* ------------------------------------------------------------------ */
static uint64_t orc_mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_fill_uniform(uint64_t seed, int64_t n, double lo, double hi,
                      double* out)
{
  const double span = hi - lo;
  for (int64_t i = 0; i < n; ++i)
  {
    uint64_t z = orc_mix64(seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL);
    double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    double t = span * u;
    out[i] = lo + t;
  }
}

/* ------------------------------------------------------------------ */
/* a1: gemm_micro  (src/linalg.hpp:45-94)                              */
/* C(m x nc, row-major, ldc) += A(m x k, col-major, lda) B(k x nc,     */
/* row-major, ldb) in one of the six loop orders ijk,ikj,jik,jki,kij,  */
/* kji (src/linalg.hpp:11-19).  Every order accumulates p = 0..k-1 in  */
/* increasing order into each C(i,j); only the traversal differs.      */
/* ------------------------------------------------------------------ */
void orc_gemm_micro(int k, int m, int nc, int lda, int ldb, int ldc,
                    int order, const double* a, const double* b, double* c)
{
#define A_(i, p) a[(i) + (p) * lda]
#define B_(p, j) b[(p) * ldb + (j)]
#define C_(i, j) c[(i) * ldc + (j)]
  switch (order)
  {
  case 0: /* ijk */
    for (int i = 0; i < m; i++)
      for (int j = 0; j < nc; j++)
        for (int p = 0; p < k; p++)
          C_(i, j) += A_(i, p) * B_(p, j);
    break;
  case 1: /* ikj */
    for (int i = 0; i < m; i++)
      for (int p = 0; p < k; p++)
        for (int j = 0; j < nc; j++)
          C_(i, j) += A_(i, p) * B_(p, j);
    break;
  case 2: /* jik */
    for (int j = 0; j < nc; j++)
      for (int i = 0; i < m; i++)
        for (int p = 0; p < k; p++)
          C_(i, j) += A_(i, p) * B_(p, j);
    break;
  case 3: /* jki */
    for (int j = 0; j < nc; j++)
      for (int p = 0; p < k; p++)
        for (int i = 0; i < m; i++)
          C_(i, j) += A_(i, p) * B_(p, j);
    break;
  case 4: /* kij */
    for (int p = 0; p < k; p++)
      for (int i = 0; i < m; i++)
        for (int j = 0; j < nc; j++)
          C_(i, j) += A_(i, p) * B_(p, j);
    break;
  default: /* kji */
    for (int p = 0; p < k; p++)
      for (int j = 0; j < nc; j++)
        for (int i = 0; i < m; i++)
          C_(i, j) += A_(i, p) * B_(p, j);
    break;
  }
#undef A_
#undef B_
#undef C_
}

/* ------------------------------------------------------------------ */
/* a2: gemm_blocked  (src/linalg.hpp:101-182)                          */
/* C(m x n, col-major) = A(m x k, col-major) B(k x n, row-major).      */
/* C is ASSIGNED block by block (src/linalg.hpp:136), remainder blocks */
/* handled as in src/linalg.hpp:140-181.  MB/NB = 0 means "whole".     */
/* ------------------------------------------------------------------ */
static void orc_block(int k, int m, int n, int order, int bx, int by,
                      const double* Aik, const double* Bpj, double* Cij,
                      double* tmp)
{
  memset(tmp, 0, sizeof(double) * (size_t)bx * (size_t)by);
  orc_gemm_micro(k, bx, by, m, n, by, order, Aik, Bpj, tmp);
  for (int i = 0; i < bx; i++)
    for (int j = 0; j < by; j++)
      Cij[j * m + i] = tmp[i * by + j];
}

void orc_gemm_blocked(int k, int m, int n, int order, int MB, int NB,
                      const double* a, const double* b, double* c)
{
  const int bx = (MB == 0) ? m : MB;
  const int by = (NB == 0) ? n : NB;
  const int Nm = m / bx, Nn = n / by;
  const int mrem = m % bx, nrem = n % by;
  double* tmp = (double*)malloc(sizeof(double) * (size_t)bx * (size_t)by);
  for (int jb = 0; jb < Nn; jb++)
    for (int ib = 0; ib < Nm; ib++)
      orc_block(k, m, n, order, bx, by, a + ib * bx, b + jb * by,
                c + jb * (m * by) + ib * bx, tmp);
  if (mrem > 0)
    for (int jb = 0; jb < Nn; jb++)
      orc_block(k, m, n, order, mrem, by, a + Nm * bx, b + jb * by,
                c + jb * (m * by) + Nm * bx, tmp);
  if (nrem > 0)
    for (int ib = 0; ib < Nm; ib++)
      orc_block(k, m, n, order, bx, nrem, a + ib * bx, b + Nn * by,
                c + Nn * (m * by) + ib * bx, tmp);
  if (mrem > 0 && nrem > 0)
    orc_block(k, m, n, order, mrem, nrem, a + Nm * bx, b + Nn * by,
              c + Nn * (m * by) + Nm * bx, tmp);
  free(tmp);
}

/* ------------------------------------------------------------------ */
/* One 1-D contraction step ("rotation"), the unit the operators are   */
/* built from (src/mass.hpp:36-47 call gemm_blocked with k=m=N, n=N^2).*/
/* X is [S][R] (contracted index slowest), T is [S][D] row-major, and  */
/* out[r][d] = sum_s T[s][d] X[s][r]: the contracted index disappears  */
/* and the new index becomes the fastest, so three steps restore the   */
/* axis order.  Accumulation order s = 0..S-1 as in gemm_micro.        */
/* ------------------------------------------------------------------ */
void orc_rot(int S, int D, int R, const double* T, const double* X,
             double* out)
{
  for (int r = 0; r < R; r++)
    for (int d = 0; d < D; d++)
    {
      double acc = 0.0;
      for (int s = 0; s < S; s++)
        acc += T[s * D + d] * X[s * R + r];
      out[r * D + d] = acc;
    }
}

/* Transpose of an S x D row-major table into D x S. */
static void orc_transpose(int S, int D, const double* T, double* Tt)
{
  for (int s = 0; s < S; s++)
    for (int d = 0; d < D; d++)
      Tt[d * S + s] = T[s * D + d];
}

/* chain(T0,T1,T2) X: three rotations with source extent S and target  */
/* extent D in every direction; T_j acts on axis j.                    */
static void orc_chain(int S, int D, const double* T0, const double* T1,
                      const double* T2, const double* X, double* out,
                      double* w0, double* w1)
{
  orc_rot(S, D, S * S, T0, X, w0);   /* [b][c][d0]   */
  orc_rot(S, D, S * D, T1, w0, w1);  /* [c][d0][d1]  */
  orc_rot(S, D, D * D, T2, w1, out); /* [d0][d1][d2] */
}

/* ------------------------------------------------------------------ */
/* a5: transform  (src/stiffness.hpp:14-29), in place per point.       */
/* ------------------------------------------------------------------ */
void orc_transform(int nq, const double* G, double* fw0, double* fw1,
                   double* fw2)
{
  for (int iq = 0; iq < nq; ++iq)
  {
    const double* g = G + 6 * iq;
    const double w0 = fw0[iq], w1 = fw1[iq], w2 = fw2[iq];
    fw0[iq] = g[5] * w0 + g[4] * w1 + g[2] * w2;
    fw1[iq] = g[4] * w0 + g[3] * w1 + g[1] * w2;
    fw2[iq] = g[2] * w0 + g[1] * w1 + g[0] * w2;
  }
}

/* ------------------------------------------------------------------ */
/* a4 literal: mass_operator  (src/mass.hpp:10-49).  Q = N = P+1, the  */
/* same table in both directions (phi_t aliases phi, :20-21), detJ has */
/* N^3 entries per cell (:27), W assigned (:47).                       */
/* ------------------------------------------------------------------ */
void orc_mass_literal(int P, const double* phi, const double* U, double* W,
                      const double* detJ, int64_t num_cells)
{
  const int N = P + 1, nd = N * N * N;
  double* t0 = (double*)malloc(sizeof(double) * nd);
  double* t1 = (double*)malloc(sizeof(double) * nd);
  double* w0 = (double*)malloc(sizeof(double) * nd);
  for (int64_t cell = 0; cell < num_cells; cell++)
  {
    const double* Uc = U + cell * nd;
    const double* Dc = detJ + cell * nd;
    double* Wc = W + cell * nd;
    orc_chain(N, N, phi, phi, phi, Uc, w0, t0, t1);
    for (int i = 0; i < nd; i++)
      w0[i] = w0[i] * Dc[i];
    orc_chain(N, N, phi, phi, phi, w0, Wc, t0, t1);
  }
  free(t0);
  free(t1);
  free(w0);
}

/* ------------------------------------------------------------------ */
/* a6 literal: stiffness_operator  (src/stiffness.hpp:34-99).          */
/* phi2 = [phi | dphi] with 2 N^2 entries (:44-45).  The three forward  */
/* chains are Wx = chain(g,phi,phi) U (:68-70), Wy = chain(phi,g,phi) U */
/* (:72-74) and Wz = chain(phi,phi,phi) U (:76-77: built on temp0 =     */
/* phi U, so no derivative).  After transform (:79) only the last       */
/* backward chain survives (:95-97): W = chain(phi,phi,g) Wz'.          */
/* ------------------------------------------------------------------ */
void orc_stiffness_literal(int P, const double* phi2, const double* U,
                           double* W, const double* G, int64_t num_cells)
{
  const int N = P + 1, nd = N * N * N;
  const double* ph = phi2;
  const double* gr = phi2 + N * N;
  double* t0 = (double*)malloc(sizeof(double) * nd);
  double* t1 = (double*)malloc(sizeof(double) * nd);
  double* wx = (double*)malloc(sizeof(double) * nd);
  double* wy = (double*)malloc(sizeof(double) * nd);
  double* wz = (double*)malloc(sizeof(double) * nd);
  for (int64_t cell = 0; cell < num_cells; cell++)
  {
    const double* Uc = U + cell * nd;
    const double* Gc = G + cell * nd * 6;
    double* Wc = W + cell * nd;
    orc_chain(N, N, gr, ph, ph, Uc, wx, t0, t1);
    orc_chain(N, N, ph, gr, ph, Uc, wy, t0, t1);
    orc_chain(N, N, ph, ph, ph, Uc, wz, t0, t1);
    orc_transform(nd, Gc, wx, wy, wz);
    orc_chain(N, N, ph, ph, gr, wz, Wc, t0, t1);
  }
  free(t0);
  free(t1);
  free(wx);
  free(wy);
  free(wz);
}

/* ------------------------------------------------------------------ */
/* a3: apply_contractions  (src/tmm.hpp:91-126).                       */
/* A[a0][a1][a2] = sum I0(a0,b0) I1(a1,b1) I2(a2,b2) B[b0][b1][b2]     */
/* with I read as [b][a] (I[b*Na+a]) when Tr == 0 (src/tmm.hpp:83) and */
/* as [a][b] (I[a*Nb+b]) when Tr != 0 (src/tmm.hpp:76).  Output        */
/* assigned.  Restated directly as three rotations.                    */
/* ------------------------------------------------------------------ */
void orc_tmm_apply(int Na, int Nb, int Tr, const double* I0, const double* I1,
                   const double* I2, const double* B, double* A)
{
  /* Rotation tables are [src][dst] = [b][a]. */
  const int sz = Na * Nb;
  double* T0 = (double*)malloc(sizeof(double) * sz);
  double* T1 = (double*)malloc(sizeof(double) * sz);
  double* T2 = (double*)malloc(sizeof(double) * sz);
  if (Tr)
  {
    orc_transpose(Na, Nb, I0, T0);
    orc_transpose(Na, Nb, I1, T1);
    orc_transpose(Na, Nb, I2, T2);
  }
  else
  {
    memcpy(T0, I0, sizeof(double) * sz);
    memcpy(T1, I1, sizeof(double) * sz);
    memcpy(T2, I2, sizeof(double) * sz);
  }
  int big = Na > Nb ? Na : Nb;
  double* w0 = (double*)malloc(sizeof(double) * big * big * big);
  double* w1 = (double*)malloc(sizeof(double) * big * big * big);
  orc_rot(Nb, Na, Nb * Nb, T0, B, w0);
  orc_rot(Nb, Na, Nb * Na, T1, w0, w1);
  orc_rot(Nb, Na, Na * Na, T2, w1, A);
  free(T0);
  free(T1);
  free(T2);
  free(w0);
  free(w1);
}

/* ------------------------------------------------------------------ */
/* Correct-mode operators, Q independent of N.                         */
/* Forward with tables as stored [dof][q]; backward with the physical  */
/* transpose [q][dof] (what src/mass.hpp:20-21 meant by phi_t).        */
/* ------------------------------------------------------------------ */
typedef struct
{
  int N, Q;
  double *ph, *gr, *pht, *grt;
  double *f0, *f1, *w0, *w1, *w2, *wm, *b0, *b1, *acc;
} orc_work;

static void orc_work_init(orc_work* w, int N, int Q, const double* tab)
{
  const int big = (N > Q ? N : Q);
  const size_t b3 = (size_t)big * big * big;
  w->N = N;
  w->Q = Q;
  w->ph = (double*)malloc(sizeof(double) * N * Q);
  w->gr = (double*)malloc(sizeof(double) * N * Q);
  w->pht = (double*)malloc(sizeof(double) * N * Q);
  w->grt = (double*)malloc(sizeof(double) * N * Q);
  memcpy(w->ph, tab, sizeof(double) * N * Q);
  memcpy(w->gr, tab + N * Q, sizeof(double) * N * Q);
  orc_transpose(N, Q, w->ph, w->pht);
  orc_transpose(N, Q, w->gr, w->grt);
  double** bufs[] = {&w->f0, &w->f1, &w->w0, &w->w1, &w->w2,
                     &w->wm, &w->b0, &w->b1, &w->acc};
  for (int i = 0; i < 9; i++)
    *bufs[i] = (double*)malloc(sizeof(double) * b3);
}

static void orc_work_free(orc_work* w)
{
  free(w->ph);
  free(w->gr);
  free(w->pht);
  free(w->grt);
  free(w->f0);
  free(w->f1);
  free(w->w0);
  free(w->w1);
  free(w->w2);
  free(w->wm);
  free(w->b0);
  free(w->b1);
  free(w->acc);
}

/* op: 0 = mass, 1 = stiffness, 2 = stiffness + mass (K+M) */
static void orc_cell_apply(orc_work* w, int op, const double* Uc, double* Wc,
                           const double* Gc, const double* Dc)
{
  const int N = w->N, Q = w->Q;
  const int nd = N * N * N, nq = Q * Q * Q;
  if (op == 0 || op == 2)
    orc_chain(N, Q, w->ph, w->ph, w->ph, Uc, w->wm, w->f0, w->f1);
  if (op >= 1)
  {
    orc_chain(N, Q, w->gr, w->ph, w->ph, Uc, w->w0, w->f0, w->f1);
    orc_chain(N, Q, w->ph, w->gr, w->ph, Uc, w->w1, w->f0, w->f1);
    orc_chain(N, Q, w->ph, w->ph, w->gr, Uc, w->w2, w->f0, w->f1);
    orc_transform(nq, Gc, w->w0, w->w1, w->w2);
  }
  if (op == 0 || op == 2)
    for (int i = 0; i < nq; i++)
      w->wm[i] *= Dc[i];
  for (int i = 0; i < nd; i++)
    Wc[i] = 0.0;
  if (op >= 1)
  {
    orc_chain(Q, N, w->grt, w->pht, w->pht, w->w0, w->acc, w->b0, w->b1);
    for (int i = 0; i < nd; i++)
      Wc[i] += w->acc[i];
    orc_chain(Q, N, w->pht, w->grt, w->pht, w->w1, w->acc, w->b0, w->b1);
    for (int i = 0; i < nd; i++)
      Wc[i] += w->acc[i];
    orc_chain(Q, N, w->pht, w->pht, w->grt, w->w2, w->acc, w->b0, w->b1);
    for (int i = 0; i < nd; i++)
      Wc[i] += w->acc[i];
  }
  if (op == 0 || op == 2)
  {
    orc_chain(Q, N, w->pht, w->pht, w->pht, w->wm, w->acc, w->b0, w->b1);
    for (int i = 0; i < nd; i++)
      Wc[i] += w->acc[i];
  }
}

/* Cell-local operator: W[c] = A_c U[c] for every cell (W assigned).
 * tab = [phi | dphi] (dphi may be absent for op 0), G [C][Q^3][6],
 * detJw [C][Q^3]. */
void orc_cell_operator(int op, int N, int Q, const double* tab,
                       const double* U, double* W, const double* G,
                       const double* detJw, int64_t num_cells)
{
  orc_work w;
  double* tab2 = NULL;
  if (op == 0)
  {
    /* mass needs only phi; pad a zero gradient so the workspace is uniform */
    tab2 = (double*)calloc((size_t)2 * N * Q, sizeof(double));
    memcpy(tab2, tab, sizeof(double) * N * Q);
    tab = tab2;
  }
  orc_work_init(&w, N, Q, tab);
  const int nd = N * N * N, nq = Q * Q * Q;
  for (int64_t c = 0; c < num_cells; c++)
    orc_cell_apply(&w, op, U + c * nd, W + c * nd,
                   G ? G + c * nq * 6 : NULL, detJw ? detJw + c * nq : NULL);
  orc_work_free(&w);
  free(tab2);
}

/* ------------------------------------------------------------------ */
/* a8: structured hexahedral dofmap.  Cells (ci,cj,ck) of an           */
/* nx x ny x nz box, cell id (ci*ny + cj)*nz + ck; global dofs are the  */
/* lexicographic (nx P+1) x (ny P+1) x (nz P+1) grid with the a axis   */
/* slowest; dofmap[cell][a][b][c] = ((ci P+a) Gy + cj P+b) Gz + ck P+c. */
/* ------------------------------------------------------------------ */
void orc_dofmap_structured(int nx, int ny, int nz, int P, int32_t* dofmap)
{
  const int N = P + 1, nd = N * N * N;
  const int64_t Gy = (int64_t)ny * P + 1, Gz = (int64_t)nz * P + 1;
  for (int ci = 0; ci < nx; ci++)
    for (int cj = 0; cj < ny; cj++)
      for (int ck = 0; ck < nz; ck++)
      {
        int64_t cell = ((int64_t)ci * ny + cj) * nz + ck;
        int32_t* dm = dofmap + cell * nd;
        for (int a = 0; a < N; a++)
          for (int b = 0; b < N; b++)
            for (int c = 0; c < N; c++)
              dm[(a * N + b) * N + c] =
                  (int32_t)((((int64_t)ci * P + a) * Gy + cj * P + b) * Gz +
                            ck * P + c);
      }
}

/* ------------------------------------------------------------------ */
/* a9: affine geometry.  J[cell] is 3x3 row-major, J[r][s] = dx_r/dxi_s;*/
/* w3[q] the tensor-product weight of point q.  detJw = |det J| w and   */
/* G = |det J| w J^-1 J^-T packed (S22,S12,S02,S11,S01,S00).            */
/* ------------------------------------------------------------------ */
void orc_geometry_affine(int64_t num_cells, int nq3, const double* J,
                         const double* w3, double* G, double* detJw)
{
  for (int64_t c = 0; c < num_cells; c++)
  {
    const double* j = J + 9 * c;
    /* cofactor inverse */
    double det = j[0] * (j[4] * j[8] - j[5] * j[7]) -
                 j[1] * (j[3] * j[8] - j[5] * j[6]) +
                 j[2] * (j[3] * j[7] - j[4] * j[6]);
    double inv[9];
    inv[0] = (j[4] * j[8] - j[5] * j[7]) / det;
    inv[1] = (j[2] * j[7] - j[1] * j[8]) / det;
    inv[2] = (j[1] * j[5] - j[2] * j[4]) / det;
    inv[3] = (j[5] * j[6] - j[3] * j[8]) / det;
    inv[4] = (j[0] * j[8] - j[2] * j[6]) / det;
    inv[5] = (j[2] * j[3] - j[0] * j[5]) / det;
    inv[6] = (j[3] * j[7] - j[4] * j[6]) / det;
    inv[7] = (j[1] * j[6] - j[0] * j[7]) / det;
    inv[8] = (j[0] * j[4] - j[1] * j[3]) / det;
    /* K = J^-1 J^-T : K[r][s] = sum_t inv[r][t] inv[s][t] */
    double K[3][3];
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++)
        K[r][s] = inv[3 * r + 0] * inv[3 * s + 0] +
                  inv[3 * r + 1] * inv[3 * s + 1] +
                  inv[3 * r + 2] * inv[3 * s + 2];
    const double adet = fabs(det);
    for (int q = 0; q < nq3; q++)
    {
      const double s = adet * w3[q];
      if (G)
      {
        double* g = G + (c * nq3 + q) * 6;
        g[0] = s * K[2][2];
        g[1] = s * K[1][2];
        g[2] = s * K[0][2];
        g[3] = s * K[1][1];
        g[4] = s * K[0][1];
        g[5] = s * K[0][0];
      }
      if (detJw)
        detJw[c * nq3 + q] = s;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Continuous operator y = A x through the dofmap (y overwritten).     */
/* ------------------------------------------------------------------ */
void orc_global_operator(int op, int N, int Q, const double* tab,
                         const double* x, double* y, int64_t ndofs,
                         const int32_t* dofmap, const double* G,
                         const double* detJw, int64_t num_cells)
{
  const int nd = N * N * N, nq = Q * Q * Q;
  orc_work w;
  double* tab2 = NULL;
  if (op == 0)
  {
    tab2 = (double*)calloc((size_t)2 * N * Q, sizeof(double));
    memcpy(tab2, tab, sizeof(double) * N * Q);
    tab = tab2;
  }
  orc_work_init(&w, N, Q, tab);
  double* u = (double*)malloc(sizeof(double) * nd);
  double* v = (double*)malloc(sizeof(double) * nd);
  for (int64_t i = 0; i < ndofs; i++)
    y[i] = 0.0;
  for (int64_t c = 0; c < num_cells; c++)
  {
    const int32_t* dm = dofmap + c * nd;
    for (int i = 0; i < nd; i++)
      u[i] = x[dm[i]];
    orc_cell_apply(&w, op, u, v, G ? G + c * nq * 6 : NULL,
                   detJw ? detJw + c * nq : NULL);
    for (int i = 0; i < nd; i++)
      y[dm[i]] += v[i];
  }
  free(u);
  free(v);
  orc_work_free(&w);
  free(tab2);
}

/* Diagonal of the continuous operator, cell by cell:
 * diag_c[i] = e_i^T A_c e_i evaluated directly at the quadrature points. */
void orc_global_diagonal(int op, int N, int Q, const double* tab,
                         double* diag, int64_t ndofs, const int32_t* dofmap,
                         const double* G, const double* detJw,
                         int64_t num_cells)
{
  const int nd = N * N * N, nq = Q * Q * Q;
  const double* ph = tab;
  const double* gr = tab + N * Q;
  for (int64_t i = 0; i < ndofs; i++)
    diag[i] = 0.0;
  for (int64_t c = 0; c < num_cells; c++)
  {
    const int32_t* dm = dofmap + c * nd;
    for (int a = 0; a < N; a++)
      for (int b = 0; b < N; b++)
        for (int cc = 0; cc < N; cc++)
        {
          double s = 0.0;
          for (int qa = 0; qa < Q; qa++)
            for (int qb = 0; qb < Q; qb++)
              for (int qc = 0; qc < Q; qc++)
              {
                const int q = (qa * Q + qb) * Q + qc;
                const double pa = ph[a * Q + qa], pb = ph[b * Q + qb],
                             pc = ph[cc * Q + qc];
                if (op >= 1)
                {
                  const double d0 = gr[a * Q + qa] * pb * pc;
                  const double d1 = pa * gr[b * Q + qb] * pc;
                  const double d2 = pa * pb * gr[cc * Q + qc];
                  const double* g = G + (c * nq + q) * 6;
                  s += g[5] * d0 * d0 + g[3] * d1 * d1 + g[0] * d2 * d2 +
                       2.0 * (g[4] * d0 * d1 + g[2] * d0 * d2 +
                              g[1] * d1 * d2);
                }
                if (op == 0 || op == 2)
                {
                  const double m = pa * pb * pc;
                  s += detJw[c * nq + q] * m * m;
                }
              }
          diag[dm[(a * N + b) * N + cc]] += s;
        }
  }
}

/* ------------------------------------------------------------------ */
/* a11: Jacobi-preconditioned CG in FP64 on the continuous operator.   */
/* Solves A u = f from u = 0.  Stops when ||r||_2 <= rtol ||f||_2 or   */
/* after maxit iterations.  Returns the iteration count; the final     */
/* relative residual is stored in *relres.                             */
/* ------------------------------------------------------------------ */
int orc_pcg(int op, int N, int Q, const double* tab, const double* f,
            double* u, int64_t ndofs, const int32_t* dofmap, const double* G,
            const double* detJw, int64_t num_cells, double rtol, int maxit,
            double* relres)
{
  double* r = (double*)malloc(sizeof(double) * ndofs);
  double* z = (double*)malloc(sizeof(double) * ndofs);
  double* p = (double*)malloc(sizeof(double) * ndofs);
  double* Ap = (double*)malloc(sizeof(double) * ndofs);
  double* dg = (double*)malloc(sizeof(double) * ndofs);
  orc_global_diagonal(op, N, Q, tab, dg, ndofs, dofmap, G, detJw, num_cells);
  double fnorm = 0.0;
  for (int64_t i = 0; i < ndofs; i++)
  {
    u[i] = 0.0;
    r[i] = f[i];
    z[i] = r[i] / dg[i];
    p[i] = z[i];
    fnorm += f[i] * f[i];
  }
  fnorm = sqrt(fnorm);
  double rz = 0.0, rr = 0.0;
  for (int64_t i = 0; i < ndofs; i++)
  {
    rz += r[i] * z[i];
    rr += r[i] * r[i];
  }
  int it = 0;
  if (fnorm == 0.0)
  {
    *relres = 0.0;
    goto done;
  }
  while (sqrt(rr) > rtol * fnorm && it < maxit)
  {
    orc_global_operator(op, N, Q, tab, p, Ap, ndofs, dofmap, G, detJw,
                        num_cells);
    double pAp = 0.0;
    for (int64_t i = 0; i < ndofs; i++)
      pAp += p[i] * Ap[i];
    const double alpha = rz / pAp;
    double rz_new = 0.0;
    rr = 0.0;
    for (int64_t i = 0; i < ndofs; i++)
    {
      u[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
      z[i] = r[i] / dg[i];
      rz_new += r[i] * z[i];
      rr += r[i] * r[i];
    }
    const double beta = rz_new / rz;
    rz = rz_new;
    for (int64_t i = 0; i < ndofs; i++)
      p[i] = z[i] + beta * p[i];
    it++;
  }
  *relres = sqrt(rr) / fnorm;
done:
  free(r);
  free(z);
  free(p);
  free(Ap);
  free(dg);
  return it;
}
