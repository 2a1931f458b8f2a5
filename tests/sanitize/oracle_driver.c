/* Exercises every entry point of the CPU oracle (oracle/femtc_oracle.c) on
 * small inputs, for AddressSanitizer / UndefinedBehaviorSanitizer builds
 * (tests/test_sanitizers.py).  The values are arbitrary: this checks memory
 * and arithmetic safety, not results. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void orc_fill_uniform(uint64_t seed, int64_t n, double lo, double hi,
                      double* out);
void orc_gemm_micro(int k, int m, int nc, int lda, int ldb, int ldc,
                    int order, const double* a, const double* b, double* c);
void orc_gemm_blocked(int k, int m, int n, int order, int MB, int NB,
                      const double* a, const double* b, double* c);
void orc_transform(int nq, const double* G, double* w0, double* w1,
                   double* w2);
void orc_mass_literal(int P, const double* phi, const double* U, double* W,
                      const double* detJ, int64_t num_cells);
void orc_stiffness_literal(int P, const double* phi2, const double* U,
                           double* W, const double* G, int64_t num_cells);
void orc_tmm_apply(int Na, int Nb, int Tr, const double* I0, const double* I1,
                   const double* I2, const double* B, double* A);
void orc_cell_operator(int op, int N, int Q, const double* tab,
                       const double* U, double* W, const double* G,
                       const double* detJw, int64_t num_cells);
void orc_dofmap_structured(int nx, int ny, int nz, int P, int32_t* dofmap);
void orc_geometry_affine(int64_t num_cells, int nq3, const double* J,
                         const double* w3, double* G, double* detJw);
void orc_global_operator(int op, int N, int Q, const double* tab,
                         const double* x, double* y, int64_t ndofs,
                         const int32_t* dofmap, const double* G,
                         const double* detJw, int64_t num_cells);
void orc_global_diagonal(int op, int N, int Q, const double* tab,
                         double* diag, int64_t ndofs, const int32_t* dofmap,
                         const double* G, const double* detJw,
                         int64_t num_cells);
int orc_pcg(int op, int N, int Q, const double* tab, const double* f,
            double* u, int64_t ndofs, const int32_t* dofmap, const double* G,
            const double* detJw, int64_t num_cells, double rtol, int maxit,
            double* relres);

static double* vec(uint64_t seed, int64_t n, double lo, double hi)
{
  double* v = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  orc_fill_uniform(seed, n, lo, hi, v);
  return v;
}

int main(void)
{
  /* GEMM building blocks: every loop order and block size */
  for (int P = 1; P <= 6; ++P)
  {
    const int N = P + 1;
    double* a = vec(1, N * N, -1, 1);
    double* b = vec(2, N * N * N, -1, 1);
    double* c = vec(3, N * N * N, 0, 0);
    for (int order = 0; order < 6; ++order)
    {
      orc_gemm_micro(N, N, N * N, N, N * N, N * N, order, a, b, c);
      const int blocks[4] = {0, 4, 8, 16};
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          orc_gemm_blocked(N, N, N * N, order, blocks[i], blocks[j], a, b, c);
    }
    free(a);
    free(b);
    free(c);
  }
  /* literal reference operators, tmm, cell operators with Q != N */
  for (int P = 1; P <= 5; ++P)
  {
    const int N = P + 1, C = 3;
    double* phi2 = vec(4, 2 * N * N, -1, 1);
    double* U = vec(5, C * N * N * N, -1, 1);
    double* W = vec(6, C * N * N * N, 0, 0);
    double* detJ = vec(7, C * N * N * N, 0.5, 1.5);
    double* G = vec(8, C * N * N * N * 6, 0, 1);
    orc_mass_literal(P, phi2, U, W, detJ, C);
    orc_stiffness_literal(P, phi2, U, W, G, C);
    for (int Q = N; Q <= N + 2; ++Q)
    {
      double* tab = vec(9, 2 * N * Q, -1, 1);
      double* Gq = vec(10, C * Q * Q * Q * 6, 0, 1);
      double* Dq = vec(11, C * Q * Q * Q, 0.5, 1.5);
      for (int op = 0; op < 3; ++op)
        orc_cell_operator(op, N, Q, tab, U, W, Gq, Dq, C);
      double* I = vec(12, N * Q, -1, 1);
      double* A = vec(13, Q * Q * Q, 0, 0);
      orc_tmm_apply(Q, N, 0, I, I, I, U, A);
      orc_tmm_apply(N, Q, 1, I, I, I, A, W);
      orc_transform(Q * Q * Q, Gq, A, A, A);
      free(tab);
      free(Gq);
      free(Dq);
      free(I);
      free(A);
    }
    free(phi2);
    free(U);
    free(W);
    free(detJ);
    free(G);
  }
  /* continuous operators, diagonal and PCG on a 2 x 3 x 2 mesh */
  for (int P = 1; P <= 4; ++P)
  {
    const int N = P + 1, Q = P + 2, nx = 2, ny = 3, nz = 2;
    const int64_t C = nx * ny * nz;
    const int64_t ng = (int64_t)(nx * P + 1) * (ny * P + 1) * (nz * P + 1);
    int32_t* dm = (int32_t*)malloc(sizeof(int32_t) * C * N * N * N);
    orc_dofmap_structured(nx, ny, nz, P, dm);
    double* J = vec(14, 9 * C, -0.1, 0.1);
    for (int64_t c = 0; c < C; ++c)
      for (int d = 0; d < 3; ++d)
        J[9 * c + 4 * d] += 1.0;
    double* w3 = vec(15, Q * Q * Q, 0.001, 0.01);
    double* G = vec(0, C * Q * Q * Q * 6, 0, 0);
    double* D = vec(0, C * Q * Q * Q, 0, 0);
    orc_geometry_affine(C, Q * Q * Q, J, w3, G, D);
    double* tab = vec(16, 2 * N * Q, -1, 1);
    double* x = vec(17, ng, -1, 1);
    double* y = vec(18, ng, 0, 0);
    for (int op = 0; op < 3; ++op)
    {
      orc_global_operator(op, N, Q, tab, x, y, ng, dm, G, D, C);
      orc_global_diagonal(op, N, Q, tab, y, ng, dm, G, D, C);
    }
    double rel = 0.0;
    (void)orc_pcg(2, N, Q, tab, x, y, ng, dm, G, D, C, 1e-8, 50, &rel);
    free(dm);
    free(J);
    free(w3);
    free(G);
    free(D);
    free(tab);
    free(x);
    free(y);
  }
  printf("oracle driver done\n");
  return 0;
}
