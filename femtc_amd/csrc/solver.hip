// Device-resident Jacobi-preconditioned conjugate gradients on the continuous
// operator A = M, K or K+M (applied through the dofmap), and FP64 defect
// correction around a lower-precision inner PCG.
//
// Plain solve (inner dtype FP64): PCG on A u = f in FP64.
// Refined solve (inner dtype FP32): the outer loop keeps u, f and the
// residual r = f - A u in FP64 (geometry, tables and the operator in FP64);
// each step solves A_32 e = r / ||r|| with FP32 vectors, FP32 tables and an
// FP32 copy of the geometry to a loose tolerance, then u += ||r|| e and the
// residual is recomputed with the FP64 operator.  Every dot product and norm
// is accumulated in FP64, in a fixed order (per-block partial sums, then one
// block), so a solve is deterministic up to the operator's atomic scatter.
//
// The CG scalars live on the device (PcgState): the host queues batches of
// iterations without reading anything back; a reduction kernel raises
// `done` on convergence, breakdown or the iteration cap, and every kernel of
// the iteration (the operator included) returns at once when it is set.  The
// host checks `done` once per batch.  p.Ap is not a separate pass over p and
// Ap: the operator kernel returns per-wave partials of the quadrature-point
// energy sum_q (d.S d + detJw u^2) = sum_e p_e^T A_e p_e, which is p^T A p.
// With the assembled scatter (structured dofmap) the operator's cell pass
// stores its cell values without atomics and the CG update kernel sums them
// per dof itself, so Ap is never stored.
#include <algorithm>
#include <vector>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "assemble.hpp"
#include "common.hpp"

namespace femtc
{
namespace
{

constexpr int RB = 256;    // threads per block, vector kernels
constexpr int NBLK = 1024; // blocks of the vector kernels (fixed: the partial
                           // sums are always combined in the same order)

struct PcgState
{
  double rz, rr, pAp, alpha, beta, tol2, rr0;
  int it, maxit, done, status;
  // alpha of this iteration is valid and e += alpha p is still to be
  // applied (by k_pupdate, which reads p anyway)
  int epend;
};

// Sum of NV values over the block of BT threads; the result is valid in
// thread 0.
template <int NV, int BT = RB>
__device__ inline void block_sum(double (&v)[NV])
{
  __shared__ double sh[NV][BT / 64];
#pragma unroll
  for (int k = 0; k < NV; ++k)
  {
    double x = v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      x += __shfl_down(x, o, 64);
    v[k] = x;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
  {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      sh[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0)
  {
#pragma unroll
    for (int k = 0; k < NV; ++k)
    {
      double s = 0.0;
      for (int j = 0; j < BT / 64; ++j)
        s += sh[k][j];
      v[k] = s;
    }
  }
}

template <int NV, int BT = RB>
__device__ inline void store_partials(double (&v)[NV], double* part)
{
  block_sum<NV, BT>(v);
  if (threadIdx.x == 0)
  {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      part[k * NBLK + blockIdx.x] = v[k];
  }
}

// Combine the NBLK partials of NV sums (one block of RB threads).
template <int NV>
__device__ inline void combine_partials(const double* part, double (&v)[NV])
{
#pragma unroll
  for (int k = 0; k < NV; ++k)
  {
    double s = 0.0;
    for (int j = threadIdx.x; j < NBLK; j += RB)
      s += part[k * NBLK + j];
    v[k] = s;
  }
  block_sum<NV>(v);
}

// The Jacobi preconditioner D^-1 of a PCG in precision V is held as
// Precond<V>::T: FP64 for the FP64 solve, and bfloat16 (the upper half of
// the FP32 value, rounded to nearest even) for the FP32 inner solve of the
// refinement.  Any fixed positive diagonal is a valid preconditioner; the
// rounded one is within 2^-9 of D^-1 entry by entry, which moves the
// condition number of the preconditioned operator by under 0.4 %, and the
// outer FP64 loop fixes the accuracy of the result.  The streaming passes
// then read 2 bytes per dof for it instead of 4 (the p-update at 289^3
// dofs: 95 -> 84 us; the refined solve with the assembly update reading
// them too: 107.2 -> 105.3 ms).
template <typename V>
struct Precond
{
  using T = V;
};
template <>
struct Precond<float>
{
  using T = uint16_t;
};

__device__ inline double precond_value(double d) { return d; }
__device__ inline float precond_value(uint16_t d)
{
  return __uint_as_float((uint32_t)d << 16);
}

#define GRID_LOOP(i, n)                                                      \
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < (n);          \
       i += (int64_t)NBLK * RB)

// ---------------------------------------------------------------- PCG ----

// e = 0, r = rhs, p = z = D^-1 r, Ap = 0; partials of r.z and r.r
template <typename V>
__global__ void __launch_bounds__(RB)
    k_start(const V* __restrict__ rhs, V* __restrict__ e, V* __restrict__ r,
            V* __restrict__ p, V* __restrict__ Ap,
            const typename Precond<V>::T* __restrict__ dinv, int64_t n,
            double* __restrict__ part)
{
  double s[2] = {0.0, 0.0};
  GRID_LOOP(i, n)
  {
    const V ri = rhs[i];
    const V zi = (V)precond_value(dinv[i]) * ri;
    e[i] = V(0);
    r[i] = ri;
    p[i] = zi;
    Ap[i] = V(0);
    s[0] += (double)ri * (double)zi;
    s[1] += (double)ri * (double)ri;
  }
  store_partials<2>(s, part);
}

__global__ void __launch_bounds__(RB)
    k_reduce_start(const double* __restrict__ part, PcgState* st, double rtol,
                   int maxit)
{
  double v[2];
  combine_partials<2>(part, v);
  if (threadIdx.x == 0)
  {
    PcgState s{};
    s.rz = v[0];
    s.rr = v[1];
    s.rr0 = v[1];
    s.tol2 = rtol * rtol * v[1];
    s.maxit = maxit;
    s.it = 0;
    s.status = FEMTC_OK;
    s.done = 0;
    if (!(v[1] > s.tol2) || v[1] == 0.0)
      s.done = 1; // zero right-hand side (or nothing to do)
    if (!__builtin_isfinite(v[0]) || !__builtin_isfinite(v[1]))
    {
      s.done = 1;
      s.status = FEMTC_ERR_BREAKDOWN;
    }
    *st = s;
  }
}

// First stage of p.Ap: the operator kernel's per-wave quadrature-point
// energies p_e^T A_e p_e (npart of them) summed into NBLK block partials.
__global__ void __launch_bounds__(RB)
    k_energy_partials(const double* __restrict__ e, int64_t npart,
                      double* __restrict__ part,
                      const PcgState* __restrict__ st)
{
  if (st->done)
    return;
  double s[1] = {0.0};
  GRID_LOOP(i, npart) { s[0] += e[i]; }
  store_partials<1>(s, part);
}

// alpha = r.z / p.Ap (block partials combined in a fixed order)
__global__ void __launch_bounds__(RB)
    k_alpha(const double* __restrict__ part, PcgState* st)
{
  if (st->done)
  {
    if (threadIdx.x == 0)
      st->epend = 0; // the last step's e += alpha p has been applied
    return;
  }
  double v[1];
  combine_partials<1>(part, v);
  if (threadIdx.x == 0)
  {
    st->pAp = v[0];
    if (!(v[0] > 0.0) || !__builtin_isfinite(v[0]))
    {
      st->done = 1;
      st->status = FEMTC_ERR_BREAKDOWN;
      st->epend = 0;
    }
    else
    {
      st->alpha = st->rz / v[0];
      st->epend = 1;
    }
  }
}

// r -= alpha Ap, Ap = 0 (ready for the next scatter); partials of r.z and
// r.r with z = D^-1 r (e += alpha p waits for k_pupdate, which reads p)
template <typename V>
__global__ void __launch_bounds__(RB)
    k_update(V* __restrict__ r, V* __restrict__ Ap,
             const typename Precond<V>::T* __restrict__ dinv, int64_t n,
             const PcgState* __restrict__ st, double* __restrict__ part)
{
  if (st->done)
    return;
  const V a = (V)st->alpha;
  double s[2] = {0.0, 0.0};
  GRID_LOOP(i, n)
  {
    const V ri = r[i] - a * Ap[i];
    r[i] = ri;
    Ap[i] = V(0);
    const V zi = (V)precond_value(dinv[i]) * ri;
    s[0] += (double)ri * (double)zi;
    s[1] += (double)ri * (double)ri;
  }
  store_partials<2>(s, part);
}

// The same update with the assembled scatter: Ap_i is summed from the cell
// pass's lattice values (no Ap vector, no atomics, nothing to clear).  The
// fixed grid's waves take the global dof rows in turn (one row per wave,
// lanes along the row), so the partial sums are still in a fixed order.
// Its workgroups are ARB threads: twice the waves of the other vector
// kernels, for more loads in flight.
constexpr int ARB = 512;

template <typename V, int P>
__global__ void __launch_bounds__(ARB)
    k_assemble_update(const V* __restrict__ W, V* __restrict__ r,
                      const typename Precond<V>::T* __restrict__ dinv,
                      LatticeGrid L,
                      const PcgState* __restrict__ st,
                      double* __restrict__ part)
{
  if (st->done)
    return;
  const V a = (V)st->alpha;
  double s[2] = {0.0, 0.0};
  const long long nrows = lattice_global_rows(L);
  const int NZ = L.ncz * P + 1;
  const int lane = threadIdx.x & 63;
  // the row is the same on every lane of a wave: made wave-uniform, its
  // index arithmetic and the lattice row pointers live in scalar registers
  const int wave0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (long long row = (long long)blockIdx.x * (ARB / 64) + wave0;
       row < nrows; row += (long long)NBLK * (ARB / 64))
  {
    const V* rows[4];
    const int mask = lattice_rows<V, P>(W, (int)row, L, rows);
    const int nr = lattice_rows_front(rows, mask);
    V* rr = r + row * NZ;
    const typename Precond<V>::T* dr = dinv + row * NZ;
    with_row_count(nr, [&](auto NRc) {
      constexpr int NR = decltype(NRc)::value;
#pragma unroll 2
      for (int gk = lane; gk < NZ; gk += 64)
      {
        const V ri =
            at(rr, gk) - a * lattice_row_sum<V, P, NR>(rows, gk, L.ncz);
        at(rr, gk) = ri;
        const V zi = (V)precond_value(at(dr, gk)) * ri;
        s[0] += (double)ri * (double)zi;
        s[1] += (double)ri * (double)ri;
      }
    });
  }
  store_partials<2, ARB>(s, part);
}

__global__ void __launch_bounds__(RB)
    k_beta(const double* __restrict__ part, PcgState* st)
{
  if (st->done)
    return;
  double v[2];
  combine_partials<2>(part, v);
  if (threadIdx.x == 0)
  {
    const double rzn = v[0];
    st->beta = rzn / st->rz;
    st->rz = rzn;
    st->rr = v[1];
    st->it += 1;
    if (!__builtin_isfinite(v[0]) || !__builtin_isfinite(v[1]))
    {
      st->done = 1;
      st->status = FEMTC_ERR_BREAKDOWN;
    }
    else if (v[1] <= st->tol2)
    {
      st->done = 1;
      st->status = FEMTC_OK;
    }
    else if (st->it >= st->maxit)
    {
      st->done = 1;
      st->status = FEMTC_ERR_NOT_CONVERGED;
    }
  }
}

// e += alpha p (this iteration's step), then p = D^-1 r + beta p unless the
// iteration ended the solve
template <typename V>
__global__ void __launch_bounds__(RB)
    k_pupdate(V* __restrict__ p, V* __restrict__ e, const V* __restrict__ r,
              const typename Precond<V>::T* __restrict__ dinv, int64_t n,
              const PcgState* __restrict__ st)
{
  if (!st->epend)
    return;
  const V a = (V)st->alpha;
  if (st->done)
  {
    GRID_LOOP(i, n) { e[i] += a * p[i]; }
    return;
  }
  const V b = (V)st->beta;
  GRID_LOOP(i, n)
  {
    const V pi = p[i];
    e[i] += a * pi;
    p[i] = (V)precond_value(dinv[i]) * r[i] + b * pi;
  }
}

// ------------------------------------------------------- outer loop ----

// partial sums of x.x
__global__ void __launch_bounds__(RB)
    k_norm2(const double* __restrict__ x, int64_t n, double* __restrict__ part)
{
  double s[1] = {0.0};
  GRID_LOOP(i, n) { s[0] += x[i] * x[i]; }
  store_partials<1>(s, part);
}

__global__ void __launch_bounds__(RB)
    k_combine1(const double* __restrict__ part, double* out)
{
  double v[1];
  combine_partials<1>(part, v);
  if (threadIdx.x == 0)
    *out = v[0];
}

// r = f - Au (FP64); partial sums of r.r
__global__ void __launch_bounds__(RB)
    k_residual(const double* __restrict__ f, const double* __restrict__ Au,
               double* __restrict__ r, int64_t n, double* __restrict__ part)
{
  double s[1] = {0.0};
  GRID_LOOP(i, n)
  {
    const double ri = f[i] - Au[i];
    r[i] = ri;
    s[0] += ri * ri;
  }
  store_partials<1>(s, part);
}

template <typename V>
__global__ void __launch_bounds__(RB)
    k_scale_to(const double* __restrict__ x, double s, V* __restrict__ y,
               int64_t n)
{
  GRID_LOOP(i, n) { y[i] = (V)(x[i] * s); }
}

// u += s e
template <typename V>
__global__ void __launch_bounds__(RB)
    k_correct(double* __restrict__ u, const V* __restrict__ e, double s,
              int64_t n)
{
  GRID_LOOP(i, n) { u[i] += s * (double)e[i]; }
}

template <typename A, typename B>
__global__ void __launch_bounds__(RB)
    k_convert(const A* __restrict__ x, B* __restrict__ y, int64_t n)
{
  GRID_LOOP(i, n) { y[i] = (B)x[i]; }
}

// D^-1 for the FP32 inner solve: FP64 -> FP32 -> bfloat16, rounded to
// nearest even (the values are finite: k_invert writes 1/d or 0)
__global__ void __launch_bounds__(RB)
    k_precond_bf16(const double* __restrict__ x, uint16_t* __restrict__ y,
                   int64_t n)
{
  GRID_LOOP(i, n)
  {
    const uint32_t u = __float_as_uint((float)x[i]);
    const uint32_t h = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    y[i] = (uint16_t)h;
  }
}

__global__ void __launch_bounds__(RB)
    k_invert(const double* __restrict__ d, double* __restrict__ dinv,
             int64_t n, int* __restrict__ bad)
{
  GRID_LOOP(i, n)
  {
    const double v = d[i];
    if (!(v > 0.0))
      bad[0] = 1; // every writer stores the same value
    dinv[i] = v > 0.0 ? 1.0 / v : 0.0;
  }
}

// --------------------------------------------------------- diagonal ----
//
// diag(A) of the continuous operator, summed cell by cell:
//   D_c[a][b][c] = sum_q sum_t F_t(q) X_t(a,qa) Y_t(b,qb) Z_t(c,qc)
// with seven terms t (X, Y, Z each one of phi^2, g^2 or phi g):
//   S00 (g^2,  phi^2, phi^2)   S11 (phi^2, g^2, phi^2)   S22 (phi^2, phi^2, g^2)
//   2 S01 (phi g, phi g, phi^2)  2 S02 (phi g, phi^2, phi g)
//   2 S12 (phi^2, phi g, phi g)  detJw (phi^2, phi^2, phi^2)
// each evaluated by sum factorization (qc, then qb, then qa).  One workgroup
// per cell; the tables and partial tensors live in LDS.  Setup kernel.
__device__ inline double tab_kind(const double* ph, const double* gr, int kind,
                                  int i)
{
  const double p = ph[i], g = gr[i];
  return kind == 0 ? p * p : (kind == 1 ? g * g : p * g);
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_diag(int op, int N, int Q, const T* __restrict__ tab,
           const T* __restrict__ G, const T* __restrict__ detJw,
           const T* __restrict__ cellgeo, const int32_t* __restrict__ dofmap,
           int64_t num_cells, double* __restrict__ diag)
{
  extern __shared__ double sm[];
  double* ph = sm;
  double* gr = ph + N * Q;
  double* F = gr + N * Q;
  double* T1 = F + Q * Q * Q;     // [qa][qb][c]
  double* T2 = T1 + Q * Q * N;    // [qa][b][c]
  double* wt = T2 + Q * N * N;    // Gauss weights (affine cells)
  const int64_t cell = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int Q3 = Q * Q * Q, N3 = N * N * N, NN = N * N;
  // with cellgeo, tables are [phi | dphi | w] and the geometry of point q is
  // w(qa) w(qb) w(qc) times the cell's 7 values
  const int ntab = 2 * N * Q + (cellgeo ? Q : 0);
  for (int i = tid; i < ntab; i += nt)
    sm[i < 2 * N * Q ? i : (int)(wt - sm) + i - 2 * N * Q] = (double)tab[i];
  // term -> (G slot, factor, X, Y, Z); slot 6 = detJw
  const int slot[7] = {5, 3, 0, 4, 2, 1, 6};
  const double fac[7] = {1, 1, 1, 2, 2, 2, 1};
  const int kx[7] = {1, 0, 0, 2, 2, 0, 0};
  const int ky[7] = {0, 1, 0, 2, 0, 2, 0};
  const int kz[7] = {0, 0, 1, 0, 2, 2, 0};
  double acc[3] = {0.0, 0.0, 0.0}; // dofs tid, tid+nt, tid+2nt (N3 <= 729)
#pragma unroll
  for (int t = 0; t < 7; ++t)
  {
    if (t < 6 && op == FEMTC_OP_MASS)
      continue;
    if (t == 6 && op == FEMTC_OP_STIFFNESS)
      continue;
    __syncthreads(); // previous term's readers are done
    __syncthreads(); // tables and weights are in LDS
    for (int q = tid; q < Q3; q += nt)
    {
      if (cellgeo)
        F[q] = fac[t] * wt[q / (Q * Q)] * wt[(q / Q) % Q] * wt[q % Q] *
               (double)cellgeo[cell * CELLGEO + slot[t]];
      else
        F[q] = fac[t] * (slot[t] == 6 ? (double)detJw[cell * Q3 + q]
                                      : (double)G[(cell * Q3 + q) * 6 + slot[t]]);
    }
    __syncthreads();
    for (int i = tid; i < Q * Q * N; i += nt)
    {
      const int c = i % N, qab = i / N;
      double s = 0.0;
      for (int qc = 0; qc < Q; ++qc)
        s += F[qab * Q + qc] * tab_kind(ph, gr, kz[t], c * Q + qc);
      T1[i] = s;
    }
    __syncthreads();
    for (int i = tid; i < Q * NN; i += nt)
    {
      const int c = i % N, b = (i / N) % N, qa = i / NN;
      double s = 0.0;
      for (int qb = 0; qb < Q; ++qb)
        s += T1[(qa * Q + qb) * N + c] * tab_kind(ph, gr, ky[t], b * Q + qb);
      T2[i] = s;
    }
    __syncthreads();
    for (int k = 0; k < 3; ++k)
    {
      const int i = tid + k * nt;
      if (i < N3)
      {
        const int bc = i % NN, a = i / NN;
        double s = 0.0;
        for (int qa = 0; qa < Q; ++qa)
          s += T2[qa * NN + bc] * tab_kind(ph, gr, kx[t], a * Q + qa);
        acc[k] += s;
      }
    }
  }
  for (int k = 0; k < 3; ++k)
  {
    const int i = tid + k * nt;
    if (i < N3)
      unsafeAtomicAdd(diag + dofmap[cell * N3 + i], acc[k]);
  }
}

size_t diag_lds_bytes(int N, int Q)
{
  return sizeof(double) *
         (size_t)(2 * N * Q + Q * Q * Q + Q * Q * N + Q * N * N + Q);
}

int launch_diag(int op, int dtype, int N, int Q, const void* tab,
                const void* G, const void* detJw, const int32_t* dofmap,
                int64_t num_cells, double* diag, int64_t ndofs,
                hipStream_t s, const void* cellgeo = nullptr)
{
  int st = fill_zero(diag, ndofs, FEMTC_F64, s);
  if (st != FEMTC_OK)
    return st;
  if (num_cells == 0)
    return FEMTC_OK;
  if (num_cells > 0x7fffffffLL)
    return FEMTC_ERR_INVALID_ARGUMENT;
  const size_t lds = diag_lds_bytes(N, Q);
  if (dtype == FEMTC_F64)
    hipLaunchKernelGGL(k_diag<double>, dim3((unsigned)num_cells), dim3(256),
                       lds, s, op, N, Q, (const double*)tab,
                       (const double*)G, (const double*)detJw,
                       (const double*)cellgeo, dofmap, num_cells, diag);
  else
    hipLaunchKernelGGL(k_diag<float>, dim3((unsigned)num_cells), dim3(256),
                       lds, s, op, N, Q, (const float*)tab, (const float*)G,
                       (const float*)detJw, (const float*)cellgeo, dofmap,
                       num_cells, diag);
  return hip_status(hipGetLastError());
}

// ----------------------------------------------------------- solver ----

template <typename V>
constexpr int dtype_of()
{
  return sizeof(V) == 8 ? FEMTC_F64 : FEMTC_F32;
}

} // namespace
} // namespace femtc

using namespace femtc;

// The solver object behind the opaque femtc_solver handle.
struct femtc_solver
{
  femtc_problem prob{};
  int inner_dtype = FEMTC_F64;
  int N = 0, Q = 0;
  int64_t n = 0;
  hipStream_t stream = nullptr;
  // FP64
  double* diag = nullptr;
  double* dinv64 = nullptr;
  double* r64 = nullptr;
  double* Au64 = nullptr;
  double* part = nullptr;
  double* dscal = nullptr; // device scalar for norms
  int* dbad = nullptr;
  PcgState* st = nullptr;
  PcgState* hst = nullptr; // pinned host copy
  double* hscal = nullptr; // pinned host scalar
  // inner precision (FP64 aliases the FP64 buffers)
  void* rhs = nullptr;
  void* e = nullptr;
  void* r = nullptr;
  void* p = nullptr;
  void* Ap = nullptr;
  void* dinv = nullptr;   // D^-1 as Precond<V>::T
  void* tab = nullptr;    // inner-precision tables and geometry
  void* G = nullptr;     // per point, or per cell (inner_affine)
  void* detJw = nullptr;
  // affine cells (problem.J given): the inner operator computes its
  // geometry from 7 values per cell; so does the FP64 operator when the
  // problem has no per-point geometry (outer_affine)
  bool inner_affine = false;
  bool outer_affine = false;
  double* tabw64 = nullptr; // [phi | dphi | w], FP64
  double* cg64 = nullptr;   // [C][7], FP64
  double* epart = nullptr;  // per-wave p.Ap partials of the inner operator
  int64_t nepart = 0;
  int64_t nepart_cap = 0;   // partials the epart allocation holds
  int check_every = 16;
  // assembled scatter of the inner operator (problem.grid, structured
  // dofmap): its lattice buffer in the inner precision
  bool inner_assembled = false;
  void* lattice = nullptr;
  // structured z-run scatter of the inner operator (the atomic scatter's
  // Ap and k_update)
  bool inner_zrun = false;
  // even-odd table sets (symmetric tables, nq = degree + 1): FP64 for the
  // FP64 operators, inner precision for the inner one (aliases eo64 for an
  // FP64 inner solve)
  bool even_odd = false;
  void* eo64 = nullptr;
  void* eo_in = nullptr;
  // inner operator on the matrix cores (inner_dtype FEMTC_BF16 or
  // FEMTC_BF16_ROUNDED): -1 when not, else the MfmaReq split flag; its
  // operand set (tables as A operands, then the weights) on the device
  int mfma_split = -1;
  void* mfma_ops = nullptr;
  // the batch of check_every iterations as a hipGraph (captured on
  // cap_stream at the first solve, replayed on `stream`)
  bool use_graph = true;
  bool graph_failed = false;
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t graph = nullptr;
  // device bytes allocated for the workspace (femtc_solver_bytes)
  size_t bytes = 0;
};

namespace femtc
{
namespace
{

// Device allocation of solver S, counted in S->bytes.
int dev_alloc(femtc_solver* S, void** p, size_t bytes)
{
  if (bytes == 0)
    bytes = 8;
  if (hipMalloc(p, bytes) != hipSuccess)
  {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(FEMTC_ERR_OUT_OF_MEMORY,
                "solver workspace allocation failed (" +
                    std::to_string(bytes) + " bytes)");
  }
  S->bytes += bytes;
  return FEMTC_OK;
}

void dev_free(void* p)
{
  if (p)
    (void)hipFree(p);
}

void destroy(femtc_solver* S)
{
  if (!S)
    return;
  const bool own_inner = S->inner_dtype != FEMTC_F64;
  dev_free(S->diag);
  dev_free(S->dinv64);
  dev_free(S->r64);
  dev_free(S->Au64);
  dev_free(S->part);
  dev_free(S->dscal);
  dev_free(S->dbad);
  dev_free(S->st);
  if (S->hst)
    (void)hipHostFree(S->hst);
  if (S->hscal)
    (void)hipHostFree(S->hscal);
  dev_free(S->rhs);
  dev_free(S->e);
  dev_free(S->p);
  dev_free(S->Ap);
  if (own_inner)
  {
    dev_free(S->r);
    dev_free(S->dinv);
    dev_free(S->tab);
    dev_free(S->G);
    dev_free(S->detJw);
  }
  if (S->eo_in != S->eo64)
    dev_free(S->eo_in);
  dev_free(S->eo64);
  dev_free(S->tabw64);
  dev_free(S->cg64);
  dev_free(S->epart);
  dev_free(S->lattice);
  dev_free(S->mfma_ops);
  if (S->graph)
    (void)hipGraphExecDestroy(S->graph);
  if (S->cap_stream)
    (void)hipStreamDestroy(S->cap_stream);
  delete S;
}

// The operator request of the inner solve (precision V) or of the FP64
// outer residual.
ApplyReq inner_req(const femtc_solver* S, int dtype, const void* x, void* y);
ApplyReq outer_req(const femtc_solver* S, const void* x, void* y);

ApplyReq operator_req(const femtc_solver* S, int dtype, const void* tab,
                      const void* G, const void* detJw, const void* x,
                      void* y)
{
  ApplyReq a;
  a.op = S->prob.op;
  a.dtype = dtype;
  a.N = S->N;
  a.Q = S->Q;
  a.gather = 1;
  a.tab = tab;
  a.x = x;
  a.y = y;
  a.G = G;
  a.detJw = detJw;
  a.dofmap = S->prob.dofmap;
  a.num_cells = S->prob.num_cells;
  a.stream = S->stream;
  if (S->even_odd)
    a.eo_tab = dtype == FEMTC_F64 ? S->eo64 : S->eo_in;
  return a;
}

ApplyReq inner_req(const femtc_solver* S, int dtype, const void* x, void* y)
{
  ApplyReq a = operator_req(S, dtype, S->tab, S->G, S->detJw, x, y);
  if (S->inner_affine)
  {
    a.op += OP_MASS_AFF; // the same operator with per-cell geometry
    a.detJw = nullptr;
  }
  if (S->inner_assembled)
  {
    // the cell pass only: the CG update kernel assembles
    a.assembled = 2;
    a.lattice = S->lattice;
    a.ncx = S->prob.grid[0];
    a.ncy = S->prob.grid[1];
    a.ncz = S->prob.grid[2];
  }
  else if (S->inner_zrun)
  {
    a.zrun = 1;
    a.ncx = S->prob.grid[0];
    a.ncy = S->prob.grid[1];
    a.ncz = S->prob.grid[2];
  }
  return a;
}

// Energy partials of the inner operator's z-run launch in precision dtype:
// one per wave of each z-segment's workgroup.
int zrun_partials(const femtc_solver* S, int dtype, int64_t* npart)
{
  ApplyReq a = operator_req(S, dtype, S->tab, S->G, S->detJw, nullptr,
                            nullptr);
  if (S->inner_affine)
    a.op += OP_MASS_AFF;
  a.eo_tab = nullptr;
  a.zrun = 1;
  int shape[2] = {0, 0};
  a.zrun_shape = shape;
  const int st = run(a);
  if (st != FEMTC_OK)
    return st;
  const int64_t nzs = (S->prob.grid[2] + shape[0] - 1) / shape[0];
  *npart = (int64_t)S->prob.grid[0] * S->prob.grid[1] * nzs * shape[1];
  return FEMTC_OK;
}

ApplyReq outer_req(const femtc_solver* S, const void* x, void* y)
{
  if (S->outer_affine)
  {
    ApplyReq a = operator_req(S, FEMTC_F64, S->tabw64, S->cg64, nullptr, x, y);
    a.op += OP_MASS_AFF;
    return a;
  }
  return operator_req(S, FEMTC_F64, S->prob.tables, S->prob.G, S->prob.detJw,
                      x, y);
}

// The inner operator of request a: the fused kernels, or the cell pass on
// the matrix cores into the lattice (the CG update assembles it).
int run_inner(const femtc_solver* S, const ApplyReq& a)
{
  if (S->mfma_split < 0)
    return run(a);
  MfmaReq m;
  m.N = m.Q = S->N;
  m.split = S->mfma_split;
  m.amat = S->mfma_ops;
  m.w = reinterpret_cast<const float*>(
      static_cast<const unsigned char*>(S->mfma_ops) +
      (size_t)mfma_table_matrices(S->mfma_split) * 64 * 16);
  m.x = static_cast<const float*>(a.x);
  m.lattice = static_cast<float*>(S->lattice);
  m.cg = static_cast<const float*>(S->G);
  m.ks = S->prob.op == FEMTC_OP_MASS ? 0.f : 1.f;
  m.km = S->prob.op == FEMTC_OP_STIFFNESS ? 0.f : 1.f;
  m.ncx = S->prob.grid[0];
  m.ncy = S->prob.grid[1];
  m.ncz = S->prob.grid[2];
  m.skip = a.skip;
  m.energy = a.energy;
  m.stream = a.stream;
  return launch_mfma(m);
}

LatticeGrid lattice_grid(const femtc_solver* S)
{
  return LatticeGrid{S->prob.grid[0], S->prob.grid[1], S->prob.grid[2],
                     S->N - 1};
}

// The vector update of a CG iteration in precision V after the operator:
// k_update on the atomically scattered Ap, or k_assemble_update on the
// assembled scatter's cell values.
template <typename V>
int launch_update(femtc_solver* S, hipStream_t s)
{
  if (S->inner_assembled)
  {
    const V* W = (const V*)S->lattice;
    V* r = (V*)S->r;
    const auto* dinv = (const typename Precond<V>::T*)S->dinv;
    const LatticeGrid L = lattice_grid(S);
    const PcgState* st = S->st;
    switch (L.P)
    {
#define FEMTC_ASM_UPDATE(PP)                                                  \
  case PP:                                                                    \
    hipLaunchKernelGGL((k_assemble_update<V, PP>), dim3(NBLK), dim3(ARB), 0, \
                       s, W, r, dinv, L, st, S->part);                        \
    break;
      FEMTC_ASM_UPDATE(1)
      FEMTC_ASM_UPDATE(2)
      FEMTC_ASM_UPDATE(3)
      FEMTC_ASM_UPDATE(4)
      FEMTC_ASM_UPDATE(5)
      FEMTC_ASM_UPDATE(6)
      FEMTC_ASM_UPDATE(7)
      FEMTC_ASM_UPDATE(8)
#undef FEMTC_ASM_UPDATE
    default:
      return FEMTC_ERR_DEGREE_NOT_SUPPORTED;
    }
  }
  else
    hipLaunchKernelGGL(k_update<V>, dim3(NBLK), dim3(RB), 0, s, (V*)S->r,
                       (V*)S->Ap, (const typename Precond<V>::T*)S->dinv,
                       S->n,
                       (const PcgState*)S->st, S->part);
  return hip_status(hipGetLastError());
}

// Mean time in ms of reps operator-and-update steps (after one warm-up).
template <typename V>
int time_step(femtc_solver* S, int reps, float* ms)
{
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess)
    return fail(FEMTC_ERR_HIP, "event creation failed");
  if (hipEventCreate(&e1) != hipSuccess)
  {
    (void)hipEventDestroy(e0);
    return fail(FEMTC_ERR_HIP, "event creation failed");
  }
  const ApplyReq a = inner_req(S, dtype_of<V>(), S->p, S->Ap);
  int st = FEMTC_OK;
  for (int i = 0; i <= reps && st == FEMTC_OK; ++i)
  {
    if (i == 1 && hipEventRecord(e0, S->stream) != hipSuccess)
      st = FEMTC_ERR_HIP;
    if (st == FEMTC_OK)
      st = run_inner(S, a);
    if (st == FEMTC_OK)
      st = launch_update<V>(S, S->stream);
  }
  if (st == FEMTC_OK && (hipEventRecord(e1, S->stream) != hipSuccess ||
                         hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(ms, e0, e1) != hipSuccess))
    st = FEMTC_ERR_HIP;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st == FEMTC_ERR_HIP)
    return fail(st, std::string("solver scatter timing: ") +
                        hipGetErrorString(hipGetLastError()));
  *ms /= reps;
  return st;
}

// Energy partials one operator launch writes: one per wave of every cell
// group (the assembled scatter's groups hold assembled_cells cells).
int64_t energy_partials(int64_t C, int Q, bool assembled, int dtype)
{
  const int64_t Qq = (int64_t)Q * Q;
  const int64_t cpg =
      assembled ? assembled_cells(Q, dtype == FEMTC_F64 ? 8 : 4)
                : direct_cells(Q);
  const int64_t waves = (cpg * Qq + 63) / 64;
  return (C + cpg - 1) / cpg * waves;
}

// Scatter and contraction form of the operators.  The scatters allowed by
// problem.grid/scatter (the assembled and z-run ones only for the
// structured dofmap; FEMTC_SCATTER_AUTO allows all three) and the forms
// available (even-odd when the tables passed the check, direct and
// assembled scatters) are each timed over one CG iteration step on this
// problem, and the fastest pair is kept (vectors zeroed and alpha = 0 for
// the timing; the solve re-initialises everything).
int choose_variant(femtc_solver* S, int structured)
{
  if (S->mfma_split >= 0)
  {
    // the matrix-core cell pass writes the lattice: assembled scatter only
    if (!structured)
      return fail(FEMTC_ERR_INVALID_ARGUMENT,
                  "BF16 inner operator: the dofmap is not the structured "
                  "numbering of the grid");
    if (S->prob.scatter == FEMTC_SCATTER_DIRECT ||
        S->prob.scatter == FEMTC_SCATTER_ZRUN)
      return fail(FEMTC_ERR_INVALID_ARGUMENT,
                  "BF16 inner operator: assembled scatter only");
    S->inner_assembled = true;
    S->even_odd = false;
    S->nepart = S->prob.num_cells; // one energy partial per cell (wave)
    return dev_alloc(S, &S->lattice,
                     4 * (size_t)lattice_size(S->prob.grid[0], S->prob.grid[1],
                                              S->prob.grid[2], S->N - 1));
  }
  const int mode = S->prob.scatter;
  const bool have_eo = S->eo64 && S->eo_in;
  bool try_asm = structured && (mode == FEMTC_SCATTER_AUTO ||
                                mode == FEMTC_SCATTER_ASSEMBLED);
  const bool try_zrun = structured && S->Q <= 8 &&
                        (mode == FEMTC_SCATTER_AUTO ||
                         mode == FEMTC_SCATTER_ZRUN);
  if (try_asm)
  {
    const size_t es = S->inner_dtype == FEMTC_F64 ? 8 : 4;
    if (dev_alloc(S, &S->lattice,
                  es * (size_t)lattice_size(S->prob.grid[0], S->prob.grid[1],
                                            S->prob.grid[2], S->N - 1)))
    {
      if (mode == FEMTC_SCATTER_ASSEMBLED)
        return FEMTC_ERR_OUT_OF_MEMORY;
      try_asm = false; // no room: keep the atomic scatter
    }
  }
  const bool try_direct =
      mode == FEMTC_SCATTER_DIRECT || mode == FEMTC_SCATTER_AUTO ||
      (mode == FEMTC_SCATTER_ASSEMBLED && !try_asm);
  // even-odd forms to try: none, the even-odd one only (FEMTC_EVEN_ODD_ON),
  // or both; the z-run scatter has no even-odd form
  const int e0 = have_eo && S->prob.even_odd == FEMTC_EVEN_ODD_ON ? 1 : 0;
  const int e1 = have_eo ? 2 : 1;
  // (scatter: 0 direct, 1 assembled, 2 z-run; even-odd)
  std::vector<std::pair<int, bool>> cand;
  for (int a = 0; a < 3; ++a)
    if (a == 0 ? try_direct : a == 1 ? try_asm : try_zrun)
    {
      if (a == 2)
      {
        if (e0 == 0)
          cand.push_back({2, false});
        continue;
      }
      for (int e = e0; e < e1; ++e)
        cand.push_back({a, e == 1});
    }
  if (cand.empty())
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "the z-run scatter has no even-odd form");
  size_t best = 0;
  int st = FEMTC_OK;
  if (cand.size() > 1)
  {
    hipStream_t s = S->stream;
    const int dt = S->inner_dtype;
    for (void* v : {S->e, S->r, S->p, S->Ap})
      if (st == FEMTC_OK)
        st = fill_zero(v, S->n, dt, s);
    PcgState h{};
    if (st == FEMTC_OK &&
        (hipMemcpyAsync(S->st, &h, sizeof(h), hipMemcpyHostToDevice, s) !=
             hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
      st = fail(FEMTC_ERR_HIP, "solver: state upload failed");
    float tbest = 0.f;
    for (size_t c = 0; c < cand.size() && st == FEMTC_OK; ++c)
    {
      S->inner_assembled = cand[c].first == 1;
      S->inner_zrun = cand[c].first == 2;
      S->even_odd = cand[c].second;
      float t = 0.f;
      st = dt == FEMTC_F64 ? time_step<double>(S, 3, &t)
                           : time_step<float>(S, 3, &t);
      if (c == 0 || t < tbest)
      {
        tbest = t;
        best = c;
      }
    }
  }
  S->inner_assembled = cand[best].first == 1;
  S->inner_zrun = cand[best].first == 2;
  S->even_odd = cand[best].second;
  S->nepart = energy_partials(S->prob.num_cells, S->Q, S->inner_assembled,
                              S->inner_dtype);
  if (S->inner_zrun && st == FEMTC_OK)
    st = zrun_partials(S, S->inner_dtype, &S->nepart);
  if (st == FEMTC_OK && S->nepart > S->nepart_cap)
  {
    // the z-run launch has a wave per segment slot: more partials
    dev_free(S->epart);
    S->bytes -= 8 * (size_t)S->nepart_cap;
    S->epart = nullptr;
    S->nepart_cap = 0;
    if ((st = dev_alloc(S, (void**)&S->epart, 8 * S->nepart)) == FEMTC_OK)
      S->nepart_cap = S->nepart;
  }
  if (!S->inner_assembled && S->lattice)
  {
    dev_free(S->lattice);
    S->lattice = nullptr;
    S->bytes -= (S->inner_dtype == FEMTC_F64 ? 8 : 4) *
                (size_t)lattice_size(S->prob.grid[0], S->prob.grid[1],
                                     S->prob.grid[2], S->N - 1);
  }
  return st;
}

int sync_read(femtc_solver* S, void* host, const void* dev, size_t bytes)
{
  if (hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, S->stream) !=
          hipSuccess ||
      hipStreamSynchronize(S->stream) != hipSuccess)
    return fail(FEMTC_ERR_HIP, std::string("solver: ") +
                                   hipGetErrorString(hipGetLastError()));
  return FEMTC_OK;
}

// Host to device on the solver's stream, complete on return.  Solver
// creation makes no call on the legacy default stream, so that a server
// can create one solver on a slot's thread while another slot's thread
// captures a second solver's CG batch into a graph.
int sync_write(femtc_solver* S, void* dev, const void* host, size_t bytes)
{
  if (hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, S->stream) !=
          hipSuccess ||
      hipStreamSynchronize(S->stream) != hipSuccess)
    return fail(FEMTC_ERR_HIP, std::string("solver: ") +
                                   hipGetErrorString(hipGetLastError()));
  return FEMTC_OK;
}

#define LAUNCH_CHECK()                                                        \
  do                                                                          \
  {                                                                           \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess)                                                     \
      return fail(FEMTC_ERR_HIP,                                              \
                  std::string("solver kernel launch: ") +                     \
                      hipGetErrorString(e_));                                 \
  } while (0)

// check_every CG iterations in precision V, queued on stream s: operator
// (with the p.Ap energies), alpha, update, beta, p-update.  Every kernel
// returns at once after the state's `done` is set.
template <typename V>
int enqueue_batch(femtc_solver* S, hipStream_t s)
{
  ApplyReq a = inner_req(S, dtype_of<V>(), S->p, S->Ap);
  a.skip = &S->st->done;
  a.energy = S->epart; // p.Ap from the operator kernel itself
  a.stream = s;
  for (int k = 0; k < S->check_every; ++k)
  {
    int st = run_inner(S, a);
    if (st != FEMTC_OK)
      return st;
    hipLaunchKernelGGL(k_energy_partials, dim3(NBLK), dim3(RB), 0, s,
                       (const double*)S->epart, S->nepart, S->part,
                       (const PcgState*)S->st);
    hipLaunchKernelGGL(k_alpha, dim3(1), dim3(RB), 0, s,
                       (const double*)S->part, S->st);
    if ((st = launch_update<V>(S, s)) != FEMTC_OK)
      return st;
    hipLaunchKernelGGL(k_beta, dim3(1), dim3(RB), 0, s,
                       (const double*)S->part, S->st);
    hipLaunchKernelGGL(k_pupdate<V>, dim3(NBLK), dim3(RB), 0, s, (V*)S->p,
                       (V*)S->e, (const V*)S->r,
                       (const typename Precond<V>::T*)S->dinv, S->n,
                       (const PcgState*)S->st);
  }
  return hip_status(hipGetLastError());
}

// One batch on the solver's stream: replayed from a hipGraph captured at
// the first call (on a private stream of default priority; the graph holds
// only kernel nodes), or queued kernel by kernel when graphs are off or
// capture failed.
template <typename V>
int launch_batch(femtc_solver* S)
{
  if (S->use_graph && !S->graph && !S->graph_failed)
  {
    hipGraph_t g = nullptr;
    bool ok = S->cap_stream ||
              hipStreamCreateWithFlags(&S->cap_stream, hipStreamNonBlocking) ==
                  hipSuccess;
    if (ok)
      ok = hipStreamBeginCapture(S->cap_stream,
                                 hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok)
    {
      const int st = enqueue_batch<V>(S, S->cap_stream);
      ok = hipStreamEndCapture(S->cap_stream, &g) == hipSuccess &&
           st == FEMTC_OK && g &&
           hipGraphInstantiate(&S->graph, g, nullptr, nullptr, 0) ==
               hipSuccess;
    }
    if (g)
      (void)hipGraphDestroy(g);
    if (!ok)
    {
      S->graph = nullptr;
      S->graph_failed = true;
      (void)hipGetLastError();
    }
  }
  if (S->graph)
    return hipGraphLaunch(S->graph, S->stream) == hipSuccess
               ? FEMTC_OK
               : fail(FEMTC_ERR_HIP, std::string("solver graph launch: ") +
                                         hipGetErrorString(hipGetLastError()));
  const int st = enqueue_batch<V>(S, S->stream);
  if (st == FEMTC_ERR_HIP)
    return fail(st, std::string("solver kernel launch: ") +
                        hipGetErrorString(hipGetLastError()));
  return st;
}

// PCG in precision V on A e = rhs (inner buffers of S), from e = 0.
template <typename V>
int pcg_run(femtc_solver* S, double rtol, int maxit, int* iters, int* status)
{
  hipStream_t s = S->stream;
  const int64_t n = S->n;
  hipLaunchKernelGGL(k_start<V>, dim3(NBLK), dim3(RB), 0, s, (const V*)S->rhs,
                     (V*)S->e, (V*)S->r, (V*)S->p, (V*)S->Ap,
                     (const typename Precond<V>::T*)S->dinv, n, S->part);
  hipLaunchKernelGGL(k_reduce_start, dim3(1), dim3(RB), 0, s,
                     (const double*)S->part, S->st, rtol, maxit);
  LAUNCH_CHECK();
  for (;;)
  {
    int st = launch_batch<V>(S);
    if (st != FEMTC_OK)
      return st;
    st = sync_read(S, S->hst, S->st, sizeof(PcgState));
    if (st != FEMTC_OK)
      return st;
    if (S->hst->done)
      break;
  }
  *iters = S->hst->it;
  *status = S->hst->status;
  return FEMTC_OK;
}

// ||f - A u|| in FP64; leaves the residual in S->r64.
int true_residual(femtc_solver* S, const double* f, const double* u,
                  double* rnorm)
{
  hipStream_t s = S->stream;
  int st = fill_zero(S->Au64, S->n, FEMTC_F64, s);
  if (st != FEMTC_OK)
    return st;
  ApplyReq a = outer_req(S, u, S->Au64);
  st = run(a);
  if (st != FEMTC_OK)
    return st;
  hipLaunchKernelGGL(k_residual, dim3(NBLK), dim3(RB), 0, s, f,
                     (const double*)S->Au64, S->r64, S->n, S->part);
  hipLaunchKernelGGL(k_combine1, dim3(1), dim3(RB), 0, s,
                     (const double*)S->part, S->dscal);
  LAUNCH_CHECK();
  st = sync_read(S, S->hscal, S->dscal, sizeof(double));
  if (st != FEMTC_OK)
    return st;
  *rnorm = std::sqrt(*S->hscal);
  return FEMTC_OK;
}

int norm2(femtc_solver* S, const double* x, double* out)
{
  hipLaunchKernelGGL(k_norm2, dim3(NBLK), dim3(RB), 0, S->stream, x, S->n,
                     S->part);
  hipLaunchKernelGGL(k_combine1, dim3(1), dim3(RB), 0, S->stream,
                     (const double*)S->part, S->dscal);
  LAUNCH_CHECK();
  int st = sync_read(S, S->hscal, S->dscal, sizeof(double));
  if (st != FEMTC_OK)
    return st;
  *out = std::sqrt(*S->hscal);
  return FEMTC_OK;
}

int check_problem(const femtc_problem* P)
{
  if (!P)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "problem is NULL");
  if (P->op < FEMTC_OP_MASS || P->op > FEMTC_OP_KPM)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "op must be a femtc_op");
  int st = check_shape_fused(P->degree, P->nq, FEMTC_MODE_CORRECT);
  if (st != FEMTC_OK)
    return st;
  if (P->num_cells < 0 || P->ndofs < 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "negative size");
  if (!P->tables || (P->num_cells > 0 && !P->dofmap))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "tables and dofmap are required");
  if (P->op != FEMTC_OP_MASS && !P->G && !P->J)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "G (or the cell Jacobians J) is required for this op");
  if (P->op != FEMTC_OP_STIFFNESS && !P->detJw && !P->J)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "detJw (or the cell Jacobians J) is required for this op");
  const bool nogrid = P->grid[0] == 0 && P->grid[1] == 0 && P->grid[2] == 0;
  if (!nogrid && (P->grid[0] < 1 || P->grid[1] < 1 || P->grid[2] < 1 ||
                  (int64_t)P->grid[0] * P->grid[1] * P->grid[2] !=
                      P->num_cells))
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "grid must be all zero or positive with grid[0]*grid[1]*grid[2] "
                "== num_cells");
  if (P->scatter < FEMTC_SCATTER_AUTO || P->scatter > FEMTC_SCATTER_ZRUN)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "scatter must be a femtc_scatter");
  if ((P->scatter == FEMTC_SCATTER_ASSEMBLED ||
       P->scatter == FEMTC_SCATTER_ZRUN) &&
      nogrid)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "the assembled and z-run scatters need the structured grid");
  if (P->scatter == FEMTC_SCATTER_ZRUN && P->nq > 8)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "the z-run scatter needs nq <= 8");
  if (P->even_odd < FEMTC_EVEN_ODD_AUTO || P->even_odd > FEMTC_EVEN_ODD_ON)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "even_odd must be FEMTC_EVEN_ODD_AUTO, _OFF or _ON");
  return FEMTC_OK;
}

} // namespace
} // namespace femtc

extern "C" {

int femtc_global_diagonal(int op, int dtype, int degree, int nq,
                          const void* tables, double* diag, int64_t ndofs,
                          const int32_t* dofmap, const void* G,
                          const void* detJw, int64_t num_cells, void* stream)
{
  int st = check_shape_fused(degree, nq, FEMTC_MODE_CORRECT);
  if (st != FEMTC_OK)
    return st;
  if ((st = check_dtype(dtype)) != FEMTC_OK)
    return st;
  if (op < FEMTC_OP_MASS || op > FEMTC_OP_KPM)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "op must be a femtc_op");
  if (!diag || !tables || (num_cells > 0 && !dofmap) || ndofs < 0 ||
      num_cells < 0)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "missing buffer or bad size");
  if (op != FEMTC_OP_MASS && !G)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "G is required for this op");
  if (op != FEMTC_OP_STIFFNESS && !detJw)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "detJw is required for this op");
  st = launch_diag(op, dtype, degree + 1, nq, tables, G, detJw, dofmap,
                   num_cells, diag, ndofs, (hipStream_t)stream);
  if (st == FEMTC_ERR_HIP)
    return fail(st, std::string("diagonal kernel: ") +
                        hipGetErrorString(hipGetLastError()));
  return st;
}

int femtc_solver_create(femtc_solver** out, const femtc_problem* prob,
                        int inner_dtype, void* stream)
{
  if (!out)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  int st = check_problem(prob);
  if (st != FEMTC_OK)
    return st;
  const bool bf = inner_dtype == FEMTC_BF16 || inner_dtype == FEMTC_BF16_ROUNDED;
  if (inner_dtype != FEMTC_F64 && inner_dtype != FEMTC_F32 && !bf)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "inner_dtype must be FEMTC_F64 (plain PCG), FEMTC_F32, "
                "FEMTC_BF16 or FEMTC_BF16_ROUNDED (inner PCG with FP64 defect "
                "correction)");
  if (bf && (!prob->J || prob->grid[0] < 1 || prob->nq != prob->degree + 1 ||
             prob->degree > 7))
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "BF16 inner operator: needs the cell Jacobians J, the "
                "structured grid, nq = degree + 1 and degree <= 7");
  femtc_solver* S = new femtc_solver;
  S->prob = *prob;
  S->inner_dtype = inner_dtype;
  S->mfma_split = bf ? (inner_dtype == FEMTC_BF16 ? 1 : 0) : -1;
  S->N = prob->degree + 1;
  S->Q = prob->nq;
  S->n = prob->ndofs;
  S->stream = (hipStream_t)stream;
  const int64_t n = S->n;
  const int64_t C = prob->num_cells;
  const int64_t Q3 = (int64_t)S->Q * S->Q * S->Q;
  const size_t vs = inner_dtype == FEMTC_F64 ? 8 : 4;
  hipStream_t s = S->stream;
  do
  {
    if ((st = dev_alloc(S, (void**)&S->diag, 8 * n)) ||
        (st = dev_alloc(S, (void**)&S->dinv64, 8 * n)) ||
        (st = dev_alloc(S, (void**)&S->r64, 8 * n)) ||
        (st = dev_alloc(S, (void**)&S->Au64, 8 * n)) ||
        (st = dev_alloc(S, (void**)&S->part, 8 * 2 * NBLK)) ||
        (st = dev_alloc(S, (void**)&S->dscal, 8)) ||
        (st = dev_alloc(S, (void**)&S->dbad, sizeof(int))) ||
        (st = dev_alloc(S, (void**)&S->st, sizeof(PcgState))) ||
        (st = dev_alloc(S, &S->rhs, vs * n)) || (st = dev_alloc(S, &S->e, vs * n)) ||
        (st = dev_alloc(S, &S->p, vs * n)) || (st = dev_alloc(S, &S->Ap, vs * n)))
      break;
    // one energy partial per wave of the operator launch, whose cell
    // groups depend on the scatter (set again once it is chosen)
    S->nepart = bf ? C
                   : std::max(energy_partials(C, S->Q, false, inner_dtype),
                              energy_partials(C, S->Q, true, inner_dtype));
    if ((st = dev_alloc(S, (void**)&S->epart, 8 * S->nepart)))
      break;
    S->nepart_cap = S->nepart;
    if (hipHostMalloc((void**)&S->hst, sizeof(PcgState)) != hipSuccess ||
        hipHostMalloc((void**)&S->hscal, sizeof(double)) != hipSuccess)
    {
      st = fail(FEMTC_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
      break;
    }
    // affine cells: [phi | dphi | w] and the 7 per-cell values in FP64
    const bool aff = prob->J != nullptr;
    S->inner_affine = aff;
    S->outer_affine =
        aff && ((prob->op != FEMTC_OP_MASS && !prob->G) ||
                (prob->op != FEMTC_OP_STIFFNESS && !prob->detJw));
    const int64_t ntab = 2 * (int64_t)S->N * S->Q;
    if (aff)
    {
      if ((st = dev_alloc(S, (void**)&S->tabw64, 8 * (ntab + S->Q))) ||
          (st = dev_alloc(S, (void**)&S->cg64, 8 * CELLGEO * C)))
        break;
      std::vector<double> xq(S->Q), wq(S->Q);
      femtc_gauss_rule(S->Q, xq.data(), wq.data());
      if (hipMemcpyAsync(S->tabw64, prob->tables, 8 * ntab,
                         hipMemcpyDeviceToDevice, s) != hipSuccess)
      {
        st = fail(FEMTC_ERR_HIP, "solver: table copy failed");
        break;
      }
      if ((st = sync_write(S, S->tabw64 + ntab, wq.data(), 8 * S->Q)))
        break;
      if ((st = femtc_geometry_cell(C, prob->J, S->cg64, FEMTC_F64, s)))
        break;
    }
    // every dofmap entry in range before the first kernel scatters
    // through it
    {
      int inrange = 0;
      if ((st = check_dofmap_range(prob->dofmap,
                                   C * S->N * S->N * S->N, n, s,
                                   &inrange)) != FEMTC_OK)
        break;
      if (!inrange)
      {
        st = fail(FEMTC_ERR_INVALID_ARGUMENT,
                  "a dofmap entry is outside [0, ndofs)");
        break;
      }
    }
    // Jacobi preconditioner from the FP64 operator
    if (S->outer_affine)
      st = launch_diag(prob->op, FEMTC_F64, S->N, S->Q, S->tabw64, nullptr,
                       nullptr, prob->dofmap, C, S->diag, n, s, S->cg64);
    else
      st = launch_diag(prob->op, FEMTC_F64, S->N, S->Q, prob->tables,
                       prob->G, prob->detJw, prob->dofmap, C, S->diag, n, s);
    if (st != FEMTC_OK)
      break;
    if ((st = fill_zero(S->dbad, 1, FEMTC_F32, s)) != FEMTC_OK)
      break;
    hipLaunchKernelGGL(k_invert, dim3(NBLK), dim3(RB), 0, s,
                       (const double*)S->diag, S->dinv64, n, S->dbad);
    if (inner_dtype == FEMTC_F64)
    {
      S->r = S->r64;
      S->dinv = S->dinv64;
      S->tab = aff ? (void*)S->tabw64 : (void*)prob->tables;
      S->G = aff ? (void*)S->cg64 : (void*)prob->G;
      S->detJw = aff ? nullptr : (void*)prob->detJw;
    }
    else if (aff)
    {
      // FP32 inner operator on affine cells: 7 values per cell
      if ((st = dev_alloc(S, &S->r, 4 * n)) || (st = dev_alloc(S, &S->dinv, 2 * n)) ||
          (st = dev_alloc(S, &S->tab, 4 * (ntab + S->Q))) ||
          (st = dev_alloc(S, &S->G, 4 * CELLGEO * C)))
        break;
      hipLaunchKernelGGL(k_precond_bf16, dim3(NBLK), dim3(RB), 0, s,
                         (const double*)S->dinv64, (uint16_t*)S->dinv, n);
      hipLaunchKernelGGL((k_convert<double, float>), dim3(NBLK), dim3(RB), 0,
                         s, (const double*)S->tabw64, (float*)S->tab,
                         ntab + S->Q);
      hipLaunchKernelGGL((k_convert<double, float>), dim3(NBLK), dim3(RB), 0,
                         s, (const double*)S->cg64, (float*)S->G,
                         (int64_t)CELLGEO * C);
    }
    else
    {
      if ((st = dev_alloc(S, &S->r, 4 * n)) || (st = dev_alloc(S, &S->dinv, 2 * n)) ||
          (st = dev_alloc(S, &S->tab, 4 * 2 * S->N * S->Q)))
        break;
      hipLaunchKernelGGL(k_precond_bf16, dim3(NBLK), dim3(RB), 0, s,
                         (const double*)S->dinv64, (uint16_t*)S->dinv, n);
      hipLaunchKernelGGL((k_convert<double, float>), dim3(NBLK), dim3(RB), 0,
                         s, prob->tables, (float*)S->tab,
                         (int64_t)2 * S->N * S->Q);
      if (prob->G)
      {
        if ((st = dev_alloc(S, &S->G, 4 * 6 * Q3 * C)))
          break;
        hipLaunchKernelGGL((k_convert<double, float>), dim3(NBLK), dim3(RB),
                           0, s, prob->G, (float*)S->G, 6 * Q3 * C);
      }
      if (prob->detJw)
      {
        if ((st = dev_alloc(S, &S->detJw, 4 * Q3 * C)))
          break;
        hipLaunchKernelGGL((k_convert<double, float>), dim3(NBLK), dim3(RB),
                           0, s, prob->detJw, (float*)S->detJw, Q3 * C);
      }
    }
    if (bf)
    {
      // the matrix-core operand set, from the FP64 tables read back once
      std::vector<double> th(ntab);
      size_t ob = 0;
      if ((st = femtc_mfma_tables_bytes(S->mfma_split, &ob)))
        break;
      std::vector<unsigned char> oh(ob);
      if ((st = sync_read(S, th.data(), prob->tables, 8 * ntab)) ||
          (st = dev_alloc(S, &S->mfma_ops, ob)) ||
          (st = femtc_mfma_tables_host(prob->degree, prob->nq,
                                       S->mfma_split, th.data(),
                                       oh.data())) ||
          (st = sync_write(S, S->mfma_ops, oh.data(), ob)))
        break;
    }
    hipError_t he = hipGetLastError();
    if (he != hipSuccess)
    {
      st = fail(FEMTC_ERR_HIP, std::string("solver setup: ") +
                                   hipGetErrorString(he));
      break;
    }
    int bad = 0;
    if ((st = sync_read(S, &bad, S->dbad, sizeof(int))) != FEMTC_OK)
      break;
    if (bad)
    {
      st = fail(FEMTC_ERR_INVALID_ARGUMENT,
                "the operator diagonal has a non-positive entry (a dof that "
                "no cell touches, or geometry that is not positive definite)");
      break;
    }
    // even-odd table sets when the tables are symmetric (the weights of the
    // Gauss rule ride along for the per-cell geometry operators); whether
    // they are used is timed below with the scatter
    if (prob->even_odd != FEMTC_EVEN_ODD_OFF && S->Q == S->N && !bf)
    {
      std::vector<double> xq(S->Q), wq(S->Q);
      femtc_gauss_rule(S->Q, xq.data(), wq.data());
      if ((st = make_even_odd(prob->tables, FEMTC_F64, S->N, S->Q, wq.data(),
                              FEMTC_F64, s, &S->eo64)) != FEMTC_OK)
        break;
      S->eo_in = S->eo64;
      if (S->eo64 && inner_dtype != FEMTC_F64 &&
          (st = make_even_odd(prob->tables, FEMTC_F64, S->N, S->Q, wq.data(),
                              inner_dtype, s, &S->eo_in)) != FEMTC_OK)
        break;
    }
    // scatter and contraction form of the operators
    int structured = 0;
    if (prob->grid[0] > 0 &&
        (st = verify_structured(prob->dofmap, prob->grid[0], prob->grid[1],
                                prob->grid[2], prob->degree, n, s,
                                &structured)) != FEMTC_OK)
      break;
    if ((prob->scatter == FEMTC_SCATTER_ASSEMBLED ||
         prob->scatter == FEMTC_SCATTER_ZRUN) &&
        !structured)
    {
      st = fail(FEMTC_ERR_INVALID_ARGUMENT,
                "assembled or z-run scatter: the dofmap is not the "
                "structured numbering of the grid");
      break;
    }
    if ((st = choose_variant(S, structured)) != FEMTC_OK)
      break;
  } while (false);
  if (st != FEMTC_OK)
  {
    destroy(S);
    return st;
  }
  *out = S;
  return FEMTC_OK;
}

int femtc_solver_destroy(femtc_solver* solver)
{
  destroy(solver);
  return FEMTC_OK;
}

int femtc_solver_bytes(const femtc_solver* solver, size_t* bytes)
{
  if (!solver || !bytes)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "NULL argument");
  *bytes = solver->bytes;
  return FEMTC_OK;
}

int femtc_solver_diagonal(const femtc_solver* solver, const double** diag)
{
  if (!solver || !diag)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "NULL argument");
  *diag = solver->diag;
  return FEMTC_OK;
}

int femtc_solver_set_batch(femtc_solver* solver, int check_every,
                           int use_graph)
{
  if (!solver)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (check_every < 1 || check_every > 1024)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "check_every must be in [1, 1024]");
  if (solver->graph &&
      (check_every != solver->check_every || !use_graph))
  {
    if (hipStreamSynchronize(solver->stream) != hipSuccess)
      return fail(FEMTC_ERR_HIP, "solver: stream synchronisation failed");
    (void)hipGraphExecDestroy(solver->graph);
    solver->graph = nullptr;
  }
  solver->check_every = check_every;
  solver->use_graph = use_graph != 0;
  solver->graph_failed = false;
  return FEMTC_OK;
}

int femtc_solver_scatter(const femtc_solver* solver, int* scatter)
{
  if (!solver || !scatter)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "NULL argument");
  *scatter = solver->inner_assembled ? FEMTC_SCATTER_ASSEMBLED
             : solver->inner_zrun       ? FEMTC_SCATTER_ZRUN
                                        : FEMTC_SCATTER_DIRECT;
  return FEMTC_OK;
}

int femtc_solver_even_odd(const femtc_solver* solver, int* even_odd)
{
  if (!solver || !even_odd)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "NULL argument");
  *even_odd = solver->even_odd ? 1 : 0;
  return FEMTC_OK;
}

int femtc_solver_solve(femtc_solver* S, const double* f, double* u,
                       double rtol, double inner_rtol, int maxit,
                       femtc_solve_info* info)
{
  if (!S || !f || !u)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!(rtol > 0.0) || maxit < 1)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "rtol must be > 0, maxit >= 1");
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = S->stream;
  const int64_t n = S->n;
  femtc_solve_info inf{};
  inf.status = FEMTC_OK;
  double fn = 0.0, rn_final = 0.0;
  int st = norm2(S, f, &fn);
  if (st != FEMTC_OK)
    return st;
  if (S->inner_dtype == FEMTC_F64)
  {
    // plain PCG in FP64: rhs = f, solution into u
    if (hipMemcpyAsync(S->rhs, f, 8 * n, hipMemcpyDeviceToDevice, s) !=
        hipSuccess)
      return fail(FEMTC_ERR_HIP, "copy of f failed");
    int it = 0, pst = FEMTC_OK;
    st = pcg_run<double>(S, rtol, maxit, &it, &pst);
    if (st != FEMTC_OK)
      return st;
    if (hipMemcpyAsync(u, S->e, 8 * n, hipMemcpyDeviceToDevice, s) !=
        hipSuccess)
      return fail(FEMTC_ERR_HIP, "copy of u failed");
    inf.inner_iterations = it;
    inf.outer_iterations = 1;
    inf.status = pst;
    // report the true FP64 residual of the returned u
    if ((st = true_residual(S, f, u, &rn_final)) != FEMTC_OK)
      return st;
  }
  else
  {
    if (!(inner_rtol > 0.0 && inner_rtol < 1.0))
      return fail(FEMTC_ERR_INVALID_ARGUMENT, "inner_rtol must be in (0, 1)");
    if ((st = fill_zero(u, n, FEMTC_F64, s)) != FEMTC_OK)
      return st;
    double rn = fn; // r = f for u = 0
    rn_final = rn;
    if (hipMemcpyAsync(S->r64, f, 8 * n, hipMemcpyDeviceToDevice, s) !=
        hipSuccess)
      return fail(FEMTC_ERR_HIP, "copy of f failed");
    const int max_outer = 60;
    int stall = 0;
    while (fn > 0.0 && rn > rtol * fn)
    {
      if (inf.outer_iterations >= max_outer || inf.inner_iterations >= maxit)
      {
        inf.status = FEMTC_ERR_NOT_CONVERGED;
        break;
      }
      // inner: A_32 e = r / ||r||
      hipLaunchKernelGGL(k_scale_to<float>, dim3(NBLK), dim3(RB), 0, s,
                         (const double*)S->r64, 1.0 / rn, (float*)S->rhs, n);
      LAUNCH_CHECK();
      // the last step only has to close the remaining gap to rtol: solve
      // it no tighter than needed (with a safety factor of 3)
      const double tau = std::max(inner_rtol, 0.3 * rtol * fn / rn);
      int it = 0, pst = FEMTC_OK;
      st = pcg_run<float>(S, std::min(tau, 0.5), maxit - inf.inner_iterations,
                          &it, &pst);
      if (st != FEMTC_OK)
        return st;
      inf.inner_iterations += it;
      inf.outer_iterations += 1;
      if (pst == FEMTC_ERR_BREAKDOWN)
      {
        inf.status = pst;
        break;
      }
      hipLaunchKernelGGL(k_correct<float>, dim3(NBLK), dim3(RB), 0, s, u,
                         (const float*)S->e, rn, n);
      LAUNCH_CHECK();
      double rn_new = 0.0;
      if ((st = true_residual(S, f, u, &rn_new)) != FEMTC_OK)
        return st;
      stall = rn_new > 0.5 * rn ? stall + 1 : 0;
      rn = rn_new;
      rn_final = rn;
      if (stall >= 3)
      {
        inf.status = FEMTC_ERR_NOT_CONVERGED; // refinement stagnates
        break;
      }
    }
  }
  inf.relres = fn > 0.0 ? rn_final / fn : 0.0;
  inf.seconds = std::chrono::duration<double>(
                    std::chrono::steady_clock::now() - t0)
                    .count();
  if (info)
    *info = inf;
  if (inf.status != FEMTC_OK)
    return fail(inf.status,
                "solve stopped: " +
                    std::string(femtc_status_string(inf.status)) +
                    ", relative residual " + std::to_string(inf.relres));
  return FEMTC_OK;
}

} // extern "C"
