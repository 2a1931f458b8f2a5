// General-degree cell operators: N = 10..26 dofs per direction (degrees
// 9..25, the top of the reference's dispatch ladder, src/mass.hpp:52-138 and
// src/stiffness.hpp:102-188), Q in {N, N+1, N+2}, FP64 and FP32, cell-local
// or gathered through a dofmap with an atomic scatter.  Mass, stiffness and
// K+M in correct mode (stored or affine geometry) go to the line kernels
// (apply_line.hpp) where one exists for (dtype, op, N, Q) -- N = 10..22 in
// FP64 stiffness, up to 26 for the rest (line_fits) -- unless
// femtc_set_high_degree_kernel asks for this one; the reference modes and
// the shapes without a line kernel run here.
//
// The fused kernels hold a whole cell per wave in registers and LDS, which
// stops at N = 9.  Here one workgroup (256 to 768 threads) takes one cell at a time
// and walks it one quadrature plane qa at a time, so LDS holds only a few
// planes (at most 60 KB at N = 26, Q = 28 in FP64) and each thread keeps
// the output column it owns in registers:
//
//   F_a  columns bc = b N + c    A0 = phi_a U,  A1 = g_a U   (U from global)
//   F_b  entries (qb, c)         B0 = phi_b A0, B1 = g_b A0, B2 = phi_b A1
//   F_c  entries (qb, qc)        u = phi_c B0,  d_c = g_c B0,
//                                d_b = phi_c B1, d_a = phi_c B2
//        pointwise               (v_a, v_b, v_c) = S(q) (d_a, d_b, d_c),
//                                v_m = detJw u
//   B_c  entries (qb, c')        D1 = phi_c' v_a, D2 = phi_c' v_b,
//                                D3 = g_c' v_c + phi_c' v_m
//   B_b  columns b'c'            E1 = phi_b' D1, E2 = g_b' D2 + phi_b' D3
//   B_a  registers               w[a'] += g_a'(qa) E1 + phi_a'(qa) E2
//
// which is W = sum_ij B_i^T S_ij B_j U + B^T detJw B U, as in the fused
// kernels (the contraction order differs, which changes only rounding).
// Forward tables are read [dof][q]; backward tables are their transpose in
// correct mode and the table as stored in reference mode (Q == N), where
// the literal stiffness keeps v_c = G2 d_a + G1 d_b + G0 u only
// (src/stiffness.hpp:27,68-97) and the literal mass is the correct chain
// with the aliased table (src/mass.hpp:20-21,36-47).
#include <algorithm>
#include <atomic>

#include "apply_line.hpp"
#include "common.hpp"

namespace femtc
{
namespace
{

constexpr int GEN_NMAX = 26; // dofs per direction
// Workgroup shapes: GT threads, each owning one output column of up to NMAX
// values in registers: (256, 16) for N <= 16, (512, 22) for N <= 22,
// (768, 26) above.  One column per thread keeps the accumulator small, so
// more waves stay resident.
static_assert(256 >= 16 * 16 && 512 >= 22 * 22 && 768 >= GEN_NMAX * GEN_NMAX,
              "one column per thread");

__device__ inline void gen_atomic_add(double* p, double v)
{
  unsafeAtomicAdd(p, v);
}
__device__ inline void gen_atomic_add(float* p, float v)
{
  unsafeAtomicAdd(p, v);
}

// LDS elements of one workgroup; the layout is the one gen_cells sets up.
inline size_t gen_lds_elems(int N, int Q, bool ref, bool aff)
{
  const size_t NQ = (size_t)N * Q;
  const size_t X = std::max<size_t>(2 * (size_t)N * N, 4 * (size_t)Q * Q);
  return (ref ? 4 : 2) * NQ + (aff ? Q : 0) + X + 3 * NQ;
}

template <typename T, int OP, bool REF, bool GATHER, int KMAX, int NMAX,
          int GT>
__global__ void __launch_bounds__(GT)
    gen_cells(const T* __restrict__ tab, const T* __restrict__ x,
              T* __restrict__ y, const T* __restrict__ G,
              const T* __restrict__ detJw, const int32_t* __restrict__ dofmap,
              int64_t C, int N, int Q)
{
  constexpr bool AFF = OP >= OP_MASS_AFF;
  constexpr int OPB = AFF ? OP - OP_MASS_AFF : OP;
  constexpr bool LIT = OPB == OP_STIFF_LIT;
  constexpr bool FULL = OPB == OP_STIFF || OPB == OP_KPM;
  constexpr bool MASSTERM = OPB == OP_MASS || OPB == OP_KPM;
  constexpr bool GRAD = FULL || LIT;
  static_assert(!LIT || REF, "the literal stiffness is a reference mode");
  static_assert(!AFF || !REF, "affine geometry runs in correct mode");

  extern __shared__ __align__(16) unsigned char gen_smem[];
  T* sm = (T*)gen_smem;
  const int NQ = N * Q, N2 = N * N, Q2 = Q * Q;
  const int Q3 = Q2 * Q, N3 = N2 * N;
  T* F0 = sm;           // phi [d][q]
  T* F1 = F0 + NQ;      // g   [d][q]
  T* R0 = F0;           // backward tables [d][q]
  T* R1 = F1;
  T* end = F1 + NQ;
  if constexpr (REF)
  {
    R0 = end;
    R1 = R0 + NQ;
    end = R1 + NQ;
  }
  T* wts = end; // quadrature weights (affine geometry)
  if constexpr (AFF)
    end += Q;
  T* X = end; // A0 A1 (F_a -> F_b), then v_a v_b v_c v_m (F_c -> B_c)
  T* Y = X + (2 * N2 > 4 * Q2 ? 2 * N2 : 4 * Q2); // B0..B2, then D1..D3
  T* A0 = X;
  T* A1 = X + N2;
  T* Ca = X;
  T* Cb = X + Q2;
  T* Cc = X + 2 * Q2;
  T* Cm = X + 3 * Q2;
  T* B0 = Y;
  T* B1 = Y + NQ;
  T* B2 = Y + 2 * NQ;
  T* D1 = Y;
  T* D2 = Y + NQ;
  T* D3 = Y + 2 * NQ;

  const int tid = threadIdx.x;
  for (int i = tid; i < NQ; i += GT)
  {
    F0[i] = tab[i];
    if constexpr (GRAD)
      F1[i] = tab[NQ + i];
    if constexpr (REF)
    {
      // the table as stored (Q == N): R[d][q] = T[q][d]
      const int d = i / Q, q = i - d * Q;
      R0[i] = tab[q * N + d];
      if constexpr (GRAD)
        R1[i] = tab[N2 + q * N + d];
    }
  }
  if constexpr (AFF)
    for (int i = tid; i < Q; i += GT)
      wts[i] = tab[2 * NQ + i];
  __syncthreads();

  for (int64_t cell = blockIdx.x; cell < C; cell += gridDim.x)
  {
    T w[KMAX][NMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
#pragma unroll
      for (int a = 0; a < NMAX; ++a)
        w[k][a] = T(0);
    const int64_t cb = cell * (int64_t)N3;
    T cg[AFF ? CELLGEO : 1];
    if constexpr (AFF)
    {
#pragma unroll
      for (int k = 0; k < CELLGEO; ++k)
        cg[k] = G[cell * CELLGEO + k];
    }

    for (int qa = 0; qa < Q; ++qa)
    {
      // F_a: thread per column, U streamed from global (L2-resident)
#pragma unroll
      for (int k = 0; k < KMAX; ++k)
      {
        const int col = tid + k * GT;
        if (col < N2)
        {
          T s0 = T(0), s1 = T(0);
#pragma unroll 2
          for (int a = 0; a < N; ++a)
          {
            const int64_t li = cb + (int64_t)a * N2 + col;
            const T u = GATHER ? x[dofmap[li]] : x[li];
            s0 += F0[a * Q + qa] * u;
            if constexpr (GRAD)
              s1 += F1[a * Q + qa] * u;
          }
          A0[col] = s0;
          if constexpr (GRAD)
            A1[col] = s1;
        }
      }
      __syncthreads();
      // F_b: entries (qb, c)
      for (int e = tid; e < NQ; e += GT)
      {
        const int qb = e / N, c = e - qb * N;
        T s0 = T(0), s1 = T(0), s2 = T(0);
        for (int b = 0; b < N; ++b)
        {
          const T p = F0[b * Q + qb];
          const T a0 = A0[b * N + c];
          s0 += p * a0;
          if constexpr (GRAD)
          {
            s1 += F1[b * Q + qb] * a0;
            s2 += p * A1[b * N + c];
          }
        }
        B0[e] = s0;
        if constexpr (GRAD)
        {
          B1[e] = s1;
          B2[e] = s2;
        }
      }
      __syncthreads();
      // F_c and the pointwise stage: entries (qb, qc)
      for (int e = tid; e < Q2; e += GT)
      {
        const int qb = e / Q, qc = e - qb * Q;
        T u = T(0), dc = T(0), db = T(0), da = T(0);
        for (int c = 0; c < N; ++c)
        {
          const T p = F0[c * Q + qc];
          const T b0 = B0[qb * N + c];
          u += p * b0;
          if constexpr (FULL)
            dc += F1[c * Q + qc] * b0;
          if constexpr (GRAD)
          {
            db += p * B1[qb * N + c];
            da += p * B2[qb * N + c];
          }
        }
        const int64_t q = cell * (int64_t)Q3 + (int64_t)(qa * Q + qb) * Q + qc;
        T g[6];
        T dw = T(0);
        if constexpr (AFF)
        {
          const T wq = wts[qa] * wts[qb] * wts[qc];
#pragma unroll
          for (int k = 0; k < 6; ++k)
            g[k] = wq * cg[k];
          dw = wq * cg[6];
        }
        else
        {
          if constexpr (GRAD)
          {
#pragma unroll
            for (int k = 0; k < 6; ++k)
              g[k] = G[q * 6 + k];
          }
          if constexpr (MASSTERM)
            dw = detJw[q];
        }
        if constexpr (FULL)
        {
          // S = [[G5,G4,G2],[G4,G3,G1],[G2,G1,G0]] on (d_a, d_b, d_c)
          Ca[e] = g[5] * da + g[4] * db + g[2] * dc;
          Cb[e] = g[4] * da + g[3] * db + g[1] * dc;
          Cc[e] = g[2] * da + g[1] * db + g[0] * dc;
        }
        else if constexpr (LIT)
          Cc[e] = g[2] * da + g[1] * db + g[0] * u;
        if constexpr (MASSTERM)
          Cm[e] = dw * u;
      }
      __syncthreads();
      // B_c: entries (qb, c')
      for (int e = tid; e < NQ; e += GT)
      {
        const int qb = e / N, c = e - qb * N;
        T s1 = T(0), s2 = T(0), s3 = T(0);
        for (int qc = 0; qc < Q; ++qc)
        {
          const T p = R0[c * Q + qc];
          const int ci = qb * Q + qc;
          if constexpr (FULL)
          {
            s1 += p * Ca[ci];
            s2 += p * Cb[ci];
          }
          if constexpr (GRAD)
            s3 += R1[c * Q + qc] * Cc[ci];
          if constexpr (MASSTERM)
            s3 += p * Cm[ci];
        }
        if constexpr (FULL)
        {
          D1[e] = s1;
          D2[e] = s2;
        }
        D3[e] = s3;
      }
      __syncthreads();
      // B_b and B_a: thread per output column, into registers.  The next
      // plane's F_a writes X only (last read in B_c, before the barrier
      // above) and its F_b writes Y after the barrier that follows F_a.
#pragma unroll
      for (int k = 0; k < KMAX; ++k)
      {
        const int col = tid + k * GT;
        if (col < N2)
        {
          const int b = col / N, c = col - b * N;
          T e1 = T(0), e2 = T(0);
          for (int qb = 0; qb < Q; ++qb)
          {
            const T p = R0[b * Q + qb];
            const int di = qb * N + c;
            if constexpr (FULL)
            {
              e1 += p * D1[di];
              e2 += R1[b * Q + qb] * D2[di];
            }
            e2 += p * D3[di];
          }
#pragma unroll
          for (int a = 0; a < NMAX; ++a)
          {
            if (a < N)
            {
              T v = R0[a * Q + qa] * e2;
              if constexpr (FULL)
                v += R1[a * Q + qa] * e1;
              w[k][a] += v;
            }
          }
        }
      }
    }

#pragma unroll
    for (int k = 0; k < KMAX; ++k)
    {
      const int col = tid + k * GT;
      if (col < N2)
      {
#pragma unroll
        for (int a = 0; a < NMAX; ++a)
        {
          if (a < N)
          {
            const int64_t li = cb + (int64_t)a * N2 + col;
            if constexpr (GATHER)
              gen_atomic_add(y + dofmap[li], w[k][a]);
            else
              y[li] = w[k][a];
          }
        }
      }
    }
    // the next cell's F_a rewrites X, last read in this cell's B_c
  }
}

template <typename T, int OP, bool REF, bool GATHER, int KMAX, int NMAX,
          int GT>
int gen_launch_shape(const ApplyReq& r)
{
  constexpr bool AFF = OP >= OP_MASS_AFF;
  if (r.N > NMAX || r.N * r.N > KMAX * GT)
    return FEMTC_ERR_INVALID_ARGUMENT;
  const size_t bytes = gen_lds_elems(r.N, r.Q, REF, AFF) * sizeof(T);
  if (bytes > 64 * 1024)
    return FEMTC_ERR_INVALID_ARGUMENT;
  const int64_t grid64 = std::min<int64_t>(r.num_cells, 8192);
  hipLaunchKernelGGL((gen_cells<T, OP, REF, GATHER, KMAX, NMAX, GT>),
                     dim3((unsigned)grid64), dim3(GT), bytes, r.stream,
                     (const T*)r.tab, (const T*)r.x, (T*)r.y, (const T*)r.G,
                     (const T*)r.detJw, r.dofmap, r.num_cells, r.N, r.Q);
  return hipGetLastError() == hipSuccess ? FEMTC_OK : FEMTC_ERR_HIP;
}

template <typename T, int OP, bool REF, bool GATHER>
int gen_launch(const ApplyReq& r)
{
  if (r.N <= 16)
    return gen_launch_shape<T, OP, REF, GATHER, 1, 16, 256>(r);
  if (r.N <= 22)
    return gen_launch_shape<T, OP, REF, GATHER, 1, 22, 512>(r);
  return gen_launch_shape<T, OP, REF, GATHER, 1, GEN_NMAX, 768>(r);
}

template <typename T, int OP, bool REF>
int gen_gather(const ApplyReq& r)
{
  return r.gather ? gen_launch<T, OP, REF, true>(r)
                  : gen_launch<T, OP, REF, false>(r);
}

template <typename T>
int gen_op(const ApplyReq& r)
{
  const bool ref = r.reference_mode != 0;
  if (ref && r.Q != r.N)
    return FEMTC_ERR_INVALID_ARGUMENT;
  switch (r.op)
  {
  case OP_MASS:
    return ref ? gen_gather<T, OP_MASS, true>(r)
               : gen_gather<T, OP_MASS, false>(r);
  case OP_STIFF:
    return ref ? gen_gather<T, OP_STIFF_LIT, true>(r)
               : gen_gather<T, OP_STIFF, false>(r);
  case OP_KPM:
    return ref ? FEMTC_ERR_INVALID_ARGUMENT : gen_gather<T, OP_KPM, false>(r);
  case OP_MASS_AFF:
    return ref ? FEMTC_ERR_INVALID_ARGUMENT
               : gen_gather<T, OP_MASS_AFF, false>(r);
  case OP_STIFF_AFF:
    return ref ? FEMTC_ERR_INVALID_ARGUMENT
               : gen_gather<T, OP_STIFF_AFF, false>(r);
  case OP_KPM_AFF:
    return ref ? FEMTC_ERR_INVALID_ARGUMENT
               : gen_gather<T, OP_KPM_AFF, false>(r);
  }
  return FEMTC_ERR_INVALID_ARGUMENT;
}

std::atomic<int> g_high_degree_mode{FEMTC_HIGH_DEGREE_AUTO};

bool line_disabled()
{
  return g_high_degree_mode.load(std::memory_order_relaxed) ==
         FEMTC_HIGH_DEGREE_GENERAL;
}

// The line kernels (apply_line.hpp) for N = 10..26 in correct mode: mass,
// stiffness and K+M with per-point geometry or on affine cells, cell-local
// or gathered.
// Returns FEMTC_ERR_INVALID_ARGUMENT when no line kernel covers the request
// (the general kernel then runs).
int line_launch(const ApplyReq& r)
{
  if (r.reference_mode || r.op == OP_STIFF_LIT)
    return FEMTC_ERR_INVALID_ARGUMENT;
  if (line_disabled())
    return FEMTC_ERR_INVALID_ARGUMENT;
  int dev = 0, max_lds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                            dev) != hipSuccess)
    return FEMTC_ERR_HIP;
  switch (r.N)
  {
  case 10: return launch_line_n10(r, max_lds);
  case 11: return launch_line_n11(r, max_lds);
  case 12: return launch_line_n12(r, max_lds);
  case 13: return launch_line_n13(r, max_lds);
  case 14: return launch_line_n14(r, max_lds);
  case 15: return launch_line_n15(r, max_lds);
  case 16: return launch_line_n16(r, max_lds);
  case 17: return launch_line_n17(r, max_lds);
  case 18: return launch_line_n18(r, max_lds);
  case 19: return launch_line_n19(r, max_lds);
  case 20: return launch_line_n20(r, max_lds);
  case 21: return launch_line_n21(r, max_lds);
  case 22: return launch_line_n22(r, max_lds);
  case 23: return launch_line_n23(r, max_lds);
  case 24: return launch_line_n24(r, max_lds);
  case 25: return launch_line_n25(r, max_lds);
  case 26: return launch_line_n26(r, max_lds);
  }
  return FEMTC_ERR_INVALID_ARGUMENT;
}

} // namespace

int launch_apply_general(const ApplyReq& r)
{
  if (r.N < 2 || r.N > GEN_NMAX)
    return FEMTC_ERR_DEGREE_NOT_SUPPORTED;
  if (r.Q < r.N || r.Q > r.N + 2)
    return FEMTC_ERR_INVALID_ARGUMENT;
  // plans, the solver and batches stay on the fused kernels
  if (r.blocked || r.assembled || r.eo_tab || r.energy || r.skip)
    return FEMTC_ERR_INVALID_ARGUMENT;
  if (r.num_cells <= 0)
    return FEMTC_OK;
  const int st = line_launch(r);
  if (st != FEMTC_ERR_INVALID_ARGUMENT)
    return st;
  if (r.dtype == FEMTC_F64)
    return gen_op<double>(r);
  if (r.dtype == FEMTC_F32)
    return gen_op<float>(r);
  return FEMTC_ERR_INVALID_ARGUMENT;
}

} // namespace femtc

namespace femtc
{
namespace
{
// line_fits for a runtime (dtype, op, N, Q), N = LINE_NMIN..LINE_NMAX,
// Q - N in 0..2 (constexpr evaluation only; no kernel is instantiated)
template <int N>
bool line_fits_n(int dtype, int op, int dq)
{
  if constexpr (N > LINE_NMAX)
    return false;
  else
  {
    auto pick = [&](auto t) {
      using T = decltype(t);
      auto byop = [&](auto qc) {
        constexpr int Q = N + decltype(qc)::value;
        return op == OP_MASS    ? line_fits<T, OP_MASS, N, Q>()
               : op == OP_STIFF ? line_fits<T, OP_STIFF, N, Q>()
                                : line_fits<T, OP_KPM, N, Q>();
      };
      return dq == 0   ? byop(std::integral_constant<int, 0>{})
             : dq == 1 ? byop(std::integral_constant<int, 1>{})
                       : byop(std::integral_constant<int, 2>{});
    };
    return dtype == FEMTC_F64 ? pick(double{}) : pick(float{});
  }
}

template <int N>
bool line_fits_rt(int dtype, int op, int n, int dq)
{
  if constexpr (N > LINE_NMAX)
    return false;
  else
    return n == N ? line_fits_n<N>(dtype, op, dq)
                  : line_fits_rt<N + 1>(dtype, op, n, dq);
}
} // namespace
} // namespace femtc

extern "C" int femtc_apply_kernel(int op, int dtype, int degree, int nq,
                                  int mode, int* kernel)
{
  using namespace femtc;
  if (!kernel)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "kernel must not be null");
  int st = check_shape(degree, nq, mode);
  if (st != FEMTC_OK)
    return st;
  if ((st = check_dtype(dtype)) != FEMTC_OK)
    return st;
  if (op != FEMTC_OP_MASS && op != FEMTC_OP_STIFFNESS && op != FEMTC_OP_KPM)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "op must be FEMTC_OP_*");
  const int N = degree + 1;
  if (degree <= FEMTC_MAX_FUSED_DEGREE)
    *kernel = FEMTC_KERNEL_FUSED;
  else if (mode == FEMTC_MODE_CORRECT && !line_disabled() &&
           N >= LINE_NMIN &&
           line_fits_rt<LINE_NMIN>(dtype, op, N, nq - N))
    *kernel = FEMTC_KERNEL_LINE;
  else
    *kernel = FEMTC_KERNEL_GENERAL;
  return FEMTC_OK;
}

extern "C" int femtc_set_high_degree_kernel(int mode, int* previous)
{
  if (mode != FEMTC_HIGH_DEGREE_AUTO && mode != FEMTC_HIGH_DEGREE_GENERAL)
    return femtc::fail(FEMTC_ERR_INVALID_ARGUMENT,
                       "high-degree kernel mode must be FEMTC_HIGH_DEGREE_AUTO "
                       "or FEMTC_HIGH_DEGREE_GENERAL");
  const int prev = femtc::g_high_degree_mode.exchange(mode);
  if (previous)
    *previous = prev;
  return FEMTC_OK;
}
