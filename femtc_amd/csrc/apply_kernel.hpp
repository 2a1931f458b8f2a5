// Fused sum-factorized cell operator for gfx950.
//
// One workgroup applies the operator to CPB cells.  Each cell gets Q*Q
// threads; a thread owns one "line" of a 3-D cell tensor and performs the
// 1-D contraction along that line in registers, with the table entries as
// wave-uniform (scalar) operands.  Between contractions along different axes
// the partial tensors pass through LDS.  The quadrature-point stage runs in
// the registers of thread (qa, qb), which owns the Q contiguous points
// (qa, qb, 0..Q-1), so the geometry stream G[cell][q][6] is read as one
// contiguous 48*Q-byte run per thread and never staged.
//
// Contraction schedule (forward: tables read [dof][q]; backward: the
// transpose, or the table as stored in reference mode):
//   F1  axis a  lines (b,c)   : A_p = phi_a U,      A_g = g_a U
//   F2  axis b  lines (qa,c)  : B_pp = phi_b A_p,   B_pg = g_b A_p,
//                               B_gp = phi_b A_g
//   F3  axis c  lines (qa,qb) : u = phi_c B_pp, d_c = g_c B_pp,
//                               d_b = phi_c B_pg, d_a = phi_c B_gp
//   pointwise                 : (v_a,v_b,v_c) = S(q) (d_a,d_b,d_c), v_m = detJw u
//   B3  axis c  lines (qa,qb) : R_a = phi_c' v_a, R_b = phi_c' v_b,
//                               R_0 = g_c' v_c + phi_c' v_m
//   B2  axis b  lines (qa,c)  : S_a = phi_b' R_a, S_0 = g_b' R_b + phi_b' R_0
//   B1  axis a  lines (b,c)   : W = g_a' S_a + phi_a' S_0
// which is W = sum_ij B_i^T S_ij B_j U (+ B^T detJw B U) with
// B_a = g(x)phi(x)phi etc.  This is the contraction chain of
// src/stiffness.hpp:68-97 with the intended transposes and sums; the
// reference's literal variant (OP_STIFF_LIT) keeps only the chain
// W = chain(phi,phi,g)(G2 d_a + G1 d_b + G0 u) (src/stiffness.hpp:76-97).
#pragma once

#include "common.hpp"

namespace femtc
{

template <int Q>
constexpr int cells_per_block()
{
  return (256 / (Q * Q)) < 1 ? 1 : (256 / (Q * Q));
}

template <int Q>
constexpr int block_threads()
{
  return ((cells_per_block<Q>() * Q * Q + 63) / 64) * 64;
}

__device__ inline void atomic_add(double* p, double v)
{
  unsafeAtomicAdd(p, v);
}
__device__ inline void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }

template <typename T>
__device__ inline T stream_load(const T* p)
{
  return __builtin_nontemporal_load(p);
}

template <typename T, int N, int Q, int OP, bool REF, bool GATHER>
__global__ void __launch_bounds__(block_threads<Q>())
    apply_cells(const T* __restrict__ tab, const T* __restrict__ x,
                T* __restrict__ y, const T* __restrict__ G,
                const T* __restrict__ detJw, const int32_t* __restrict__ dofmap,
                long long num_cells)
{
  constexpr int CPB = cells_per_block<Q>();
  constexpr int NL = Q * Q;
  constexpr bool LIT = (OP == OP_STIFF_LIT);
  constexpr bool MASSTERM = (OP == OP_MASS || OP == OP_KPM);
  constexpr bool GRAD = (OP != OP_MASS);
  constexpr bool DC = (OP == OP_STIFF || OP == OP_KPM); // d/dc needed
  constexpr bool UQ = MASSTERM || LIT;                  // value at q needed
  constexpr bool CHAIN_AB = DC;                         // R_a/R_b chains
  constexpr int NX = GRAD ? 2 : 1;
  constexpr int NY = GRAD ? 3 : 1;
  constexpr int SX = Q * N * N;
  constexpr int SY = Q * Q * N;
  constexpr int NN = N * N;
  constexpr int N3 = N * N * N;
  constexpr int Q3 = Q * Q * Q;
  // LDS slots
  constexpr int X_P = 0, X_G = 1;              // F1 -> F2
  constexpr int Y_PP = 0, Y_PG = 1, Y_GP = 2;  // F2 -> F3
  constexpr int Y_RA = 0, Y_RB = 1, Y_R0 = NY - 1; // B3 -> B2
  constexpr int X_SA = 0, X_S0 = NX - 1;       // B2 -> B1

  __shared__ T Xs[CPB * NX * SX];
  __shared__ T Ys[CPB * NY * SY];

  const T* __restrict__ ph = tab;
  const T* __restrict__ gr = tab + N * Q;
  // backward table entry for (dof d, point q)
  auto bph = [&](int d, int q) -> T { return REF ? ph[q * N + d] : ph[d * Q + q]; };
  auto bgr = [&](int d, int q) -> T { return REF ? gr[q * N + d] : gr[d * Q + q]; };

  const int tid = threadIdx.x;
  const int lc = tid / NL;
  const int ln = tid - lc * NL;
  const long long cell = (long long)blockIdx.x * CPB + lc;
  const bool live = (lc < CPB) && (cell < num_cells);
  T* Xc = Xs + (lc < CPB ? lc : 0) * NX * SX;
  T* Yc = Ys + (lc < CPB ? lc : 0) * NY * SY;

  // ---------------- global loads issued up front --------------------------
  T u[N];
  int32_t idx[GATHER ? N : 1];
  if (live && ln < NN)
  {
    const long long base = cell * N3 + ln;
#pragma unroll
    for (int a = 0; a < N; ++a)
    {
      if constexpr (GATHER)
      {
        idx[a] = dofmap[base + a * NN];
        u[a] = x[idx[a]];
      }
      else
        u[a] = x[base + a * NN];
    }
  }
  // geometry of the Q points (qa, qb, 0..Q-1) owned by this thread
  T g6[GRAD ? 6 * Q : 1];
  T dw[MASSTERM ? Q : 1];
  if (live)
  {
    if constexpr (GRAD)
    {
      const T* Gp = G + (cell * Q3 + (long long)ln * Q) * 6;
#pragma unroll
      for (int i = 0; i < 6 * Q; ++i)
        g6[i] = stream_load(Gp + i);
    }
    if constexpr (MASSTERM)
    {
      const T* Dp = detJw + cell * Q3 + (long long)ln * Q;
#pragma unroll
      for (int i = 0; i < Q; ++i)
        dw[i] = stream_load(Dp + i);
    }
  }

  // ---------------- F1: contract axis a ------------------------------------
  if (live && ln < NN)
  {
#pragma unroll
    for (int qa = 0; qa < Q; ++qa)
    {
      T sp = 0, sg = 0;
#pragma unroll
      for (int a = 0; a < N; ++a)
      {
        sp = fma(ph[a * Q + qa], u[a], sp);
        if constexpr (GRAD)
          sg = fma(gr[a * Q + qa], u[a], sg);
      }
      Xc[X_P * SX + qa * NN + ln] = sp;
      if constexpr (GRAD)
        Xc[X_G * SX + qa * NN + ln] = sg;
    }
  }
  __syncthreads();

  // ---------------- F2: contract axis b ------------------------------------
  if (live && ln < Q * N)
  {
    const int qa = ln / N, c = ln - (ln / N) * N;
    T ap[N], ag[N];
#pragma unroll
    for (int b = 0; b < N; ++b)
    {
      ap[b] = Xc[X_P * SX + qa * NN + b * N + c];
      if constexpr (GRAD)
        ag[b] = Xc[X_G * SX + qa * NN + b * N + c];
    }
#pragma unroll
    for (int qb = 0; qb < Q; ++qb)
    {
      T spp = 0, spg = 0, sgp = 0;
#pragma unroll
      for (int b = 0; b < N; ++b)
      {
        spp = fma(ph[b * Q + qb], ap[b], spp);
        if constexpr (GRAD)
        {
          spg = fma(gr[b * Q + qb], ap[b], spg);
          sgp = fma(ph[b * Q + qb], ag[b], sgp);
        }
      }
      const int o = (qa * Q + qb) * N + c;
      Yc[Y_PP * SY + o] = spp;
      if constexpr (GRAD)
      {
        Yc[Y_PG * SY + o] = spg;
        Yc[Y_GP * SY + o] = sgp;
      }
    }
  }
  __syncthreads();

  // ---------------- F3 + quadrature-point work + B3 -----------------------
  T bpp[N], bpg[N], bgp[N];
  if (live)
  {
#pragma unroll
    for (int c = 0; c < N; ++c)
    {
      bpp[c] = Yc[Y_PP * SY + ln * N + c];
      if constexpr (GRAD)
      {
        bpg[c] = Yc[Y_PG * SY + ln * N + c];
        bgp[c] = Yc[Y_GP * SY + ln * N + c];
      }
    }
  }
  __syncthreads(); // Ys is rewritten by B3 below

  if (live)
  {
    T va[Q], vb[Q], v0[Q], vm[Q];
#pragma unroll
    for (int qc = 0; qc < Q; ++qc)
    {
      T uq = 0, dc = 0, db = 0, da = 0;
#pragma unroll
      for (int c = 0; c < N; ++c)
      {
        if constexpr (UQ)
          uq = fma(ph[c * Q + qc], bpp[c], uq);
        if constexpr (DC)
          dc = fma(gr[c * Q + qc], bpp[c], dc);
        if constexpr (GRAD)
        {
          db = fma(ph[c * Q + qc], bpg[c], db);
          da = fma(ph[c * Q + qc], bgp[c], da);
        }
      }
      if constexpr (DC)
      {
        const T* g = g6 + 6 * qc;
        // S = [[G5,G4,G2],[G4,G3,G1],[G2,G1,G0]] acting on (d_a, d_b, d_c)
        va[qc] = g[5] * da + g[4] * db + g[2] * dc;
        vb[qc] = g[4] * da + g[3] * db + g[1] * dc;
        v0[qc] = g[2] * da + g[1] * db + g[0] * dc;
      }
      else if constexpr (LIT)
      {
        const T* g = g6 + 6 * qc;
        // third row of the transform applied to (d_a, d_b, u)
        // (src/stiffness.hpp:27 with Wz undifferentiated)
        v0[qc] = g[2] * da + g[1] * db + g[0] * uq;
      }
      if constexpr (MASSTERM)
        vm[qc] = dw[qc] * uq;
    }
    // B3: contract axis c (transposed)
#pragma unroll
    for (int c = 0; c < N; ++c)
    {
      T ra = 0, rb = 0, r0 = 0;
#pragma unroll
      for (int qc = 0; qc < Q; ++qc)
      {
        if constexpr (CHAIN_AB)
        {
          ra = fma(bph(c, qc), va[qc], ra);
          rb = fma(bph(c, qc), vb[qc], rb);
        }
        if constexpr (GRAD)
          r0 = fma(bgr(c, qc), v0[qc], r0);
        if constexpr (MASSTERM)
          r0 = fma(bph(c, qc), vm[qc], r0);
      }
      const int o = ln * N + c;
      if constexpr (CHAIN_AB)
      {
        Yc[Y_RA * SY + o] = ra;
        Yc[Y_RB * SY + o] = rb;
      }
      Yc[Y_R0 * SY + o] = r0;
    }
  }
  __syncthreads();

  // ---------------- B2: contract axis b (transposed) ------------------------
  if (live && ln < Q * N)
  {
    const int qa = ln / N, c = ln - (ln / N) * N;
    T ra[Q], rb[Q], r0[Q];
#pragma unroll
    for (int qb = 0; qb < Q; ++qb)
    {
      const int o = (qa * Q + qb) * N + c;
      if constexpr (CHAIN_AB)
      {
        ra[qb] = Yc[Y_RA * SY + o];
        rb[qb] = Yc[Y_RB * SY + o];
      }
      r0[qb] = Yc[Y_R0 * SY + o];
    }
#pragma unroll
    for (int b = 0; b < N; ++b)
    {
      T sa = 0, s0 = 0;
#pragma unroll
      for (int qb = 0; qb < Q; ++qb)
      {
        if constexpr (CHAIN_AB)
        {
          sa = fma(bph(b, qb), ra[qb], sa);
          s0 = fma(bgr(b, qb), rb[qb], s0);
        }
        s0 = fma(bph(b, qb), r0[qb], s0);
      }
      const int o = qa * NN + b * N + c;
      if constexpr (CHAIN_AB)
        Xc[X_SA * SX + o] = sa;
      Xc[X_S0 * SX + o] = s0;
    }
  }
  __syncthreads();

  // ---------------- B1: contract axis a (transposed) + store/scatter ------
  if (live && ln < NN)
  {
    T sa[Q], s0[Q];
#pragma unroll
    for (int qa = 0; qa < Q; ++qa)
    {
      if constexpr (CHAIN_AB)
        sa[qa] = Xc[X_SA * SX + qa * NN + ln];
      s0[qa] = Xc[X_S0 * SX + qa * NN + ln];
    }
    const long long base = cell * N3 + ln;
#pragma unroll
    for (int a = 0; a < N; ++a)
    {
      T w = 0;
#pragma unroll
      for (int qa = 0; qa < Q; ++qa)
      {
        if constexpr (CHAIN_AB)
          w = fma(bgr(a, qa), sa[qa], w);
        w = fma(bph(a, qa), s0[qa], w);
      }
      if constexpr (GATHER)
        atomic_add(y + idx[a], w);
      else
        y[base + a * NN] = w;
    }
  }
}

// Launch one instance; the grid covers num_cells in blocks of CPB cells.
template <typename T, int N, int Q, int OP, bool REF, bool GATHER>
int launch_cells(const ApplyReq& r)
{
  constexpr int CPB = cells_per_block<Q>();
  const long long nblocks = (r.num_cells + CPB - 1) / CPB;
  if (nblocks == 0)
    return FEMTC_OK;
  if (nblocks > 0x7fffffffLL)
    return FEMTC_ERR_INVALID_ARGUMENT;
  hipLaunchKernelGGL((apply_cells<T, N, Q, OP, REF, GATHER>),
                     dim3((unsigned)nblocks), dim3(block_threads<Q>()), 0,
                     r.stream, (const T*)r.tab, (const T*)r.x, (T*)r.y,
                     (const T*)r.G, (const T*)r.detJw, r.dofmap,
                     (long long)r.num_cells);
  return hip_status(hipGetLastError());
}

template <typename T, int N, int Q>
int dispatch_op(const ApplyReq& r)
{
  const bool g = r.gather != 0;
  switch (r.op)
  {
  case OP_MASS:
    if (r.reference_mode)
    {
      if constexpr (N == Q)
        return g ? launch_cells<T, N, Q, OP_MASS, true, true>(r)
                 : launch_cells<T, N, Q, OP_MASS, true, false>(r);
      return FEMTC_ERR_INVALID_ARGUMENT;
    }
    return g ? launch_cells<T, N, Q, OP_MASS, false, true>(r)
             : launch_cells<T, N, Q, OP_MASS, false, false>(r);
  case OP_STIFF:
    if (r.reference_mode)
    {
      if constexpr (N == Q)
        return g ? launch_cells<T, N, Q, OP_STIFF_LIT, true, true>(r)
                 : launch_cells<T, N, Q, OP_STIFF_LIT, true, false>(r);
      return FEMTC_ERR_INVALID_ARGUMENT;
    }
    return g ? launch_cells<T, N, Q, OP_STIFF, false, true>(r)
             : launch_cells<T, N, Q, OP_STIFF, false, false>(r);
  case OP_KPM:
    if (r.reference_mode)
      return FEMTC_ERR_INVALID_ARGUMENT;
    return g ? launch_cells<T, N, Q, OP_KPM, false, true>(r)
             : launch_cells<T, N, Q, OP_KPM, false, false>(r);
  }
  return FEMTC_ERR_INVALID_ARGUMENT;
}

template <int N, int Q>
int dispatch_dtype(const ApplyReq& r)
{
  if (r.dtype == FEMTC_F64)
    return dispatch_op<double, N, Q>(r);
  if (r.dtype == FEMTC_F32)
    return dispatch_op<float, N, Q>(r);
  return FEMTC_ERR_INVALID_ARGUMENT;
}

// Q in {N, N+1, N+2}
template <int N>
int dispatch_q(const ApplyReq& r)
{
  if (r.N != N)
    return FEMTC_ERR_DEGREE_NOT_SUPPORTED;
  switch (r.Q - N)
  {
  case 0: return dispatch_dtype<N, N>(r);
  case 1: return dispatch_dtype<N, N + 1>(r);
  case 2: return dispatch_dtype<N, N + 2>(r);
  }
  return FEMTC_ERR_INVALID_ARGUMENT;
}

} // namespace femtc
