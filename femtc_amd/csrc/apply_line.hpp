// Line kernels for the high degrees (the top of the reference's dispatch
// ladder, src/mass.hpp:52-138 and src/stiffness.hpp:102-188): cell-local or
// gathered mass, stiffness and K+M in correct mode, with per-point geometry
// or on affine cells (7 values per cell), N = LINE_NMIN..LINE_NMAX dofs per
// direction, Q in {N, N+1, N+2}, FP64 and FP32, for the shapes whose
// registers and LDS fit (line_fits).
//
// One workgroup takes one cell at a time.  Every contraction is done by
// threads that own a whole line of the tensor along the contracted axis,
// in registers; the 1D tables are the same for every lane, so the
// compiler reads them with scalar loads and every multiply-add takes its
// table operand from a scalar register.  LDS only moves the tensor between
// owners (one write and one read per value and stage), where the general
// kernel reads one or two LDS operands per multiply-add.
//
//   owner (b,c)   F_a   A0 = phi_a U, A1 = g_a U          U from global
//   owner (qa,c)  F_b   B0 = phi_b A0, B1 = g_b A0, B2 = phi_b A1
//   owner (qa,qb) F_c   u = phi_c B0, d_c = g_c B0, d_b = phi_c B1,
//                       d_a = phi_c B2
//                 pointwise (v_a, v_b, v_c) = S(q) (d_a, d_b, d_c),
//                       v_m = detJw u
//                 B_c   D1 = phi_c' v_a, D2 = phi_c' v_b,
//                       D3 = g_c' v_c + phi_c' v_m
//   owner (qa,c') B_b   E1 = phi_b' D1, E2 = g_b' D2 + phi_b' D3
//   owner (b',c') B_a   W = g_a' E1 + phi_a' E2              W to global
//
// Each arrow between owners is one transpose through LDS.  Two LDS slices
// alternate, so one barrier per transpose orders both the slice's write
// after its last read (two transposes back) and its read after its write;
// cells too large for two slices use one, with two barriers per transpose.
#pragma once

#include <algorithm>
#include <utility>

#include "common.hpp"

namespace femtc
{

constexpr int LINE_NMIN = 10;
constexpr int LINE_NMAX = 26;
#ifndef LINE_STIFF64_LIVE
#define LINE_STIFF64_LIVE 73
#endif
#ifndef LINE_KPM64_LIVE
#define LINE_KPM64_LIVE 66
#endif
#ifndef LINE_PREFETCH
#define LINE_PREFETCH 1
#endif
#ifndef LINE_PAIR
#define LINE_PAIR 2
#endif
#ifndef LINE_CHUNK
#define LINE_CHUNK 128
#endif

template <int N, int Q>
struct LineShape
{
  static constexpr int NN = N * N, QN = Q * N, QQ = Q * Q;
  // odd row pitch of the slices that (qa,qb) owners read or write along c:
  // lanes then hit distinct LDS banks
  static constexpr int NP = N | 1;
  static constexpr int LINES = QQ; // Q >= N
  static constexpr int TH = (LINES + 63) / 64 * 64;
  static constexpr int SLICE = Q * NN > QQ * NP ? Q * NN : QQ * NP;
};

// The stored-geometry operator an affine one computes (OP_*_AFF -> OP_*).
constexpr int line_base_op(int op)
{
  return op >= OP_MASS_AFF ? op - OP_MASS_AFF : op;
}

// Waves per SIMD a cell's registers are sized for: 3 (168 VGPRs), or 2
// (256 VGPRs) for FP64 stiffness and K+M when the values live at once --
// the three gradient lines (and u) at Q points plus one line of N -- do not
// fit 168 without spilling (bounds from the compiler's counts over the
// instantiated shapes).
template <typename T, int OP, int N, int Q>
constexpr int line_want()
{
  constexpr bool f64 = sizeof(T) == 8;
  constexpr int B = line_base_op(OP);
  return (f64 && B == OP_STIFF && 3 * Q + N > LINE_STIFF64_LIVE)  ? 2
         : (f64 && B == OP_KPM && 4 * Q + N > LINE_KPM64_LIVE) ? 2
                                                                : 3;
}

// LDS slices per cell: two alternating slices (one barrier per transpose)
// when they fit the 160 KB of a CU, else one (two barriers per transpose).
template <typename T, int N, int Q>
constexpr int line_nbuf()
{
  return 2 * (size_t)LineShape<N, Q>::SLICE * sizeof(T) <= 160 * 1024 ? 2
                                                                      : 1;
}

// LDS of one cell.
template <typename T, int N, int Q>
constexpr size_t line_cell_lds()
{
  return line_nbuf<T, N, Q>() * (size_t)LineShape<N, Q>::SLICE * sizeof(T);
}

// Registers a cell's values need: the values live at once (mass: a line of
// Q and one of N; stiffness: three of Q and one of N; K+M: four of Q and one
// of N) plus about 24 of addresses and temporaries (the compiler's counts
// over N <= 20).
template <typename T, int OP, int N, int Q>
constexpr int line_vgprs()
{
  constexpr int B = line_base_op(OP);
  constexpr int live = B == OP_MASS ? Q + N : B == OP_STIFF ? 3 * Q + N
                                                            : 4 * Q + N;
  return live * (int)(sizeof(T) / 4) + 24;
}

// VGPRs per wave with w waves per SIMD (arch VGPRs, 8-register granules).
constexpr int line_budget(int w)
{
  return w <= 2 ? 256 : (512 / w) / 8 * 8;
}

// Cells per workgroup: 2 when one cell takes 5 to 7 waves, its registers
// fit 3 waves per SIMD and two cells' LDS fits the 160 KB of a CU
// (measured: a second workgroup of 5 waves does not become resident beside
// the first, so one workgroup of two cells doubles the waves per CU), else
// 1.
template <typename T, int OP, int N, int Q>
constexpr int line_cpb()
{
  constexpr int w1 = LineShape<N, Q>::TH / 64;
  constexpr int w2 = (2 * w1 + 3) / 4; // waves per SIMD for two cells
  return (w1 >= 5 && w1 <= 7 && line_want<T, OP, N, Q>() == 3 &&
          2 * line_cell_lds<T, N, Q>() <= 160 * 1024 &&
          line_vgprs<T, OP, N, Q>() <= line_budget(w2 > 3 ? w2 : 3))
             ? LINE_PAIR
             : 1;
}

template <typename T, int OP, int N, int Q>
constexpr size_t line_lds_bytes()
{
  return line_cpb<T, OP, N, Q>() * line_cell_lds<T, N, Q>();
}

template <typename T, int OP, int N, int Q>
constexpr int line_threads()
{
  return line_cpb<T, OP, N, Q>() * LineShape<N, Q>::TH;
}

// Waves per SIMD the compiler keeps registers for: line_want, and never
// fewer than one workgroup needs.
template <typename T, int OP, int N, int Q>
constexpr int line_waves()
{
  constexpr int wg = (line_threads<T, OP, N, Q>() / 64 + 3) / 4;
  constexpr int want = line_want<T, OP, N, Q>();
  return want > wg ? want : wg;
}

// Whether a line kernel exists for (T, OP, N, Q): one cell's LDS fits a CU
// and line_vgprs fits the registers of the waves per SIMD the workgroup
// needs.  Other shapes run on the general kernel.
template <typename T, int OP, int N, int Q>
constexpr bool line_fits()
{
  return line_vgprs<T, OP, N, Q>() <= line_budget(line_waves<T, OP, N, Q>()) &&
         line_cell_lds<T, N, Q>() <= 160 * 1024 &&
         line_threads<T, OP, N, Q>() <= 1024;
}

__device__ __forceinline__ double line_fma(double a, double b, double c)
{
  return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float line_fma(float a, float b, float c)
{
  return __builtin_fmaf(a, b, c);
}

// Tables are read through a pointer in the constant address space, which
// the compiler reads with scalar loads, laundered through an empty asm per
// chunk of LINE_CHUNK bytes: each contraction then loads its table one
// chunk ahead of the multiply-adds that use it, instead of the compiler
// hoisting a whole table (up to 2 x 26 x 28 values) into scalar registers
// at once.  128-byte chunks measured up to 8 % faster than 64-byte chunks
// and than two 64-byte chunks in flight.
template <typename T>
using line_cptr = const __attribute__((address_space(4))) T*;

template <typename T>
__device__ __forceinline__ line_cptr<T> line_opaque(const T* p)
{
  line_cptr<T> c = (line_cptr<T>)p;
  asm volatile("" : "+s"(c));
  return c;
}

// f(integral_constant<int, i>) for i = 0..COUNT-1, expanded at compile time
// (register arrays indexed by i stay in registers whatever the loop size).
template <typename F, int... I>
__device__ __forceinline__ void line_for_impl(F&& f,
                                              std::integer_sequence<int, I...>)
{
  (f(std::integral_constant<int, I>{}), ...);
}
template <int COUNT, typename F>
__device__ __forceinline__ void line_static_for(F&& f)
{
  line_for_impl(f, std::make_integer_sequence<int, COUNT>{});
}

// One contraction with the table M (row-major, NI * NO values) walked in
// storage order, one chunk of K values in flight:
//   BWD false: out[o] (+)= sum_i M[i][o] in[i]  (M [NI][NO], a forward
//              table [dof][q]); consecutive products feed distinct outputs;
//   BWD true:  out[o] (+)= sum_i M[o][i] in[i]  (M [NO][NI], the same table
//              read backward); each output is one chain of products, whose
//              latency the other resident waves cover.
template <typename T, int NI, int NO, bool BWD, bool ACC = false>
__device__ __forceinline__ void line_contract(const T* Mt, const T (&in)[NI],
                                              T (&out)[NO])
{
  constexpr int TOT = NI * NO;
  constexpr int K = LINE_CHUNK / (int)sizeof(T);
  constexpr int NCH = (TOT + K - 1) / K;
  // chunks in flight ahead of the one in use (a ring of D + 1 buffers)
  constexpr int D = LINE_PREFETCH < NCH ? LINE_PREFETCH : NCH;
  T ring[D + 1][K];
  // the chunk addresses derive from a base that changes every call, so
  // they are not hoisted out of the cell loop either
  asm volatile("" : "+s"(Mt));
  auto load = [&](int k, T (&b)[K]) {
    const line_cptr<T> M = line_opaque(Mt + k * K);
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (k * K + j < TOT)
        b[j] = M[j];
  };
  auto product = [&](const T (&cur)[K], int k, int j) {
    const int f = k * K + j;
    if (f < TOT)
    {
      if constexpr (!BWD)
      {
        const int i = f / NO, o = f - (f / NO) * NO;
        out[o] = (!ACC && i == 0) ? cur[j] * in[i]
                                  : line_fma(cur[j], in[i], out[o]);
      }
      else
      {
        const int o = f / NI, i = f - (f / NI) * NI;
        out[o] = (!ACC && i == 0) ? cur[j] * in[i]
                                  : line_fma(cur[j], in[i], out[o]);
      }
    }
  };
  // pin: the product is computed before whatever follows this asm
  auto pin = [&](int k, int j, bool always) {
    const int f = k * K + j;
    if (f < TOT)
    {
      if constexpr (!BWD)
        asm volatile("" : "+v"(out[f - (f / NO) * NO]));
      else if (always || f - (f / NI) * NI == NI - 1 || j == K - 1)
        asm volatile("" : "+v"(out[f / NI]));
    }
  };
  line_static_for<D>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    load(k, ring[k]);
  });
  line_static_for<NCH>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    __builtin_amdgcn_sched_barrier(0);
    const T (&cur)[K] = ring[k % (D + 1)];
    // Scalar loads return out of order, so the wait before this chunk's
    // first product waits for every scalar load in flight.  The loads of
    // chunk k + D are issued after that product, not before, so that wait
    // does not include them.
    product(cur, k, 0);
    pin(k, 0, true);
    if constexpr (k + D < NCH)
      load(k + D, ring[(k + D) % (D + 1)]);
#pragma unroll
    for (int j = 1; j < K; ++j)
      product(cur, k, j);
    // the chunk's products are done here, before the loads issued with
    // the next chunk (ordered behind this asm)
#pragma unroll
    for (int j = 1; j < K; ++j)
      pin(k, j, false);
  });
}

template <typename T, int NI, int NO, bool ACC = false>
__device__ __forceinline__ void line_fwd(const T* M, const T (&in)[NI],
                                         T (&out)[NO])
{
  line_contract<T, NI, NO, false, ACC>(M, in, out);
}

template <typename T, int NI, int NO, bool ACC = false>
__device__ __forceinline__ void line_bwd(const T* M, const T (&in)[NI],
                                         T (&out)[NO])
{
  line_contract<T, NI, NO, true, ACC>(M, in, out);
}

// The quadrature-point stage of owner (qa,qb) = tq on its Q points
// (qa,qb,qc): (d_a, d_b, d_c) <- S(q) (d_a, d_b, d_c) with
// S = [[G5,G4,G2],[G4,G3,G1],[G2,G1,G0]] (stiffness), u <- detJw u (mass),
// one point at a time with the next point's geometry in flight.
template <typename T, int OP, int N, int Q, int NA, int NU>
__device__ __forceinline__ void line_pointwise(const T* __restrict__ G,
                                               const T* __restrict__ detJw,
                                               const T* __restrict__ tab,
                                               int64_t cell, int tq,
                                               T (&DA)[NA], T (&DB)[NA],
                                               T (&DC)[NA], T (&U)[NU])
{
  constexpr int B = line_base_op(OP);
  constexpr bool AFF = OP >= OP_MASS_AFF;
  constexpr bool FULL = B == OP_STIFF || B == OP_KPM;
  constexpr bool MT = B == OP_MASS || B == OP_KPM;
  constexpr int GV = (FULL ? 6 : 0) + (MT ? 1 : 0);
  const int64_t p0 = cell * (int64_t)(Q * Q * Q) + (int64_t)tq * Q;
  // affine cells: the cell's 7 values (detJ J^-1 J^-T packed as G, detJ;
  // the same for every lane of the wave) times the tensor-product Gauss
  // weights w[qa] w[qb] w[qc] (after the two tables)
  T cg[AFF ? CELLGEO : 1];
  T wab = T(1);
  const T* const wts = tab + 2 * N * Q;
  if constexpr (AFF)
  {
#pragma unroll
    for (int k = 0; k < CELLGEO; ++k)
      cg[k] = G[cell * CELLGEO + k];
    const int qa = tq / Q;
    wab = wts[qa] * wts[tq - qa * Q];
  }
  T gcur[GV], gnxt[GV];
  auto load = [&](int qc, T (&gb)[GV]) {
    if constexpr (AFF)
    {
      const T wq = wab * wts[qc];
      if constexpr (FULL)
      {
#pragma unroll
        for (int k = 0; k < 6; ++k)
          gb[k] = wq * cg[k];
      }
      if constexpr (MT)
        gb[GV - 1] = wq * cg[6];
    }
    else
    {
      if constexpr (FULL)
      {
        const T* g = G + (p0 + qc) * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k)
          gb[k] = g[k];
      }
      if constexpr (MT)
        gb[GV - 1] = detJw[p0 + qc];
    }
  };
  load(0, gcur);
  line_static_for<Q>([&](auto qcc) {
    constexpr int qc = decltype(qcc)::value;
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (qc + 1 < Q)
      load(qc + 1, gnxt);
    if constexpr (FULL)
    {
      const T da = DA[qc], db = DB[qc], dcv = DC[qc];
      DA[qc] = gcur[5] * da + gcur[4] * db + gcur[2] * dcv;
      DB[qc] = gcur[4] * da + gcur[3] * db + gcur[1] * dcv;
      DC[qc] = gcur[2] * da + gcur[1] * db + gcur[0] * dcv;
      asm volatile("" : "+v"(DA[qc]), "+v"(DB[qc]), "+v"(DC[qc]));
    }
    if constexpr (MT)
    {
      U[qc] *= gcur[GV - 1];
      asm volatile("" : "+v"(U[qc]));
    }
    if constexpr (qc + 1 < Q)
    {
#pragma unroll
      for (int k = 0; k < GV; ++k)
        gcur[k] = gnxt[k];
    }
  });
}

template <typename T, int OP, int N, int Q>
__global__ void __launch_bounds__((line_threads<T, OP, N, Q>()))
    __attribute__((amdgpu_waves_per_eu(line_waves<T, OP, N, Q>(),
                                       line_waves<T, OP, N, Q>())))
    line_cells(const T* __restrict__ tab, const T* __restrict__ x,
               T* __restrict__ y, const T* __restrict__ G,
               const T* __restrict__ detJw,
               const int32_t* __restrict__ dofmap, int64_t C)
{
  using S = LineShape<N, Q>;
  constexpr int OPB = line_base_op(OP);
  constexpr bool FULL = OPB == OP_STIFF || OPB == OP_KPM;
  constexpr bool MT = OPB == OP_MASS || OPB == OP_KPM;
  static_assert(FULL || MT, "mass, stiffness or K+M");
  constexpr int NN = S::NN, QN = S::QN, QQ = S::QQ, NP = S::NP;
  // transposes per cell: A0 A1 B0 B1 B2 D1 D2 D3 E1 E2, or A0 B0 D3 E2;
  // an even count keeps the slice parity the same from cell to cell
  constexpr int NTR = FULL ? 10 : 4;
  static_assert(NTR % 2 == 0, "even transpose count");

  // CPB cells per workgroup, each on its own S::TH threads (whole waves)
  // and its own two LDS slices
  constexpr int CPB = line_cpb<T, OP, N, Q>();
  constexpr int NB = line_nbuf<T, N, Q>();
  extern __shared__ __align__(16) unsigned char line_smem[];
  const T* const phi = tab; // [dof][q]
  const T* const dph = tab + N * Q;

  for (int64_t c0 = (int64_t)blockIdx.x * CPB; c0 < C;
       c0 += (int64_t)gridDim.x * CPB)
  {
    // The thread index passes through an empty asm every cell, so the
    // index arithmetic is redone next to its loads and stores instead of
    // being hoisted out of the cell loop into registers held throughout.
    // Every thread runs every contraction (a thread past the owners of a
    // stage works on the last owner's line, in bounds) and only the stores
    // are guarded: the contractions stay out of divergent branches.
    int tt = threadIdx.x;
    asm volatile("" : "+v"(tt));
    // the cell of this wave (every wave lies in one cell's threads)
    const int h = CPB > 1 ? __builtin_amdgcn_readfirstlane(tt / S::TH) : 0;
    const int t = tt - h * S::TH;
    const bool live = c0 + h < C; // a second cell past the last one
    const int64_t cell = live ? c0 + h : C - 1;
    T* const slice0 = (T*)line_smem + h * NB * S::SLICE;
    T* const slice1 = slice0 + (NB - 1) * S::SLICE;
    const bool own_bc = t < NN, own_ac = t < QN, own_ab = t < QQ;
    const int tbc = own_bc ? t : NN - 1;            // (b, c)
    const int tac = own_ac ? t : QN - 1;            // (qa, c)
    const int tab_ = own_ab ? t : QQ - 1;           // (qa, qb)
    const int ac_qa = tac / N, ac_c = tac - ac_qa * N;
    // the cell's blocks of x / y / dofmap (uniform bases, 32-bit offsets)
    const int64_t cb = cell * (int64_t)(N * NN);
    const T* const xc = dofmap ? x : x + cb;
    const int32_t* const dc = dofmap ? dofmap + cb : nullptr;
    // Two slices: transpose k writes slice k % 2 and reads it after
    // barrier k.  A write may come as soon as barrier k-1 has passed (the
    // slice's last reads, transpose k-2, come before it in every thread),
    // so each array goes to LDS as early as possible and leaves the
    // registers.  One slice (NB = 1, large cells): every write waits on a
    // barrier after the previous transpose's reads (wb).
    auto slice = [&](int k) { return (k & 1) ? slice1 : slice0; };
    auto wb = [&]() {
      if constexpr (NB == 1)
        __syncthreads();
    };
    // (b,c) writes [qa][b][c]; (qa,c) reads along b
    auto w_bc = [&](int k, const T (&v)[Q]) {
      if (own_bc)
      {
#pragma unroll
        for (int q = 0; q < Q; ++q)
          slice(k)[q * NN + tbc] = v[q];
      }
    };
    auto r_ac_b = [&](int k, T (&l)[N]) {
#pragma unroll
      for (int b = 0; b < N; ++b)
        l[b] = slice(k)[ac_qa * NN + b * N + ac_c];
    };
    // (qa,c) writes [qa][qb][c] (row pitch NP); (qa,qb) reads along c
    auto w_ac_q = [&](int k, const T (&v)[Q]) {
      if (own_ac)
      {
#pragma unroll
        for (int qb = 0; qb < Q; ++qb)
          slice(k)[(ac_qa * Q + qb) * NP + ac_c] = v[qb];
      }
    };
    auto r_ab_c = [&](int k, T (&l)[N]) {
#pragma unroll
      for (int c = 0; c < N; ++c)
        l[c] = slice(k)[tab_ * NP + c];
    };
    // (qa,qb) writes [qa][qb][c'] (row pitch NP); (qa,c') reads along qb
    auto w_ab = [&](int k, const T (&v)[N]) {
      if (own_ab)
      {
#pragma unroll
        for (int c = 0; c < N; ++c)
          slice(k)[tab_ * NP + c] = v[c];
      }
    };
    auto r_ac_qb = [&](int k, T (&l)[Q]) {
#pragma unroll
      for (int qb = 0; qb < Q; ++qb)
        l[qb] = slice(k)[(ac_qa * Q + qb) * NP + ac_c];
    };
    // (qa,c') writes [qa][b'][c']; (b',c') reads along qa
    auto w_ac_b = [&](int k, const T (&v)[N]) {
      if (own_ac)
      {
#pragma unroll
        for (int b = 0; b < N; ++b)
          slice(k)[ac_qa * NN + b * N + ac_c] = v[b];
      }
    };
    auto r_bc_qa = [&](int k, T (&l)[Q]) {
#pragma unroll
      for (int qa = 0; qa < Q; ++qa)
        l[qa] = slice(k)[qa * NN + tbc];
    };

    // ---- F_a: (b,c) reads its column of U
    T A0[Q], A1[FULL ? Q : 1];
    {
      T u[N];
#pragma unroll
      for (int a = 0; a < N; ++a)
      {
        const unsigned li = (unsigned)(a * NN + tbc);
        u[a] = dc ? xc[dc[li]] : xc[li];
      }
      line_fwd<T, N, Q>(phi, u, A0);
      if constexpr (FULL)
        line_fwd<T, N, Q>(dph, u, A1);
    }

    T W[N];
    if constexpr (FULL && NB == 1)
    {
      // the same transposes on one slice, each write after the reads of
      // the one before
      T B0[Q], B1[Q], B2[Q], l[N];
      wb();
      w_bc(0, A0);
      __syncthreads();
      r_ac_b(0, l);
      line_fwd<T, N, Q>(phi, l, B0);
      line_fwd<T, N, Q>(dph, l, B1);
      wb();
      w_bc(1, A1);
      __syncthreads();
      r_ac_b(1, l);
      line_fwd<T, N, Q>(phi, l, B2);
      wb();
      w_ac_q(2, B0);
      __syncthreads();
      T U[MT ? Q : 1], DA[Q], DB[Q], DC[Q];
      r_ab_c(2, l);
      if constexpr (MT)
        line_fwd<T, N, Q>(phi, l, U);
      line_fwd<T, N, Q>(dph, l, DC);
      wb();
      w_ac_q(3, B1);
      __syncthreads();
      r_ab_c(3, l);
      line_fwd<T, N, Q>(phi, l, DB);
      wb();
      w_ac_q(4, B2);
      __syncthreads();
      r_ab_c(4, l);
      line_fwd<T, N, Q>(phi, l, DA);
      line_pointwise<T, OP, N, Q>(G, detJw, tab, cell, tab_, DA, DB, DC, U);
      T D1[N], D2[N], D3[N];
      line_bwd<T, Q, N>(dph, DC, D3);
      if constexpr (MT)
        line_bwd<T, Q, N, true>(phi, U, D3);
      line_bwd<T, Q, N>(phi, DA, D1);
      line_bwd<T, Q, N>(phi, DB, D2);
      T E1[N], E2[N], lq[Q];
      wb();
      w_ab(5, D3);
      __syncthreads();
      r_ac_qb(5, lq);
      line_bwd<T, Q, N>(phi, lq, E2);
      wb();
      w_ab(6, D1);
      __syncthreads();
      r_ac_qb(6, lq);
      line_bwd<T, Q, N>(phi, lq, E1);
      wb();
      w_ab(7, D2);
      __syncthreads();
      r_ac_qb(7, lq);
      line_bwd<T, Q, N, true>(dph, lq, E2);
      wb();
      w_ac_b(8, E1);
      __syncthreads();
      r_bc_qa(8, lq);
      line_bwd<T, Q, N>(dph, lq, W);
      wb();
      w_ac_b(9, E2);
      __syncthreads();
      r_bc_qa(9, lq);
      line_bwd<T, Q, N, true>(phi, lq, W);
    }
    else if constexpr (FULL)
    {
      // transposes 0 A0, 1 A1, 2 B0, 3 B1, 4 B2, 5 D3, 6 D1, 7 D2, 8 E1,
      // 9 E2
      T B0[Q], B1[Q], B2[Q], l[N];
      w_bc(0, A0);
      __syncthreads();
      w_bc(1, A1);
      r_ac_b(0, l); // F_b on A0
      line_fwd<T, N, Q>(phi, l, B0);
      line_fwd<T, N, Q>(dph, l, B1);
      __syncthreads();
      w_ac_q(2, B0);
      r_ac_b(1, l); // F_b on A1
      line_fwd<T, N, Q>(phi, l, B2);
      __syncthreads();
      w_ac_q(3, B1);
      // F_c: (qa,qb) contracts c-lines of B0, B1, B2
      T U[MT ? Q : 1], DA[Q], DB[Q], DC[Q];
      r_ab_c(2, l);
      if constexpr (MT)
        line_fwd<T, N, Q>(phi, l, U);
      line_fwd<T, N, Q>(dph, l, DC);
      __syncthreads();
      w_ac_q(4, B2);
      r_ab_c(3, l);
      line_fwd<T, N, Q>(phi, l, DB);
      __syncthreads();
      r_ab_c(4, l);
      line_fwd<T, N, Q>(phi, l, DA);
      line_pointwise<T, OP, N, Q>(G, detJw, tab, cell, tab_, DA, DB, DC, U);
      // B_c along c, each result to LDS as soon as it is complete
      T D1[N], D2[N], D3[N];
      line_bwd<T, Q, N>(dph, DC, D3);
      if constexpr (MT)
        line_bwd<T, Q, N, true>(phi, U, D3);
      w_ab(5, D3);
      line_bwd<T, Q, N>(phi, DA, D1);
      __syncthreads();
      w_ab(6, D1);
      // B_b: (qa,c') contracts qb-lines of D3, D1, D2
      T E1[N], E2[N], lq[Q];
      r_ac_qb(5, lq);
      line_bwd<T, Q, N>(phi, lq, E2);
      line_bwd<T, Q, N>(phi, DB, D2);
      __syncthreads();
      w_ab(7, D2);
      r_ac_qb(6, lq);
      line_bwd<T, Q, N>(phi, lq, E1);
      __syncthreads();
      w_ac_b(8, E1);
      r_ac_qb(7, lq);
      line_bwd<T, Q, N, true>(dph, lq, E2);
      __syncthreads();
      w_ac_b(9, E2);
      // B_a: (b',c') contracts qa-lines of E1, E2
      r_bc_qa(8, lq);
      line_bwd<T, Q, N>(dph, lq, W);
      __syncthreads();
      r_bc_qa(9, lq);
      line_bwd<T, Q, N, true>(phi, lq, W);
    }
    else
    {
      // transposes 0 A0, 1 B0, 2 D3, 3 E2; each write already follows the
      // reads of the transpose before
      T B0[Q], U[Q], D3[N], E2[N], l[N], lq[Q];
      wb();
      w_bc(0, A0);
      __syncthreads();
      r_ac_b(0, l);
      line_fwd<T, N, Q>(phi, l, B0);
      wb();
      w_ac_q(1, B0);
      __syncthreads();
      r_ab_c(1, l);
      line_fwd<T, N, Q>(phi, l, U);
      T none[1];
      line_pointwise<T, OP, N, Q>(G, detJw, tab, cell, tab_, none, none, none,
                                  U);
      line_bwd<T, Q, N>(phi, U, D3);
      wb();
      w_ab(2, D3);
      __syncthreads();
      r_ac_qb(2, lq);
      line_bwd<T, Q, N>(phi, lq, E2);
      wb();
      w_ac_b(3, E2);
      __syncthreads();
      r_bc_qa(3, lq);
      line_bwd<T, Q, N>(phi, lq, W);
    }
    // keep the products here rather than sunk into both store paths
#pragma unroll
    for (int a = 0; a < N; ++a)
      asm volatile("" : "+v"(W[a]));
    if (own_bc && live)
    {
      // the store offsets are computed here, not kept from the loads
      int ts = tbc;
      asm volatile("" : "+v"(ts));
      if (dc)
      {
#pragma unroll
        for (int a = 0; a < N; ++a)
          unsafeAtomicAdd(y + dc[(unsigned)(a * NN + ts)], W[a]);
      }
      else
      {
        T* const yc = y + cb;
#pragma unroll
        for (int a = 0; a < N; ++a)
          yc[(unsigned)(a * NN + ts)] = W[a];
      }
    }
  }
}

// Launch for one (N, Q): the op and dtype from the request.
template <typename T, int OP, int N, int Q>
int line_launch_op(const ApplyReq& r, int max_lds)
{
  constexpr size_t bytes = line_lds_bytes<T, OP, N, Q>();
  if (bytes > (size_t)max_lds)
    return FEMTC_ERR_INVALID_ARGUMENT;
  constexpr int CPB = line_cpb<T, OP, N, Q>();
  const int64_t grid64 =
      std::min<int64_t>((r.num_cells + CPB - 1) / CPB, 16384);
  hipLaunchKernelGGL((line_cells<T, OP, N, Q>), dim3((unsigned)grid64),
                     dim3(line_threads<T, OP, N, Q>()), bytes, r.stream,
                     (const T*)r.tab, (const T*)r.x, (T*)r.y, (const T*)r.G,
                     (const T*)r.detJw, r.gather ? r.dofmap : nullptr,
                     r.num_cells);
  return hipGetLastError() == hipSuccess ? FEMTC_OK : FEMTC_ERR_HIP;
}

template <typename T, int OP, int N, int Q>
int line_launch_fit(const ApplyReq& r, int max_lds)
{
  if constexpr (line_fits<T, OP, N, Q>())
    return line_launch_op<T, OP, N, Q>(r, max_lds);
  else
    return FEMTC_ERR_INVALID_ARGUMENT;
}

template <typename T, int N, int Q>
int line_launch_q(const ApplyReq& r, int max_lds)
{
  switch (r.op)
  {
  case OP_MASS: return line_launch_fit<T, OP_MASS, N, Q>(r, max_lds);
  case OP_STIFF: return line_launch_fit<T, OP_STIFF, N, Q>(r, max_lds);
  case OP_KPM: return line_launch_fit<T, OP_KPM, N, Q>(r, max_lds);
  case OP_MASS_AFF:
    return line_launch_fit<T, OP_MASS_AFF, N, Q>(r, max_lds);
  case OP_STIFF_AFF:
    return line_launch_fit<T, OP_STIFF_AFF, N, Q>(r, max_lds);
  case OP_KPM_AFF: return line_launch_fit<T, OP_KPM_AFF, N, Q>(r, max_lds);
  }
  return FEMTC_ERR_INVALID_ARGUMENT;
}

template <int N>
int line_dispatch(const ApplyReq& r, int max_lds)
{
  if (r.dtype == FEMTC_F64)
  {
    if (r.Q == N)
      return line_launch_q<double, N, N>(r, max_lds);
    if (r.Q == N + 1)
      return line_launch_q<double, N, N + 1>(r, max_lds);
    if (r.Q == N + 2)
      return line_launch_q<double, N, N + 2>(r, max_lds);
  }
  else if (r.dtype == FEMTC_F32)
  {
    if (r.Q == N)
      return line_launch_q<float, N, N>(r, max_lds);
    if (r.Q == N + 1)
      return line_launch_q<float, N, N + 1>(r, max_lds);
    if (r.Q == N + 2)
      return line_launch_q<float, N, N + 2>(r, max_lds);
  }
  return FEMTC_ERR_INVALID_ARGUMENT;
}

} // namespace femtc
