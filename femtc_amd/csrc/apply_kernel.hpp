// Fused sum-factorized cell operator for gfx950.
//
// Work decomposition.  Each cell gets Q*Q threads; a thread owns one "line"
// of a 3-D cell tensor and performs the 1-D contraction along that line in
// registers, with the table entries as wave-uniform (scalar) operands.
// Between contractions along different axes the partial tensors pass through
// LDS.  A workgroup is one wave holding CPB = 64/(Q*Q) cells (Q <= 8), or the
// fewest waves that hold one cell (Q = 10, 11) or three (Q = 9: 243 of 256
// lanes), so the stage barriers are cheap and the workgroups resident on a
// CU progress independently, hiding each other's memory latency.  The FP32
// cell pass of the assembled scatter at Q = 6, 7 takes 4 waves of 7 and 5
// cells instead (assembled_cells), for busy lanes.
//
// Memory access.  Every stream is read with as few, as wide and as
// contiguous wave instructions as possible (the vector-memory address path
// is the first limit this kernel hit):
//  * the geometry G[cell][q][6] of the workgroup's cells is one contiguous
//    block; it is read at the start of a step as lane-contiguous 16-byte
//    pieces (8-byte for FP32), kept in registers while the forward
//    contractions run, and at the quadrature stage written to LDS (over the
//    then-idle work buffers) and read back per point;
//  * the gather and the scatter run in thread (a, c), so for a fixed b the
//    lanes touch N runs of N consecutive global dofs.
//
// Axis order.  The forward chain contracts b, then c, then a, so the
// quadrature-point stage runs in thread (qb, qc), which owns the points
// (0..Q-1, qb, qc).
//
// Contraction schedule (forward tables read [dof][q]; backward tables are the
// transpose, or the table as stored in reference mode):
//   F1  axis b  lines (a,c)   : A_p = phi_b U,    A_g = g_b U
//   F2  axis c  lines (qb,a)  : B_pp = phi_c A_p, B_pg = g_c A_p,
//                               B_gp = phi_c A_g
//   F3  axis a  lines (qb,qc) : u = phi_a B_pp,   d_a = g_a B_pp,
//                               d_c = phi_a B_pg, d_b = phi_a B_gp
//   pointwise                 : (v_a,v_b,v_c) = S(q) (d_a,d_b,d_c), v_m = detJw u
//   B3  axis a  lines (qb,qc) : R_a = g_a' v_a + phi_a' v_m,
//                               R_b = phi_a' v_b, R_c = phi_a' v_c
//   B2  axis c  lines (qb,a)  : S_0 = phi_c' R_a + g_c' R_c, S_b = phi_c' R_b
//   B1  axis b  lines (a,c)   : W = phi_b' S_0 + g_b' S_b
// which is W = sum_ij B_i^T S_ij B_j U (+ B^T detJw B U), B_a = g(x)phi(x)phi
// etc.: the contraction chain of src/stiffness.hpp:68-97 with the intended
// transposes and sums.  The reference's literal variant (OP_STIFF_LIT) keeps
// W = chain(phi,phi,g)(G2 d_a + G1 d_b + G0 u) (src/stiffness.hpp:76-97).
//
// LDS layout.  X[qb][c][a] (F1 -> F2, B2 -> B1) and Y[a][qb][qc] (F2 -> F3,
// B3 -> B2) with the padded strides of lds_layout.hpp, which keep both lane
// patterns of each tensor (nearly) free of bank conflicts.  X and Y are never
// live together, so they share one area (a barrier separates the last read of
// one from the first write of the other), and the geometry passes through the
// same area in two halves of qa-planes at the quadrature stage.  This keeps a
// workgroup's LDS small enough that registers, not LDS, bound occupancy.
//
// Table access.  Every contraction walks the table one dof row at a time
// (forward: acc[q] += T[d][q] in[d]; backward: out[d] = sum_q T[d][q] in[q]),
// and each row is fetched with scalar loads pinned inside its iteration, so
// only a row or two of each table is live in SGPRs instead of the whole
// 2 N Q table (which does not fit and would spill).  The EO instances take
// the even-odd table set of even_odd.hpp instead and contract each line as
// two half-size products.
#pragma once

#include "assemble.hpp"
#include "common.hpp"
#include "even_odd.hpp"
#include "lds_layout.hpp"

namespace femtc
{

template <int Q>
constexpr int cells_per_block()
{
  return direct_cells(Q);
}

template <int Q>
constexpr int block_threads()
{
  return ((cells_per_block<Q>() * Q * Q + 63) / 64) * 64;
}

// Cells and threads per workgroup of an instance with scatter kind BW.
template <typename T, int Q, int BW>
constexpr int cells_per_group()
{
  return BW == SCATTER_ASSEMBLED ? assembled_cells(Q, (int)sizeof(T))
                                 : cells_per_block<Q>();
}

template <typename T, int Q, int BW>
constexpr int group_threads()
{
  return ((cells_per_group<T, Q, BW>() * Q * Q + 63) / 64) * 64;
}

// A geometry piece: two elements, 16 bytes for FP64 and 8 for FP32.  A
// cell's G block (6 Q^3 elements, always a multiple of two) is 16-byte
// aligned for FP64 and 8-byte aligned for FP32.
template <typename T>
using Piece = T __attribute__((ext_vector_type(2)));

__device__ inline void atomic_add(double* p, double v)
{
  unsafeAtomicAdd(p, v);
}
__device__ inline void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }

// Atomic add into LDS (ds_add_f64 / ds_add_f32): lanes of one instruction
// that hit the same address all count.
template <typename T>
__device__ inline void lds_add(T* p, T v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T>
__device__ inline T stream_load(const T* p)
{
  return __builtin_nontemporal_load(p);
}

// Row d of a [rows][LEN] table, loaded at this point of the program: the
// offset goes through an empty asm so the compiler cannot hoist the (scalar)
// loads of every row to the top of the kernel.  It is an unsigned byte
// offset, which the scalar loads take as their SGPR offset (with the
// element's immediate offset): no 64-bit address arithmetic per row.
template <int LEN, typename T>
__device__ inline void load_row(const T* __restrict__ tab, int off, T (&row)[LEN])
{
  unsigned boff = (unsigned)off * (unsigned)sizeof(T);
  asm volatile("" : "+s"(boff));
  const T* p =
      reinterpret_cast<const T*>(reinterpret_cast<const char*>(tab) + boff);
#pragma unroll
  for (int i = 0; i < LEN; ++i)
    row[i] = p[i];
}

// Column d of a table stored [Q][N] (reference mode backward, Q == N).
template <int LEN, int STRIDE, typename T>
__device__ inline void load_col(const T* __restrict__ tab, int off, T (&row)[LEN])
{
  unsigned boff = (unsigned)off * (unsigned)sizeof(T);
  asm volatile("" : "+s"(boff));
  const T* p =
      reinterpret_cast<const T*>(reinterpret_cast<const char*>(tab) + boff);
#pragma unroll
  for (int i = 0; i < LEN; ++i)
    row[i] = p[i * STRIDE];
}

template <typename T, int N, int Q, bool REF>
__device__ inline void load_bwd_row(const T* __restrict__ tab, int d, T (&row)[Q])
{
  if constexpr (REF)
    load_col<Q, N>(tab, d, row);
  else
    load_row<Q>(tab, d * Q, row);
}

// ---- even-odd contractions (table set of even_odd.hpp) --------------------

// A padded even-odd table row (LEN a multiple of 4, 16-element aligned
// offset) as LEN/4 four-element loads: one scalar load of 4 (FP32) or 8
// (FP64) dwords each.
template <int LEN, typename T>
__device__ inline void load_row4(const T* __restrict__ tab, int off, T (&row)[LEN])
{
  static_assert(LEN % 4 == 0, "padded row");
  typedef T V4 __attribute__((ext_vector_type(4)));
  unsigned boff = (unsigned)off * (unsigned)sizeof(T);
  asm volatile("" : "+s"(boff));
  const V4* p =
      reinterpret_cast<const V4*>(reinterpret_cast<const char*>(tab) + boff);
#pragma unroll
  for (int i = 0; i < LEN / 4; ++i)
  {
    const V4 v = p[i];
    row[4 * i] = v.x;
    row[4 * i + 1] = v.y;
    row[4 * i + 2] = v.z;
    row[4 * i + 3] = v.w;
  }
}

// e[k] = v[k] + v[K-1-k], o[k] = v[k] - v[K-1-k], e[K/2] = v[K/2] (odd K)
template <int K, typename T>
__device__ inline void eo_split(const T (&v)[K], T (&e)[(K + 1) / 2],
                                T (&o)[K / 2])
{
#pragma unroll
  for (int k = 0; k < K / 2; ++k)
  {
    e[k] = v[k] + v[K - 1 - k];
    o[k] = v[k] - v[K - 1 - k];
  }
  if constexpr (K % 2)
    e[K / 2] = v[K / 2];
}

// One table term (term dims of eo_dims(K, M, S), at element offset off):
// S = +1: ev += A e, od += B o;  S = -1: ev += A o, od += B e.
// INIT: the first term into ev/od (products instead of sums).
template <int K, int M, int S, bool INIT, typename T>
__device__ inline void eo_term(const T* __restrict__ tab, int off,
                               const T (&e)[(K + 1) / 2], const T (&o)[K / 2],
                               T (&ev)[(M + 1) / 2], T (&od)[M / 2])
{
  constexpr EoDims d = eo_dims(K, M, S);
#pragma unroll
  for (int k = 0; k < d.rowsA; ++k)
  {
    T row[d.RA];
    load_row4<d.RA>(tab, off + k * d.RA, row);
    T in;
    if constexpr (S > 0)
      in = e[k];
    else
      in = o[k];
#pragma unroll
    for (int m = 0; m < d.Mp; ++m)
      ev[m] = (INIT && k == 0) ? row[m] * in : fma(row[m], in, ev[m]);
  }
#pragma unroll
  for (int k = 0; k < d.rowsB; ++k)
  {
    T row[d.RB];
    load_row4<d.RB>(tab, off + d.rowsA * d.RA + k * d.RB, row);
    T in;
    if constexpr (S > 0)
      in = o[k];
    else
      in = e[k];
#pragma unroll
    for (int m = 0; m < d.Hm; ++m)
      od[m] = (INIT && k == 0) ? row[m] * in : fma(row[m], in, od[m]);
  }
}

// out[m] = ev[m] + od[m], out[M-1-m] = ev[m] - od[m], out[M/2] = ev[M/2]
template <int M, typename T>
__device__ inline void eo_join(const T (&ev)[(M + 1) / 2], const T (&od)[M / 2],
                               T (&out)[M])
{
#pragma unroll
  for (int m = 0; m < M / 2; ++m)
  {
    out[m] = ev[m] + od[m];
    out[M - 1 - m] = ev[m] - od[m];
  }
  if constexpr (M % 2)
    out[M / 2] = ev[M / 2];
}

// The geometry block is staged through LDS in two halves of qa-planes:
// planes [0, QA1) and [QA1, Q).
template <int Q>
constexpr int geometry_split()
{
  return (Q + 1) / 2;
}

// Lane-contiguous pieces of the qa-planes [QB0, QB0+QH) of the geometry of
// the block's cells (cell0 .. cell0+ncells_blk-1), as they are laid out in
// LDS: [cell][plane][qb][qc][6].
template <typename T, int Q, int BT, int QB0, int QH, int GPTH>
__device__ inline void load_geometry_half(const T* __restrict__ G,
                                          long long cell0,
                                          long long ncells_blk, int tid,
                                          Piece<T> (&gp)[GPTH])
{
  constexpr int Q3 = Q * Q * Q;
  constexpr int PER_CELL = 3 * QH * Q * Q; // pieces of one cell's half
  const Piece<T>* Gb = reinterpret_cast<const Piece<T>*>(G + cell0 * Q3 * 6) +
                       QB0 * 3 * Q * Q;
#pragma unroll
  for (int k = 0; k < GPTH; ++k)
  {
    const int p = k * BT + tid;
    const int c = p / PER_CELL;
    if (c < ncells_blk)
      gp[k] = stream_load(Gb + c * 3 * Q3 + (p - c * PER_CELL));
  }
}

// LDS elements of one block: the work tensors X and Y share one area (they
// are never live together), and a half of the geometry block aliases it at
// the quadrature stage.
template <typename T, int N, int Q, int OP, int BW>
constexpr int work_lds_elems()
{
  constexpr bool AFF = OP >= OP_MASS_AFF;
  constexpr int OPB = AFF ? OP - OP_MASS_AFF : OP;
  constexpr bool GRAD = OPB != OP_MASS;
  constexpr bool GEO = GRAD && !AFF;
  constexpr int CPB = cells_per_group<T, Q, BW>();
  constexpr LdsLayout LY = lds_layout<(int)sizeof(T), N, Q>();
  // per-cell strides: the tuned ones where the table has them, else the
  // cell's own footprint
  constexpr int XCELLS =
      CPB > 1 && LY.XCELL ? LY.XCELL : (GRAD ? 2 : 1) * Q * LY.XQ;
  constexpr int YCELLS =
      CPB > 1 && LY.YCELL ? LY.YCELL : (GRAD ? 3 : 1) * N * LY.YA;
  constexpr int WCELL = XCELLS > YCELLS ? XCELLS : YCELLS;
  constexpr int WORK = CPB * WCELL;
  constexpr int GELEM = GEO ? 6 * geometry_split<Q>() * Q * Q * CPB : 0;
  return WORK > GELEM ? WORK : GELEM;
}

// Synchronisation between the contraction stages.  A one-wave workgroup (or
// a multi-wave one whose cells span waves) uses the workgroup barrier; in
// the blocked scatter every wave owns whole cells and its own LDS work area,
// so a stage only has to order the wave's own LDS accesses: LDS operations of
// one wave execute in issue order, and the fences plus wave barrier keep the
// compiler from moving accesses across the stage boundary.
template <bool WAVE_LOCAL>
__device__ inline void stage_sync()
{
  if constexpr (WAVE_LOCAL)
  {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  else
    __syncthreads();
}

// Bytes of dynamic LDS of a blocked-scatter workgroup: BW per-wave work
// areas, then the block's accumulator, global-index and count arrays.
template <typename T, int N, int Q, int OP>
constexpr size_t blocked_work_bytes()
{
  return sizeof(Piece<T>) *
         (size_t)((work_lds_elems<T, N, Q, OP, 1>() + 1) / 2);
}

inline size_t blocked_box(int P, int cbx, int cby, int cbz)
{
  return (size_t)(cbx * P + 1) * (cby * P + 1) * (cbz * P + 1);
}

template <typename T>
inline size_t blocked_box_bytes(size_t box)
{
  return box * (sizeof(T) + 2 * sizeof(int32_t));
}

// Run length the z-run scatter aims for: consecutive global dofs along z
// per atomic run.  Measured on MI355X, FP64 atomics at the memory side
// sustain 61 G/s in runs of 5 consecutive addresses and 141 G/s in runs of
// 33.
constexpr int ZRUN_TARGET = 32;
// LDS of a z-run workgroup at most: four workgroups per CU.  At P=4 (FP64
// stiffness, 64^3) runs of 17, 25 and 33 dofs took 396, 392 and 381 us per
// apply; at P=6 (48^3) runs of 19 took 497 us and the 31-dof segments of a
// 64 KiB workgroup (two per CU) 915 us.
constexpr size_t ZRUN_LDS = 40 * 1024;

// Waves per workgroup of the z-run scatter, each taking CPB cells of a
// segment along z: enough that a segment's run (cells * P + 1 dofs) spans
// ZRUN_TARGET values, as long as the waves' work areas and the segment's
// dof box fit ZRUN_LDS.
template <typename T, int N, int Q, int OP>
constexpr int zrun_waves()
{
  constexpr int CPB = cells_per_block<Q>();
  constexpr int want = (ZRUN_TARGET + (N - 1) - 1) / (N - 1); // cells
  int w = (want + CPB - 1) / CPB;
  while (w > 1 && w * blocked_work_bytes<T, N, Q, OP>() +
                          sizeof(T) * N * N * ((size_t)w * CPB * (N - 1) + 1) >
                      ZRUN_LDS)
    --w;
  return w;
}

// Cells per z-segment of the z-run scatter.
template <typename T, int N, int Q, int OP>
constexpr int zrun_cells()
{
  return zrun_waves<T, N, Q, OP>() * cells_per_block<Q>();
}

// The work of one workgroup (block index blk).
// BW = 0: CPB consecutive cells, direct scatter.
// BW > 0: blocked scatter, BW waves per workgroup and one block of
// cbx x cby x cbz cells per workgroup (Q <= 8 and cbz % CPB == 0).
// BW = SCATTER_ASSEMBLED: the cell pass of the assembled scatter; every
// cell-dof value is stored (no atomics) into the lattice-ordered buffer y of
// lattice_index(), and a second pass (assemble.hpp) sums the values of each
// global dof.  Cells are numbered (ci*ncy + cj)*ncz + ck.
// EO: tab is the even-odd table set (even_odd.hpp) of symmetric tables.
template <typename T, int N, int Q, int OP, bool REF, bool GATHER, int BW,
          bool EO = false>
__device__ __forceinline__ void
apply_body(const T* __restrict__ tab, const T* __restrict__ x,
           T* __restrict__ y, const T* __restrict__ G,
           const T* __restrict__ detJw, const int32_t* __restrict__ dofmap,
           long long num_cells, const BlockArgs& ba, long long blk,
           double* __restrict__ epart = nullptr, T* __restrict__ zbox = nullptr,
           long long zcell0 = 0, int zncell = 0)
{
  constexpr bool BLK = BW > 0;
  constexpr bool ASM = BW == SCATTER_ASSEMBLED;
  constexpr bool STR = BW == SCATTER_STRUCT;
  constexpr bool ZR = BW == SCATTER_ZRUN;
  constexpr bool CLOSED = ASM || STR || ZR; // gather indices in closed form
  // multi-wave workgroups whose waves own whole cells and their own LDS
  // work area (blocked and z-run scatters)
  constexpr bool WL = BLK || ZR;
  static_assert(!WL || (GATHER && Q * Q <= 64), "blocked scatter layout");
  static_assert(!CLOSED || (GATHER && !REF), "structured scatters");
  constexpr int CPB = cells_per_group<T, Q, BW>();
  constexpr int BT = WL ? 64 : group_threads<T, Q, BW>(); // threads per group
  constexpr int NL = Q * Q;
  constexpr bool AFF = OP >= OP_MASS_AFF; // geometry from 7 values per cell
  constexpr int OPB = AFF ? OP - OP_MASS_AFF : OP;
  constexpr bool LIT = (OPB == OP_STIFF_LIT);
  constexpr bool MASSTERM = (OPB == OP_MASS || OPB == OP_KPM);
  constexpr bool GRAD = (OPB != OP_MASS);
  constexpr bool FULL = (OPB == OP_STIFF || OPB == OP_KPM); // 3 gradients
  constexpr bool GEO = GRAD && !AFF; // geometry streamed per point
  static_assert(!AFF || (!LIT && !REF && !BLK), "affine geometry variants");
  static_assert(!EO || (!LIT && !REF && !WL), "even-odd variants");
  constexpr bool UQ = MASSTERM || LIT;
  // even-odd table set: term offsets, the weights, half lengths
  constexpr int EPF = eo_offset(N, Q, EO_PHI_F), EGF = eo_offset(N, Q, EO_DPHI_F);
  constexpr int EPB = eo_offset(N, Q, EO_PHI_B), EGB = eo_offset(N, Q, EO_DPHI_B);
  constexpr int WOFF = EO ? eo_offset(N, Q, EO_WEIGHTS) : 2 * N * Q;
  constexpr int NP = (N + 1) / 2, NH = N / 2, QP = (Q + 1) / 2, QH = Q / 2;
  constexpr int NX = GRAD ? 2 : 1;
  constexpr int NY = GRAD ? 3 : 1;
  // X[qb][c][a] and Y[a][qb][qc] with bank-conflict-minimising strides
  constexpr LdsLayout LY = lds_layout<(int)sizeof(T), N, Q>();
  constexpr int XQ = LY.XQ, XC = LY.XC, YQ = LY.YQ, YA = LY.YA;
  constexpr int SX = Q * XQ; // one X tensor
  constexpr int SY = N * YA; // one Y tensor
  constexpr int XCELLS = CPB > 1 && LY.XCELL ? LY.XCELL : NX * SX;
  constexpr int YCELLS = CPB > 1 && LY.YCELL ? LY.YCELL : NY * SY;
  static_assert(XCELLS >= NX * SX && YCELLS >= NY * SY, "layout fits");
  constexpr int NN = N * N;
  constexpr int N3 = N * N * N;
  constexpr int Q2 = Q * Q;
  constexpr int Q3 = Q * Q * Q;
  // geometry of the workgroup in pieces of two elements, in two halves of
  // qa-planes: [0, QA1) and [QA1, Q); GP0/GP1 pieces, GPT0/GPT1 per thread
  constexpr int QA1 = geometry_split<Q>();
  constexpr int GP0 = GEO ? 3 * QA1 * Q2 * CPB : 0;
  constexpr int GP1 = GEO ? 3 * (Q - QA1) * Q2 * CPB : 0;
  constexpr int GPT0 = (GP0 + BT - 1) / BT;
  constexpr int GPT1 = (GP1 + BT - 1) / BT;
  // LDS: X and Y share one area per cell (WCELL); a geometry half aliases
  // the whole area at the quadrature stage, when neither is live.
  constexpr int WCELL = XCELLS > YCELLS ? XCELLS : YCELLS;
  constexpr int LDSW = (work_lds_elems<T, N, Q, OP, BW>() + 1) / 2; // pieces
  // LDS slots
  constexpr int X_P = 0, X_G = 1;              // F1 -> F2
  constexpr int Y_PP = 0, Y_PG = 1, Y_GP = 2;  // F2 -> F3
  constexpr int Y_RA = 0, Y_RB = 1, Y_RC = 2;  // B3 -> B2
  constexpr int X_S0 = 0, X_SB = 1;            // B2 -> B1

  __shared__ Piece<T> Ls_static[WL ? 1 : LDSW];
  extern __shared__ __align__(16) unsigned char femtc_lds_dyn[];
  Piece<T>* Ls_dyn = reinterpret_cast<Piece<T>*>(femtc_lds_dyn);
  // this wave's (or, unblocked, the workgroup's) work area
  const int wave = WL ? (int)(threadIdx.x >> 6) : 0;
  Piece<T>* Ls = WL ? Ls_dyn + wave * LDSW : Ls_static;
  T* L = reinterpret_cast<T*>(Ls);

  const T* __restrict__ ph = tab;
  const T* __restrict__ gr = tab + N * Q;

  const int tid = WL ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
  const int lc = tid / NL;
  const int ln = tid - lc * NL;
  const bool lane_ok = lc < CPB;
  T* Xc = L + (lane_ok ? lc : 0) * WCELL;
  T* Yc = Xc;

  // blocked scatter: the block's accumulator, global index and count of
  // every dof of its (cbx P+1)(cby P+1)(cbz P+1) box
  constexpr int P = N - 1;
  T* acc = nullptr;
  int32_t* gid = nullptr;
  int32_t* cnt = nullptr;
  int BY = 0, BZ = 0, BOX = 0;
  long long blk_ci = 0, blk_cj = 0, blk_ck = 0; // the block's first cell
  if constexpr (BLK)
  {
    BY = ba.cby * P + 1;
    BZ = ba.cbz * P + 1;
    BOX = (ba.cbx * P + 1) * BY * BZ;
    acc = reinterpret_cast<T*>(Ls_dyn + (blockDim.x >> 6) * LDSW);
    gid = reinterpret_cast<int32_t*>(acc + BOX);
    cnt = gid + BOX;
    for (int s = threadIdx.x; s < BOX; s += blockDim.x)
    {
      acc[s] = T(0);
      cnt[s] = 0;
    }
    const long long b = blk;
    const long long bk = b % ba.nbz;
    const long long bj = (b / ba.nbz) % ba.nby;
    const long long bi = b / ((long long)ba.nbz * ba.nby);
    blk_ci = bi * ba.cbx;
    blk_cj = bj * ba.cby;
    blk_ck = bk * ba.cbz;
  }

  // thread roles
  const int r1a = ln / N, r1c = ln - (ln / N) * N;  // F1/B1: (a, c)
  const int r2q = ln / N, r2a = ln - (ln / N) * N;  // F2/B2: (qb, a)
  const int yo3 = (ln / Q) * YQ + (ln - (ln / Q) * Q); // F3/B3: (qb, qc)
  const bool role1 = lane_ok && ln < NN;
  const bool role2 = lane_ok && ln < N * Q;
  const int uoff = r1a * NN + r1c; // + b*N within a cell

  {
    // the CPB consecutive cells of this wave (blocked) or workgroup
    long long cell0 = 0, ncells_blk = 0;
    int bl_i = 0, bl_j = 0, bl_k = 0; // blocked: local cell of lane group 0
    if constexpr (BLK)
    {
      const int lcb0 = wave * CPB;
      bl_k = lcb0 % ba.cbz;
      bl_j = (lcb0 / ba.cbz) % ba.cby;
      bl_i = lcb0 / (ba.cbz * ba.cby);
      const long long ci = blk_ci + bl_i, cj = blk_cj + bl_j,
                      ck0 = blk_ck + bl_k;
      const long long left = ba.ncz - ck0;
      if (lcb0 < ba.cbx * ba.cby * ba.cbz && ci < ba.ncx && cj < ba.ncy &&
          left > 0)
      {
        ncells_blk = left < CPB ? left : (long long)CPB;
        cell0 = (ci * ba.ncy + cj) * ba.ncz + ck0;
      }
    }
    else if constexpr (ZR)
    {
      // this wave's CPB cells of the segment's zncell cells from zcell0
      cell0 = zcell0 + wave * CPB;
      const long long left = zcell0 + zncell - cell0;
      ncells_blk = left < CPB ? (left > 0 ? left : 0) : (long long)CPB;
    }
    else
    {
      cell0 = blk * CPB;
      ncells_blk = num_cells - cell0 < CPB ? num_cells - cell0 : (long long)CPB;
    }
    // one cell per group (Q = 6..8, 10, 11): the cell is wave-uniform, so
    // its index arithmetic (64-bit products and divisions) runs on the
    // scalar unit
    const long long cell = CPB == 1 ? cell0 : cell0 + lc;
    const bool live = lane_ok && lc < ncells_blk;

    // ---------------- global loads issued up front ------------------------
    T u[N];
    int32_t idx[GATHER && !CLOSED ? N : 1];
    // structured scatters: the dofmap is the structured numbering (verified
    // by the caller), so the gather (and STR's scatter) indices are
    // sbase + b * sstep, computed, not loaded
    long long sbase = 0, sstep = 0;
    if (live && role1)
    {
      if constexpr (CLOSED)
      {
        // 32-bit division: the launchers keep num_cells below 2^31
        const unsigned cu = (unsigned)cell, t = cu / (unsigned)ba.ncz;
        const long long ck = cu - t * (unsigned)ba.ncz;
        const unsigned ti = t / (unsigned)ba.ncy;
        const long long cj = t - ti * (unsigned)ba.ncy, ci = ti;
        constexpr int P1 = N - 1;
        sstep = (long long)ba.ncz * P1 + 1;
        sbase = ((ci * P1 + r1a) * ((long long)ba.ncy * P1 + 1) + cj * P1) *
                    sstep +
                ck * P1 + r1c;
      }
#pragma unroll
      for (int b = 0; b < N; ++b)
      {
        if constexpr (CLOSED)
          u[b] = x[sbase + b * sstep];
        else if constexpr (GATHER)
        {
          idx[b] = dofmap[cell * N3 + uoff + b * N];
          u[b] = x[idx[b]];
        }
        else
          u[b] = x[cell * N3 + uoff + b * N];
      }
    }
    Piece<T> gp0[GEO ? GPT0 : 1], gp1[GEO ? GPT1 : 1];
    if constexpr (GEO)
    {
      load_geometry_half<T, Q, BT, 0, QA1>(G, cell0, ncells_blk, tid, gp0);
      load_geometry_half<T, Q, BT, QA1, Q - QA1>(G, cell0, ncells_blk, tid,
                                                  gp1);
    }
    // affine cells: S = detJ J^-1 J^-T (packed as G) and detJ of the cell,
    // and w(qb) w(qc) of this thread's quadrature line
    T cg[AFF ? CELLGEO : 1];
    T wbc = T(1);
    if constexpr (AFF)
    {
      if (live)
      {
#pragma unroll
        for (int k = 0; k < CELLGEO; ++k)
          cg[k] = G[cell * CELLGEO + k];
        const T* wq = tab + WOFF;
        wbc = wq[ln / Q] * wq[ln - (ln / Q) * Q];
      }
    }
    T dw[MASSTERM ? Q : 1];
    if constexpr (MASSTERM && !AFF)
    {
      if (live)
      {
        const T* Dp = detJw + cell * Q3 + ln;
#pragma unroll
        for (int qa = 0; qa < Q; ++qa)
          dw[qa] = stream_load(Dp + qa * Q2);
      }
    }

    // blocked: record which global dof sits in each box slot (every cell
    // touching a slot writes the same index when the dofmap is the
    // structured one; otherwise the scatter below falls back to atomics)
    int slot0 = 0; // box slot of (a, 0, c) of this thread's cell
    if constexpr (BLK)
    {
      slot0 = ((bl_i * P + r1a) * BY + bl_j * P) * BZ + (bl_k + lc) * P + r1c;
      if (live && role1)
      {
#pragma unroll
        for (int b = 0; b < N; ++b)
          gid[slot0 + b * BZ] = idx[b];
      }
      __syncthreads(); // box cleared and indexed for the whole block
    }

    // ---------------- F1: contract axis b ---------------------------------
    if (live && role1)
    {
      T ap[Q], ag[Q];
      if constexpr (EO)
      {
        T ue[NP], uo[NH], ev[QP], od[QH];
        eo_split<N>(u, ue, uo);
        eo_term<N, Q, 1, true>(tab, EPF, ue, uo, ev, od);
        eo_join<Q>(ev, od, ap);
        if constexpr (GRAD)
        {
          eo_term<N, Q, -1, true>(tab, EGF, ue, uo, ev, od);
          eo_join<Q>(ev, od, ag);
        }
      }
      else
      {
#pragma unroll
        for (int b = 0; b < N; ++b)
        {
          T rp[Q], rg[Q];
          load_row<Q>(ph, b * Q, rp);
          if constexpr (GRAD)
            load_row<Q>(gr, b * Q, rg);
#pragma unroll
          for (int q = 0; q < Q; ++q)
          {
            ap[q] = b == 0 ? rp[q] * u[b] : fma(rp[q], u[b], ap[q]);
            if constexpr (GRAD)
              ag[q] = b == 0 ? rg[q] * u[b] : fma(rg[q], u[b], ag[q]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < Q; ++q)
      {
        Xc[X_P * SX + q * XQ + r1c * XC + r1a] = ap[q];
        if constexpr (GRAD)
          Xc[X_G * SX + q * XQ + r1c * XC + r1a] = ag[q];
      }
    }
    stage_sync<WL>();

    // ---------------- F2: contract axis c ---------------------------------
    T bpp[Q], bpg[Q], bgp[Q];
    if (live && role2)
    {
      T inp[N], ing[N];
#pragma unroll
      for (int c = 0; c < N; ++c)
      {
        inp[c] = Xc[X_P * SX + r2q * XQ + c * XC + r2a];
        if constexpr (GRAD)
          ing[c] = Xc[X_G * SX + r2q * XQ + c * XC + r2a];
      }
      if constexpr (EO)
      {
        T pe[NP], po[NH], ev[QP], od[QH];
        eo_split<N>(inp, pe, po);
        eo_term<N, Q, 1, true>(tab, EPF, pe, po, ev, od);
        eo_join<Q>(ev, od, bpp);
        if constexpr (GRAD)
        {
          eo_term<N, Q, -1, true>(tab, EGF, pe, po, ev, od);
          eo_join<Q>(ev, od, bpg);
          T ge[NP], go[NH];
          eo_split<N>(ing, ge, go);
          eo_term<N, Q, 1, true>(tab, EPF, ge, go, ev, od);
          eo_join<Q>(ev, od, bgp);
        }
      }
      else
      {
#pragma unroll
        for (int c = 0; c < N; ++c)
        {
          T rp[Q], rg[Q];
          load_row<Q>(ph, c * Q, rp);
          if constexpr (GRAD)
            load_row<Q>(gr, c * Q, rg);
#pragma unroll
          for (int q = 0; q < Q; ++q)
          {
            bpp[q] = c == 0 ? rp[q] * inp[c] : fma(rp[q], inp[c], bpp[q]);
            if constexpr (GRAD)
            {
              bpg[q] = c == 0 ? rg[q] * inp[c] : fma(rg[q], inp[c], bpg[q]);
              bgp[q] = c == 0 ? rp[q] * ing[c] : fma(rp[q], ing[c], bgp[q]);
            }
          }
        }
      }
    }
    stage_sync<WL>(); // X is read; Y overwrites the same LDS
    if (live && role2)
    {
      const int o = r2a * YA + r2q * YQ;
#pragma unroll
      for (int q = 0; q < Q; ++q)
      {
        Yc[Y_PP * SY + o + q] = bpp[q];
        if constexpr (GRAD)
        {
          Yc[Y_PG * SY + o + q] = bpg[q];
          Yc[Y_GP * SY + o + q] = bgp[q];
        }
      }
    }
    stage_sync<WL>();

    // ---------------- F3 + quadrature-point work + B3 ---------------------
    T ipp[N], ipg[N], igp[N];
    if (live)
    {
#pragma unroll
      for (int a = 0; a < N; ++a)
      {
        ipp[a] = Yc[Y_PP * SY + a * YA + yo3];
        if constexpr (GRAD)
        {
          ipg[a] = Yc[Y_PG * SY + a * YA + yo3];
          igp[a] = Yc[Y_GP * SY + a * YA + yo3];
        }
      }
    }
    stage_sync<WL>(); // X and Y are idle from here until B3 writes Y
    // first geometry half into LDS (its registers are freed before F3)
    if constexpr (GEO)
    {
#pragma unroll
      for (int k = 0; k < GPT0; ++k)
      {
        const int p = k * BT + tid;
        if (p < GP0)
          Ls[p] = gp0[k];
      }
      stage_sync<WL>();
    }

    // F3: contract axis a
    T uq[Q], da[Q], db[Q], dc[Q];
    if (live && EO)
    {
      T pe[NP], po[NH], ev[QP], od[QH];
      eo_split<N>(ipp, pe, po);
      if constexpr (UQ)
      {
        eo_term<N, Q, 1, true>(tab, EPF, pe, po, ev, od);
        eo_join<Q>(ev, od, uq);
      }
      if constexpr (GRAD)
      {
        eo_term<N, Q, -1, true>(tab, EGF, pe, po, ev, od);
        eo_join<Q>(ev, od, da);
        T ge[NP], go[NH];
        eo_split<N>(igp, ge, go);
        eo_term<N, Q, 1, true>(tab, EPF, ge, go, ev, od);
        eo_join<Q>(ev, od, db);
      }
      if constexpr (FULL)
      {
        T ce[NP], co[NH];
        eo_split<N>(ipg, ce, co);
        eo_term<N, Q, 1, true>(tab, EPF, ce, co, ev, od);
        eo_join<Q>(ev, od, dc);
      }
    }
    else if (live)
    {
#pragma unroll
      for (int a = 0; a < N; ++a)
      {
        T rp[Q], rg[Q];
        load_row<Q>(ph, a * Q, rp);
        if constexpr (GRAD)
          load_row<Q>(gr, a * Q, rg);
#pragma unroll
        for (int q = 0; q < Q; ++q)
        {
          if constexpr (UQ)
            uq[q] = a == 0 ? rp[q] * ipp[a] : fma(rp[q], ipp[a], uq[q]);
          if constexpr (GRAD)
          {
            da[q] = a == 0 ? rg[q] * ipp[a] : fma(rg[q], ipp[a], da[q]);
            db[q] = a == 0 ? rp[q] * igp[a] : fma(rp[q], igp[a], db[q]);
          }
          if constexpr (FULL)
            dc[q] = a == 0 ? rp[q] * ipg[a] : fma(rp[q], ipg[a], dc[q]);
        }
      }
    }

    // pointwise: v = S(q) d (+ detJw u), the geometry staged through LDS one
    // half of the qa-planes at a time
    T va[Q], vb[Q], vc[Q], vm[Q];
    double eacc = 0.0; // this thread's share of x^T A x (live cells only)
    auto pointwise = [&](int q, int qa0, int per_cell) {
      T g[6];
      if constexpr (GEO)
      {
        const Piece<T>* Lg =
            Ls + lc * per_cell + ((q - qa0) * Q2 + ln) * 3;
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
          const Piece<T> pc = Lg[j];
          g[2 * j] = pc.x;
          g[2 * j + 1] = pc.y;
        }
      }
      if constexpr (AFF)
      {
        // G(q) = w(qa) w(qb) w(qc) S, detJw(q) = w(qa) w(qb) w(qc) detJ
        const T wq = tab[WOFF + q] * wbc;
        if constexpr (GRAD)
        {
#pragma unroll
          for (int k = 0; k < 6; ++k)
            g[k] = wq * cg[k];
        }
        if constexpr (MASSTERM)
          dw[q] = wq * cg[6];
      }
      if constexpr (FULL)
      {
        // S = [[G5,G4,G2],[G4,G3,G1],[G2,G1,G0]] on (d_a, d_b, d_c)
        va[q] = g[5] * da[q] + g[4] * db[q] + g[2] * dc[q];
        vb[q] = g[4] * da[q] + g[3] * db[q] + g[1] * dc[q];
        vc[q] = g[2] * da[q] + g[1] * db[q] + g[0] * dc[q];
      }
      else if constexpr (LIT)
      {
        // third row of the transform on (d_a, d_b, undifferentiated u)
        // (src/stiffness.hpp:27 with Wz from src/stiffness.hpp:76-77)
        vc[q] = g[2] * da[q] + g[1] * db[q] + g[0] * uq[q];
      }
      if constexpr (MASSTERM)
        vm[q] = dw[q] * uq[q];
      // energy x_e^T A_e x_e = sum_q d.S d + detJw u^2 (the CG's p.Ap)
      if constexpr (FULL)
        eacc += (double)(da[q] * va[q] + db[q] * vb[q] + dc[q] * vc[q]);
      if constexpr (MASSTERM)
        eacc += (double)(uq[q] * vm[q]);
    };
    if (live)
    {
#pragma unroll
      for (int q = 0; q < QA1; ++q)
        pointwise(q, 0, 3 * QA1 * Q2);
    }
    if constexpr (GEO)
    {
      stage_sync<WL>();
#pragma unroll
      for (int k = 0; k < GPT1; ++k)
      {
        const int p = k * BT + tid;
        if (p < GP1)
          Ls[p] = gp1[k];
      }
      stage_sync<WL>();
    }
    T ra_[N], rb_[N], rc_[N]; // B3 outputs, stored after the geometry is read
    if (live)
    {
#pragma unroll
      for (int q = QA1; q < Q; ++q)
        pointwise(q, QA1, 3 * (Q - QA1) * Q2);
      // B3: contract axis a (transposed)
      if constexpr (EO)
      {
        T ev[NP], od[NH];
        if constexpr (FULL)
        {
          T xe[QP], xo[QH];
          eo_split<Q>(va, xe, xo);
          eo_term<Q, N, -1, true>(tab, EGB, xe, xo, ev, od);
        }
        if constexpr (MASSTERM)
        {
          T me[QP], mo[QH];
          eo_split<Q>(vm, me, mo);
          eo_term<Q, N, 1, !FULL>(tab, EPB, me, mo, ev, od);
        }
        eo_join<N>(ev, od, ra_);
        if constexpr (FULL)
        {
          T be[QP], bo[QH];
          eo_split<Q>(vb, be, bo);
          eo_term<Q, N, 1, true>(tab, EPB, be, bo, ev, od);
          eo_join<N>(ev, od, rb_);
        }
        if constexpr (GRAD)
        {
          T ce[QP], co[QH];
          eo_split<Q>(vc, ce, co);
          eo_term<Q, N, 1, true>(tab, EPB, ce, co, ev, od);
          eo_join<N>(ev, od, rc_);
        }
      }
      else
      {
#pragma unroll
        for (int a = 0; a < N; ++a)
        {
          T rp[Q], rg[Q];
          load_bwd_row<T, N, Q, REF>(ph, a, rp);
          if constexpr (FULL)
            load_bwd_row<T, N, Q, REF>(gr, a, rg);
          T ra = 0, rb = 0, rc = 0;
#pragma unroll
          for (int q = 0; q < Q; ++q)
          {
            if constexpr (FULL)
            {
              ra = fma(rg[q], va[q], ra);
              rb = fma(rp[q], vb[q], rb);
            }
            if constexpr (GRAD)
              rc = fma(rp[q], vc[q], rc);
            if constexpr (MASSTERM)
              ra = fma(rp[q], vm[q], ra);
          }
          ra_[a] = ra;
          rb_[a] = rb;
          rc_[a] = rc;
        }
      }
    }
    if constexpr (GEO)
      stage_sync<WL>(); // the geometry in LDS is consumed; Y may be rewritten
    // energy partial of this wave (fixed order: lane tree, one slot per
    // wave of workgroup blk)
    if constexpr (!BLK && !LIT)
    {
      if (epart)
      {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
          eacc += __shfl_xor(eacc, o, 64);
        if ((threadIdx.x & 63) == 0)
          epart[blk * (blockDim.x >> 6) + (threadIdx.x >> 6)] = eacc;
      }
    }
    if (live)
    {
#pragma unroll
      for (int a = 0; a < N; ++a)
      {
        const int o = a * YA + yo3;
        if constexpr (FULL || MASSTERM)
          Yc[Y_RA * SY + o] = ra_[a];
        if constexpr (FULL)
          Yc[Y_RB * SY + o] = rb_[a];
        if constexpr (GRAD)
          Yc[Y_RC * SY + o] = rc_[a];
      }
    }
    stage_sync<WL>();

    // ---------------- B2: contract axis c (transposed) --------------------
    T ra[Q], rb[Q], rc[Q];
    if (live && role2)
    {
      const int o = r2a * YA + r2q * YQ;
#pragma unroll
      for (int q = 0; q < Q; ++q)
      {
        if constexpr (FULL || MASSTERM)
          ra[q] = Yc[Y_RA * SY + o + q];
        if constexpr (FULL)
          rb[q] = Yc[Y_RB * SY + o + q];
        if constexpr (GRAD)
          rc[q] = Yc[Y_RC * SY + o + q];
      }
    }
    stage_sync<WL>(); // Y is read; X overwrites the same LDS
    if (live && role2 && EO)
    {
      T s0[N], sb[N], ev[NP], od[NH];
      {
        T ae[QP], ao[QH];
        eo_split<Q>(ra, ae, ao);
        eo_term<Q, N, 1, true>(tab, EPB, ae, ao, ev, od);
      }
      if constexpr (GRAD)
      {
        T ce[QP], co[QH];
        eo_split<Q>(rc, ce, co);
        eo_term<Q, N, -1, false>(tab, EGB, ce, co, ev, od);
      }
      eo_join<N>(ev, od, s0);
      if constexpr (FULL)
      {
        T be[QP], bo[QH];
        eo_split<Q>(rb, be, bo);
        eo_term<Q, N, 1, true>(tab, EPB, be, bo, ev, od);
        eo_join<N>(ev, od, sb);
      }
#pragma unroll
      for (int c = 0; c < N; ++c)
      {
        const int xo = r2q * XQ + c * XC + r2a;
        Xc[X_S0 * SX + xo] = s0[c];
        if constexpr (FULL)
          Xc[X_SB * SX + xo] = sb[c];
      }
    }
    else if (live && role2)
    {
#pragma unroll
      for (int c = 0; c < N; ++c)
      {
        T rp[Q], rg[Q];
        load_bwd_row<T, N, Q, REF>(ph, c, rp);
        if constexpr (GRAD)
          load_bwd_row<T, N, Q, REF>(gr, c, rg);
        T s0 = 0, sb = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q)
        {
          if constexpr (FULL || MASSTERM)
            s0 = fma(rp[q], ra[q], s0);
          if constexpr (GRAD)
            s0 = fma(rg[q], rc[q], s0);
          if constexpr (FULL)
            sb = fma(rp[q], rb[q], sb);
        }
        const int xo = r2q * XQ + c * XC + r2a;
        Xc[X_S0 * SX + xo] = s0;
        if constexpr (FULL)
          Xc[X_SB * SX + xo] = sb;
      }
    }
    stage_sync<WL>();

    // ---------------- B1: contract axis b (transposed) + store/scatter ----
    if (live && role1)
    {
      T s0[Q], sb[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q)
      {
        s0[q] = Xc[X_S0 * SX + q * XQ + r1c * XC + r1a];
        if constexpr (FULL)
          sb[q] = Xc[X_SB * SX + q * XQ + r1c * XC + r1a];
      }
      // assembled: lattice position of (a, 0, c) of this thread's cell
      long long wpos = 0, wrow = 0;
      if constexpr (ASM)
      {
        const unsigned cu = (unsigned)cell, t = cu / (unsigned)ba.ncz;
        const long long ck = cu - t * (unsigned)ba.ncz;
        const unsigned ti = t / (unsigned)ba.ncy;
        const long long cj = t - ti * (unsigned)ba.ncy, ci = ti;
        wrow = (long long)ba.ncz * N;
        wpos = ((ci * N + r1a) * ba.ncy + cj) * N * wrow + ck * N + r1c;
      }
      // even-odd: every output of the line at once
      T wv[EO ? N : 1];
      if constexpr (EO)
      {
        T ev[NP], od[NH];
        {
          T se[QP], so[QH];
          eo_split<Q>(s0, se, so);
          eo_term<Q, N, 1, true>(tab, EPB, se, so, ev, od);
        }
        if constexpr (FULL)
        {
          T be[QP], bo[QH];
          eo_split<Q>(sb, be, bo);
          eo_term<Q, N, -1, false>(tab, EGB, be, bo, ev, od);
        }
        eo_join<N>(ev, od, wv);
      }
#pragma unroll
      for (int b = 0; b < N; ++b)
      {
        T w = 0;
        if constexpr (EO)
          w = wv[b];
        else
        {
          T rp[Q], rg[Q];
          load_bwd_row<T, N, Q, REF>(ph, b, rp);
          if constexpr (FULL)
            load_bwd_row<T, N, Q, REF>(gr, b, rg);
#pragma unroll
          for (int q = 0; q < Q; ++q)
          {
            w = fma(rp[q], s0[q], w);
            if constexpr (FULL)
              w = fma(rg[q], sb[q], w);
          }
        }
        if constexpr (BLK)
        {
          const int s = slot0 + b * BZ;
          if (gid[s] == idx[b])
          {
            atomicAdd(acc + s, w); // LDS
            atomicAdd(cnt + s, 1);
          }
          else
            atomic_add(y + idx[b], w);
        }
        else if constexpr (ASM)
          y[wpos + b * wrow] = w;
        else if constexpr (STR)
          atomic_add(y + sbase + b * sstep, w);
        else if constexpr (ZR)
        {
          constexpr int ZB = zrun_cells<T, N, Q, OP>() * P + 1;
          lds_add(zbox + (r1a * N + b) * ZB + (int)(cell - zcell0) * P + r1c,
                  w);
        }
        else if constexpr (GATHER)
          atomic_add(y + idx[b], w);
        else
          y[cell * N3 + uoff + b * N] = w;
      }
    }

    // blocked: write the box back.  A dof whose every touch came from this
    // block is stored (or added without an atomic); the rest are atomics.
    // Consecutive threads take consecutive slots, i.e. runs of consecutive
    // global dofs for the structured dofmap.
    if constexpr (BLK)
    {
      __syncthreads();
      for (int s = threadIdx.x; s < BOX; s += blockDim.x)
      {
        const int c = cnt[s];
        if (c > 0)
        {
          const int32_t g = gid[s];
          const T v = acc[s];
          const int t = ba.touch[g];
          if (c == t && t < 255)
          {
            if (ba.overwrite)
              y[g] = v;
            else
              y[g] += v;
          }
          else
            atomic_add(y + g, v);
        }
      }
    }
  }
}

// Occupancy hint (minimum waves per SIMD): the FP64 stiffness instances up
// to Q = 7 are held to 3 waves/SIMD (<= 168 VGPRs), which they reach without
// spilling; measured 3 % faster at P=4 and P=6 than the unconstrained
// allocation (170 VGPRs, 2 waves/SIMD).  K+M spills under the same bound and
// is left alone.
// Workgroups of the direct kernel are dealt to the 8 XCDs round-robin.  For
// the FP64 gathered instances at Q = 7, 8 (one cell per wave), giving each
// XCD runs of 16 consecutive cell blocks lets cells that share faces read
// their x lines through one L2: measured 0.6-2 % faster there (stiffness
// P=6: 514 -> 503 us), and slower at lower Q and in FP32, which keep the
// plain order.  1 = no remapping.
template <typename T, int Q, bool GATHER>
constexpr int xcd_chunk()
{
  return (GATHER && sizeof(T) == 8 && (Q == 7 || Q == 8)) ? 16 : 1;
}

template <typename T, int Q, int OP, int BW>
constexpr int min_waves_per_simd()
{
  return (BW == 0 && sizeof(T) == 8 && Q <= 7 &&
          (OP == OP_STIFF || OP == OP_STIFF_LIT))
             ? 3
             : 1;
}

// The z-run scatter's work on segment seg of a structured ncx x ncy x ncz
// grid (ba): zrun_cells cells of one z-line (fewer at its end), segments
// numbered (ci*ncy + cj)*nzs + zs.  Each wave takes CPB cells of the segment
// in its own LDS work area; their values are summed into the (a, b, z) box
// of the segment's dofs in LDS, and the box is then added to y with atomics,
// consecutive lanes on consecutive global dofs along z.  (Plain stores for
// the box's interior dofs, which no other workgroup touches, measured
// slower: 449 against 383 us per P=4 64^3 plan apply, 656 against 502 at
// P=6 48^3.)  Workgroups of 64 * zrun_waves threads with the dynamic LDS
// of launch_zrun.
template <typename T, int N, int Q, int OP>
__device__ __forceinline__ void
zrun_segment(const T* __restrict__ tab, const T* __restrict__ x,
             T* __restrict__ y, const T* __restrict__ G,
             const T* __restrict__ detJw, long long num_cells,
             const BlockArgs& ba, long long seg, double* __restrict__ epart)
{
  constexpr int P = N - 1;
  constexpr int S = zrun_cells<T, N, Q, OP>();
  constexpr int ZB = S * P + 1;
  constexpr int BOXN = N * N * ZB;
  constexpr int BT = 64 * zrun_waves<T, N, Q, OP>();
  extern __shared__ __align__(16) unsigned char femtc_lds_dyn[];
  T* zbox = reinterpret_cast<T*>(femtc_lds_dyn +
                                 zrun_waves<T, N, Q, OP>() *
                                     blocked_work_bytes<T, N, Q, OP>());
  const unsigned nzs = ((unsigned)ba.ncz + S - 1) / S;
  const unsigned line = (unsigned)seg / nzs; // ci*ncy + cj
  const int ck0 = (int)((unsigned)seg - line * nzs) * S;
  const int ncell = ba.ncz - ck0 < S ? ba.ncz - ck0 : S;
  const long long zcell0 = (long long)line * ba.ncz + ck0;
  for (int i = threadIdx.x; i < BOXN; i += BT)
    zbox[i] = T(0);
  __syncthreads();
  apply_body<T, N, Q, OP, false, true, SCATTER_ZRUN, false>(
      tab, x, y, G, detJw, nullptr, num_cells, ba, seg, epart, zbox, zcell0,
      ncell);
  __syncthreads(); // every wave's box sums are in
  const unsigned ci = line / (unsigned)ba.ncy;
  const unsigned cj = line - ci * (unsigned)ba.ncy;
  const long long NYg = (long long)ba.ncy * P + 1;
  const long long NZg = (long long)ba.ncz * P + 1;
  const int zlen = ncell * P + 1;
  const long long gz0 = (long long)ck0 * P;
  for (int i = threadIdx.x; i < BOXN; i += BT)
  {
    const int ab = i / ZB, z = i - ab * ZB;
    if (z < zlen)
    {
      const int a = ab / N, b = ab - a * N;
      atomic_add(y + ((ci * P + a) * NYg + cj * P + b) * NZg + gz0 + z,
                 zbox[i]);
    }
  }
}

template <typename T, int N, int Q, int OP, bool REF, bool GATHER, int BW,
          bool EO = false>
__global__ void
__launch_bounds__((BW > 0                ? 64 * BW
                   : BW == SCATTER_ZRUN ? 64 * zrun_waves<T, N, Q, OP>()
                                        : group_threads<T, Q, BW>()),
                  (min_waves_per_simd<T, Q, OP, BW>()))
    apply_cells(const T* __restrict__ tab, const T* __restrict__ x,
                T* __restrict__ y, const T* __restrict__ G,
                const T* __restrict__ detJw, const int32_t* __restrict__ dofmap,
                long long num_cells, const int* __restrict__ skip,
                BlockArgs ba, double* __restrict__ epart)
{
  if (skip && *skip)
    return;
  long long blk = blockIdx.x;
  constexpr int XC = BW == 0 ? xcd_chunk<T, Q, GATHER>() : 1;
  if constexpr (XC > 1)
  {
    // groups of 8 XC workgroups cover 8 XC consecutive blocks, XC per XCD
    const long long w = blockIdx.x;
    const long long g = w / (8 * XC), r = w % (8 * XC);
    blk = g * 8 * XC + (r % 8) * XC + r / 8;
    if (blk >= ba.nblk)
      return; // uniform over the workgroup
  }
  if constexpr (BW == SCATTER_ZRUN)
  {
    // one z-segment per workgroup, segments dealt to the XCDs in contiguous
    // ranges as below
    const long long per = (ba.nblk + 7) / 8;
    const long long seg = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (seg >= ba.nblk)
      return; // uniform over the workgroup
    zrun_segment<T, N, Q, OP>(tab, x, y, G, detJw, num_cells, ba, seg, epart);
    return;
  }
  if constexpr (BW == SCATTER_ASSEMBLED || BW == SCATTER_STRUCT)
  {
    // Workgroups are dealt to the 8 XCDs round-robin; give each XCD a
    // contiguous range of cell blocks so that the partial lines its stores
    // write, and the x lines neighbouring cells share, meet in its own L2
    // (a layout choice only: any mapping of workgroups to blocks is
    // correct).
    const long long per = (ba.nblk + 7) / 8;
    blk = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (blk >= ba.nblk)
      return; // uniform over the workgroup
  }
  apply_body<T, N, Q, OP, REF, GATHER, BW, EO>(tab, x, y, G, detJw, dofmap,
                                               num_cells, ba, blk, epart);
}

// Batched launch: many independent problems of one (T, N, Q, OP, GATHER)
// in one grid.  Problem i owns the workgroup-sized blocks
// [blk_begin_i, blk_begin_{i+1}); each workgroup handles `per_wg`
// consecutive blocks (a persistent-style knob chosen by the autotuner) and
// finds each block's problem by binary search over blk_begin.
template <typename T, int N, int Q, int OP, bool GATHER>
__global__ void __launch_bounds__(block_threads<Q>())
    apply_batch(const BatchDesc* __restrict__ items, int nitems,
                long long nblk_total, int per_wg)
{
  // the problems' first blocks, searched in LDS (one global round trip
  // instead of one per bisection step)
  __shared__ long long first_blk[BATCH_MAX_ITEMS];
  for (int i = threadIdx.x; i < nitems; i += blockDim.x)
    first_blk[i] = items[i].blk_begin;
  __syncthreads();
  for (int it = 0; it < per_wg; ++it)
  {
    const long long b = (long long)blockIdx.x * per_wg + it;
    if (b >= nblk_total)
      return; // uniform over the workgroup
    int lo = 0, hi = nitems - 1;
    while (lo < hi)
    {
      const int mid = (lo + hi + 1) >> 1;
      if (first_blk[mid] <= b)
        lo = mid;
      else
        hi = mid - 1;
    }
    const BatchDesc d = items[lo];
    apply_body<T, N, Q, OP, false, GATHER, 0>(
        (const T*)d.tab, (const T*)d.x, (T*)d.y, (const T*)d.G,
        (const T*)d.detJw, d.dofmap, d.num_cells, BlockArgs{},
        b - d.blk_begin);
    __syncthreads(); // the next block reuses the LDS work area
  }
}

// Batched z-run launch: the z-segments of many structured problems of one
// (T, N, Q, OP) in one grid, problem i owning segments [blk_begin_i,
// blk_begin_{i+1}), segments dealt to the XCDs in contiguous ranges.
template <typename T, int N, int Q, int OP>
__global__ void __launch_bounds__((64 * zrun_waves<T, N, Q, OP>()))
    apply_batch_zrun(const BatchDesc* __restrict__ items, int nitems,
                     long long nseg_total)
{
  __shared__ long long first_seg[BATCH_MAX_ITEMS];
  for (int i = threadIdx.x; i < nitems; i += blockDim.x)
    first_seg[i] = items[i].blk_begin;
  __syncthreads();
  const long long per = (nseg_total + 7) / 8;
  const long long seg = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (seg >= nseg_total)
    return; // uniform over the workgroup
  int lo = 0, hi = nitems - 1;
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (first_seg[mid] <= seg)
      lo = mid;
    else
      hi = mid - 1;
  }
  const BatchDesc d = items[lo];
  BlockArgs ba{};
  ba.ncx = d.grid[0];
  ba.ncy = d.grid[1];
  ba.ncz = d.grid[2];
  zrun_segment<T, N, Q, OP>((const T*)d.tab, (const T*)d.x, (T*)d.y,
                            (const T*)d.G, (const T*)d.detJw, d.num_cells, ba,
                            seg - d.blk_begin, nullptr);
}

template <typename T, int N, int Q, int OP>
int launch_batch_zrun(const BatchReq& r)
{
  if constexpr (Q * Q > 64)
    return FEMTC_ERR_INVALID_ARGUMENT;
  else
  {
    constexpr int W = zrun_waves<T, N, Q, OP>();
    if (r.zrun_shape)
    {
      r.zrun_shape[0] = zrun_cells<T, N, Q, OP>();
      r.zrun_shape[1] = W;
      return FEMTC_OK;
    }
    if (!r.gather || r.nitems < 1 || r.nitems > BATCH_MAX_ITEMS)
      return FEMTC_ERR_INVALID_ARGUMENT;
    if (r.nblk_total == 0)
      return FEMTC_OK;
    const long long grid = 8 * ((r.nblk_total + 7) / 8);
    if (grid > 0x7fffffffLL)
      return FEMTC_ERR_INVALID_ARGUMENT;
    constexpr size_t lds = W * blocked_work_bytes<T, N, Q, OP>() +
                           sizeof(T) * N * N *
                               (zrun_cells<T, N, Q, OP>() * (N - 1) + 1);
    hipLaunchKernelGGL((apply_batch_zrun<T, N, Q, OP>), dim3((unsigned)grid),
                       dim3(64 * W), lds, r.stream, r.items, r.nitems,
                       r.nblk_total);
    return hip_status(hipGetLastError());
  }
}

template <typename T, int N, int Q, int OP, bool GATHER>
int launch_batch_t(const BatchReq& r)
{
  if (r.nblk_total == 0)
    return FEMTC_OK;
  const int per = r.per_wg < 1 ? 1 : r.per_wg;
  const long long grid = (r.nblk_total + per - 1) / per;
  if (grid > 0x7fffffffLL)
    return FEMTC_ERR_INVALID_ARGUMENT;
  hipLaunchKernelGGL((apply_batch<T, N, Q, OP, GATHER>), dim3((unsigned)grid),
                     dim3(block_threads<Q>()), 0, r.stream, r.items, r.nitems,
                     r.nblk_total, per);
  return hip_status(hipGetLastError());
}

template <typename T, int N, int Q, int OP>
int launch_batch_g(const BatchReq& r)
{
  if (r.zrun)
    return launch_batch_zrun<T, N, Q, OP>(r);
  return r.gather ? launch_batch_t<T, N, Q, OP, true>(r)
                  : launch_batch_t<T, N, Q, OP, false>(r);
}

template <int N, int Q>
int launch_batch_nq(const BatchReq& r)
{
  const bool d = r.dtype == FEMTC_F64;
  if (r.dtype != FEMTC_F64 && r.dtype != FEMTC_F32)
    return FEMTC_ERR_INVALID_ARGUMENT;
  switch (r.op)
  {
  case OP_MASS:
    return d ? launch_batch_g<double, N, Q, OP_MASS>(r)
             : launch_batch_g<float, N, Q, OP_MASS>(r);
  case OP_STIFF:
    return d ? launch_batch_g<double, N, Q, OP_STIFF>(r)
             : launch_batch_g<float, N, Q, OP_STIFF>(r);
  case OP_KPM:
    return d ? launch_batch_g<double, N, Q, OP_KPM>(r)
             : launch_batch_g<float, N, Q, OP_KPM>(r);
  }
  return FEMTC_ERR_INVALID_ARGUMENT;
}

template <int N>
int launch_batch_n(const BatchReq& r)
{
  if (r.N != N)
    return FEMTC_ERR_DEGREE_NOT_SUPPORTED;
  switch (r.Q - N)
  {
  case 0: return launch_batch_nq<N, N>(r);
  case 1: return launch_batch_nq<N, N + 1>(r);
  case 2: return launch_batch_nq<N, N + 2>(r);
  }
  return FEMTC_ERR_INVALID_ARGUMENT;
}

// Launch one instance: one workgroup per batch of CPB cells.
template <typename T, int N, int Q, int OP, bool REF, bool GATHER,
          bool EO = false>
int launch_cells(const ApplyReq& r)
{
  constexpr int CPB = cells_per_block<Q>();
  constexpr long long XC = xcd_chunk<T, Q, GATHER>();
  const long long nblk = (r.num_cells + CPB - 1) / CPB;
  if (nblk == 0)
    return FEMTC_OK;
  if (nblk > 0x7fffff00LL)
    return FEMTC_ERR_INVALID_ARGUMENT;
  BlockArgs ba{};
  ba.nblk = (int)nblk;
  const long long grid =
      XC > 1 ? 8 * XC * ((nblk + 8 * XC - 1) / (8 * XC)) : nblk;
  hipLaunchKernelGGL((apply_cells<T, N, Q, OP, REF, GATHER, 0, EO>),
                     dim3((unsigned)grid), dim3(block_threads<Q>()), 0,
                     r.stream, (const T*)(EO ? r.eo_tab : r.tab),
                     (const T*)r.x, (T*)r.y,
                     (const T*)r.G, (const T*)r.detJw, r.dofmap,
                     (long long)r.num_cells, r.skip, ba, r.energy);
  return hip_status(hipGetLastError());
}

// Largest number of waves of a blocked-scatter workgroup.
constexpr int BLOCKED_MAX_WAVES = 8;

// Blocked scatter: one workgroup per cbx x cby x cbz block of cells.
template <typename T, int N, int Q, int OP>
int launch_blocked(const ApplyReq& r)
{
  if constexpr (Q * Q > 64)
    return FEMTC_ERR_INVALID_ARGUMENT;
  else
  {
    constexpr int CPB = cells_per_block<Q>();
    constexpr int BW = BLOCKED_MAX_WAVES;
    const int cells = r.cbx * r.cby * r.cbz;
    const int waves = (cells + CPB - 1) / CPB;
    if (r.cbx < 1 || r.cby < 1 || r.cbz < 1 || r.cbz % CPB != 0 ||
        waves > BW || !r.touch ||
        (long long)r.ncx * r.ncy * r.ncz != r.num_cells)
      return FEMTC_ERR_INVALID_ARGUMENT;
    if (r.num_cells == 0)
      return FEMTC_OK;
    BlockArgs ba{};
    ba.ncx = r.ncx;
    ba.ncy = r.ncy;
    ba.ncz = r.ncz;
    ba.cbx = r.cbx;
    ba.cby = r.cby;
    ba.cbz = r.cbz;
    const long long nbx = (r.ncx + r.cbx - 1) / r.cbx;
    ba.nby = (r.ncy + r.cby - 1) / r.cby;
    ba.nbz = (r.ncz + r.cbz - 1) / r.cbz;
    ba.touch = r.touch;
    ba.overwrite = r.overwrite;
    const long long nblk = nbx * ba.nby * ba.nbz;
    if (nblk > 0x7fffffffLL)
      return FEMTC_ERR_INVALID_ARGUMENT;
    const size_t lds =
        waves * blocked_work_bytes<T, N, Q, OP>() +
        blocked_box_bytes<T>(blocked_box(N - 1, r.cbx, r.cby, r.cbz));
    if (lds > 160 * 1024)
      return FEMTC_ERR_INVALID_ARGUMENT;
    auto kern = apply_cells<T, N, Q, OP, false, true, BW>;
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void*)kern,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
      return FEMTC_ERR_HIP;
    hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(64 * waves), lds,
                       r.stream, (const T*)r.tab, (const T*)r.x, (T*)r.y,
                       (const T*)r.G, (const T*)r.detJw, r.dofmap,
                       (long long)r.num_cells, r.skip, ba, nullptr);
    return hip_status(hipGetLastError());
  }
}

// Assembled scatter: the cell pass into r.lattice, then (assembled == 1) the
// assembly pass into y, stream-ordered.
template <typename T, int N, int Q, int OP, bool EO = false>
int launch_assembled(const ApplyReq& r)
{
  constexpr int CPB = cells_per_group<T, Q, SCATTER_ASSEMBLED>();
  if (r.ncx < 1 || r.ncy < 1 || r.ncz < 1 || !r.lattice ||
      (long long)r.ncx * r.ncy * r.ncz != r.num_cells ||
      r.num_cells > 0x7fffffffLL)
    return FEMTC_ERR_INVALID_ARGUMENT;
  const long long nblk = (r.num_cells + CPB - 1) / CPB;
  if (nblk > 0x7ffffff0LL)
    return FEMTC_ERR_INVALID_ARGUMENT;
  BlockArgs ba{};
  ba.ncx = r.ncx;
  ba.ncy = r.ncy;
  ba.ncz = r.ncz;
  ba.nblk = (int)nblk;
  const long long grid = 8 * ((nblk + 7) / 8);
  hipLaunchKernelGGL((apply_cells<T, N, Q, OP, false, true, SCATTER_ASSEMBLED,
                                  EO>),
                     dim3((unsigned)grid),
                     dim3(group_threads<T, Q, SCATTER_ASSEMBLED>()), 0,
                     r.stream, (const T*)(EO ? r.eo_tab : r.tab),
                     (const T*)r.x, (T*)r.lattice,
                     (const T*)r.G, (const T*)r.detJw, r.dofmap,
                     (long long)r.num_cells, r.skip, ba, r.energy);
  if (r.assembled == 1)
  {
    const int st = hip_status(hipGetLastError());
    if (st != FEMTC_OK)
      return st;
    return launch_assemble<T, N - 1>(
        (const T*)r.lattice, (T*)r.y, LatticeGrid{r.ncx, r.ncy, r.ncz, N - 1},
        r.overwrite ? 0 : 1, r.skip, r.stream);
  }
  return hip_status(hipGetLastError());
}

// Structured direct scatter: the atomic scatter with closed-form indices,
// workgroups taken XCD-chunked as in the assembled cell pass.
template <typename T, int N, int Q, int OP>
int launch_struct(const ApplyReq& r)
{
  constexpr int CPB = cells_per_block<Q>();
  if (r.ncx < 1 || r.ncy < 1 || r.ncz < 1 ||
      (long long)r.ncx * r.ncy * r.ncz != r.num_cells ||
      r.num_cells > 0x7fffffffLL)
    return FEMTC_ERR_INVALID_ARGUMENT;
  const long long nblk = (r.num_cells + CPB - 1) / CPB;
  BlockArgs ba{};
  ba.ncx = r.ncx;
  ba.ncy = r.ncy;
  ba.ncz = r.ncz;
  ba.nblk = (int)nblk;
  const long long grid = 8 * ((nblk + 7) / 8);
  hipLaunchKernelGGL((apply_cells<T, N, Q, OP, false, true, SCATTER_STRUCT>),
                     dim3((unsigned)grid), dim3(block_threads<Q>()), 0,
                     r.stream, (const T*)r.tab, (const T*)r.x, (T*)r.y,
                     (const T*)r.G, (const T*)r.detJw, r.dofmap,
                     (long long)r.num_cells, r.skip, ba, r.energy);
  return hip_status(hipGetLastError());
}

// Structured z-run scatter: one workgroup per z-segment of cells.
template <typename T, int N, int Q, int OP>
int launch_zrun(const ApplyReq& r)
{
  if constexpr (Q * Q > 64)
    return FEMTC_ERR_INVALID_ARGUMENT;
  else
  {
    constexpr int S = zrun_cells<T, N, Q, OP>();
    if (r.zrun_shape)
    {
      r.zrun_shape[0] = S;
      r.zrun_shape[1] = zrun_waves<T, N, Q, OP>();
      return FEMTC_OK;
    }
    if (r.ncx < 1 || r.ncy < 1 || r.ncz < 1 ||
        (long long)r.ncx * r.ncy * r.ncz != r.num_cells ||
        r.num_cells > 0x7fffffffLL)
      return FEMTC_ERR_INVALID_ARGUMENT;
    const long long nseg =
        (long long)r.ncx * r.ncy * ((r.ncz + S - 1) / S);
    if (nseg > 0x7ffffff0LL)
      return FEMTC_ERR_INVALID_ARGUMENT;
    BlockArgs ba{};
    ba.ncx = r.ncx;
    ba.ncy = r.ncy;
    ba.ncz = r.ncz;
    ba.nblk = (int)nseg;
    const long long grid = 8 * ((nseg + 7) / 8);
    constexpr int W = zrun_waves<T, N, Q, OP>();
    constexpr size_t lds = W * blocked_work_bytes<T, N, Q, OP>() +
                           sizeof(T) * N * N * (S * (N - 1) + 1);
    static_assert(lds <= 64 * 1024, "z-run workgroup LDS");
    hipLaunchKernelGGL((apply_cells<T, N, Q, OP, false, true, SCATTER_ZRUN>),
                       dim3((unsigned)grid), dim3(64 * W), lds, r.stream,
                       (const T*)r.tab, (const T*)r.x, (T*)r.y, (const T*)r.G,
                       (const T*)r.detJw, r.dofmap, (long long)r.num_cells,
                       r.skip, ba, r.energy);
    return hip_status(hipGetLastError());
  }
}

// Even-odd instances (Q == N only): direct or assembled scatter.
template <typename T, int N, int Q, int OP>
int launch_eo(const ApplyReq& r)
{
  return r.assembled ? launch_assembled<T, N, Q, OP, true>(r)
                     : launch_cells<T, N, Q, OP, false, true, true>(r);
}

template <typename T, int N, int Q>
int dispatch_op(const ApplyReq& r)
{
  const bool g = r.gather != 0;
  if (r.eo_tab)
  {
    if constexpr (N == Q)
    {
      if (!g || r.reference_mode || r.blocked)
        return FEMTC_ERR_INVALID_ARGUMENT;
      switch (r.op)
      {
      case OP_MASS: return launch_eo<T, N, Q, OP_MASS>(r);
      case OP_STIFF: return launch_eo<T, N, Q, OP_STIFF>(r);
      case OP_KPM: return launch_eo<T, N, Q, OP_KPM>(r);
      case OP_MASS_AFF: return launch_eo<T, N, Q, OP_MASS_AFF>(r);
      case OP_STIFF_AFF: return launch_eo<T, N, Q, OP_STIFF_AFF>(r);
      case OP_KPM_AFF: return launch_eo<T, N, Q, OP_KPM_AFF>(r);
      }
    }
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  if (r.assembled)
  {
    if (!g || r.reference_mode || r.blocked)
      return FEMTC_ERR_INVALID_ARGUMENT;
    switch (r.op)
    {
    case OP_MASS: return launch_assembled<T, N, Q, OP_MASS>(r);
    case OP_STIFF: return launch_assembled<T, N, Q, OP_STIFF>(r);
    case OP_KPM: return launch_assembled<T, N, Q, OP_KPM>(r);
    case OP_MASS_AFF: return launch_assembled<T, N, Q, OP_MASS_AFF>(r);
    case OP_STIFF_AFF: return launch_assembled<T, N, Q, OP_STIFF_AFF>(r);
    case OP_KPM_AFF: return launch_assembled<T, N, Q, OP_KPM_AFF>(r);
    }
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  if (r.zrun)
  {
    if (!g || r.reference_mode || r.structured_direct)
      return FEMTC_ERR_INVALID_ARGUMENT;
    switch (r.op)
    {
    case OP_MASS: return launch_zrun<T, N, Q, OP_MASS>(r);
    case OP_STIFF: return launch_zrun<T, N, Q, OP_STIFF>(r);
    case OP_KPM: return launch_zrun<T, N, Q, OP_KPM>(r);
    case OP_MASS_AFF: return launch_zrun<T, N, Q, OP_MASS_AFF>(r);
    case OP_STIFF_AFF: return launch_zrun<T, N, Q, OP_STIFF_AFF>(r);
    case OP_KPM_AFF: return launch_zrun<T, N, Q, OP_KPM_AFF>(r);
    }
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  if (r.structured_direct)
  {
    if (!g || r.reference_mode)
      return FEMTC_ERR_INVALID_ARGUMENT;
    switch (r.op)
    {
    case OP_MASS: return launch_struct<T, N, Q, OP_MASS>(r);
    case OP_STIFF: return launch_struct<T, N, Q, OP_STIFF>(r);
    case OP_KPM: return launch_struct<T, N, Q, OP_KPM>(r);
    }
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  if (r.blocked)
  {
    if (!g || r.reference_mode)
      return FEMTC_ERR_INVALID_ARGUMENT;
    switch (r.op)
    {
    case OP_MASS: return launch_blocked<T, N, Q, OP_MASS>(r);
    case OP_STIFF: return launch_blocked<T, N, Q, OP_STIFF>(r);
    case OP_KPM: return launch_blocked<T, N, Q, OP_KPM>(r);
    }
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
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
  case OP_MASS_AFF:
  case OP_STIFF_AFF:
  case OP_KPM_AFF:
    if (r.reference_mode)
      return FEMTC_ERR_INVALID_ARGUMENT;
    if (r.op == OP_MASS_AFF)
      return g ? launch_cells<T, N, Q, OP_MASS_AFF, false, true>(r)
               : launch_cells<T, N, Q, OP_MASS_AFF, false, false>(r);
    if (r.op == OP_STIFF_AFF)
      return g ? launch_cells<T, N, Q, OP_STIFF_AFF, false, true>(r)
               : launch_cells<T, N, Q, OP_STIFF_AFF, false, false>(r);
    return g ? launch_cells<T, N, Q, OP_KPM_AFF, false, true>(r)
             : launch_cells<T, N, Q, OP_KPM_AFF, false, false>(r);
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
