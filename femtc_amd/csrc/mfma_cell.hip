// The K+M cell pass with bfloat16 operands on the matrix cores (MFMA) and
// FP32 accumulation: the inner operator of the BF16 defect-correction solve.
//
// It computes the same quantity as the assembled-scatter cell pass of
// apply_kernel.hpp on affine cells (OP_KPM_AFF with SCATTER_ASSEMBLED): for
// every cell of a structured grid, W_e = (K_e + M_e) x_e with the geometry
// from 7 values per cell times the tensor-product Gauss weights, stored
// without atomics into the lattice buffer (assemble.hpp), plus one
// quadrature-point energy partial x_e^T A_e x_e per wave.  The contraction
// chain is the one of src/stiffness.hpp:68-97 and src/mass.hpp:36-47 with
// true transposes and summed gradient chains.
//
// Work decomposition.  One wave per cell.  Each of the six 1-D contraction
// stages is a set of v_mfma_f32_16x16x32_bf16 products: the table is the A
// operand (16 output rows x 32 K slots, built once on the host), and the
// cell tensor is the B operand, 16 "lines" (the tensor's two other indices)
// per tile, so a stage over N*N, N*Q or Q*Q lines takes 4 tiles at N = Q = 7.
// The 32 K slots are four groups of 8: a group holds the contracted index of
// one operand plane (N or Q <= 8 values, the rest zero in A).
//
// Precision.  A bfloat16 product of a table entry and a value has 8
// significant bits.  Both schemes here split every table entry t into
// t_hi + t_lo (two bfloat16, t accurate to about 2^-17):
//  * SPLIT = false: the cell values are rounded to bfloat16 once per stage,
//    groups (t_hi x, t_lo x, t_hi y, t_lo y) for one or two inputs x, y; the
//    operator is accurate to about 3e-3;
//  * SPLIT = true: the values are split too, x = x_hi + x_lo, and a product
//    is t_hi x_hi + t_hi x_lo + t_lo x_hi (groups 0..2; group 3 unused); the
//    operator is accurate to about 1e-5, close to FP32.
// The D accumulators are FP32; the pointwise stage and the energy are FP32
// with the energy summed in FP64 per wave.
//
// Data movement.  A stage's D tile gives each lane 4 consecutive output
// rows of one line (rows 4(g%2)+i of block g/2, g = lane/16), which are
// scattered as bfloat16 into LDS in the next stage's layout: every tensor
// is [line][8 slots] (16 bytes a line, one 16-byte read per lane per tile),
// with a second plane for the low parts when SPLIT.  The third forward stage
// is organised so that one lane holds all four fields (u, d_a, d_b, d_c) of
// its points: its A rows are (u, d_a, d_b, d_c) x 4 quadrature indices, two
// D tiles per line tile, and the pointwise geometry then runs in registers.
//
// LDS: two regions of 7 tensors in all (7 KB, 14 KB split), stage inputs
// alternating between them.
#include "assemble.hpp"
#include "common.hpp"

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace femtc
{
namespace
{

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// A-operand matrices of the table set, in this order.
// SPLIT = false:
//   0 F    rows q: phi (block 0), phi' (block 1); groups (hi, lo, 0, 0)
//   1 X0   F3 quad 0, inputs (T2, T2, T2b, T2b): rows 4 qa + (u, d_a, d_b, -)
//   2 Y0   F3 quad 0, input T2c in groups 0,1: rows 4 qa + 3 (d_c)
//   3 X1, 4 Y1: the same for qa = 4..7
//   5 BPG  rows d: groups (phi_hi, phi_lo, phi'_hi, phi'_lo), block 0
//   6 BPP  block 0 rows d: (phi_hi, phi_lo, 0, 0); block 1: (0, 0, phi_hi, phi_lo)
// SPLIT = true (groups (hi.x_hi, hi.x_lo, lo.x_hi, 0)):
//   0 F, then F3 quads: 1 T2->(u, d_a), 2 T2b->d_b, 3 T2c->d_c (qa 0..3),
//   4..6 the same for qa 4..7, then 7 Bp0 (phi^T, block 0), 8 Bg0 (phi'^T,
//   block 0), 9 Bp1 (phi^T, block 1), 10 Bpp (phi^T, both blocks)
constexpr int nmat(bool split) { return split ? 11 : 7; }

// Every cell tensor lives in LDS as 64 lines x 8 slots of bfloat16 (16 bytes
// a line): line 8 i + j for its two uncontracted indices (i, j), slot k for
// the contracted one, each index padded to 8.  A stage is then 4 tiles of 16
// lines, and every LDS address a lane touches is its own base plus
// compile-time offsets.  Every stage writes all 64 lines x 8 slots of its
// outputs before the next stage reads them, so no value is read unwritten:
// padding rows of A are zero (pad outputs are 0), pad lines of the gathered
// tensor are zero and their D columns only ever reach pad places, and pad
// quadrature points have weight 0.
constexpr int LINES = 64;
constexpr int TILES = 4;
constexpr int PLANE = LINES * 16; // bytes of one plane of a tensor

template <bool SPLIT>
struct MfmaLayout
{
  static constexpr int PL = SPLIT ? 2 : 1;
  static constexpr int TB = PLANE * PL; // bytes of one tensor
  // region A: u | T2, T2b, T2c | V1, V2, V3; region B: T1, T1g | va, vb,
  // vc, vm | X1, X2 -- stage inputs alternate between the two
  static constexpr int RA = 3 * TB;
  static constexpr int BYTES = 7 * TB;
  static constexpr int U = 0;
  static constexpr int T2(int k) { return k * TB; }
  static constexpr int V(int k) { return k * TB; }
  static constexpr int T1(int k) { return RA + k * TB; }
  static constexpr int VP(int k) { return RA + k * TB; }
  static constexpr int X(int k) { return RA + k * TB; }
};

__device__ inline f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline float u2f(uint32_t v) { return __builtin_bit_cast(float, v); }

// Two values rounded to bfloat16 (nearest even), packed lo | hi << 16.
__device__ inline uint32_t pack_bf16(float a, float b)
{
  const bf16x2 p = __builtin_convertvector((f32x2){a, b}, bf16x2);
  return __builtin_bit_cast(uint32_t, p);
}

__device__ inline void st16(unsigned char* lds, int off, uint32_t v)
{
  *reinterpret_cast<uint16_t*>(lds + off) = (uint16_t)v;
}
__device__ inline void st16hi(unsigned char* lds, int off, uint32_t v)
{
  *reinterpret_cast<uint16_t*>(lds + off) = (uint16_t)(v >> 16);
}

// The four D rows of a lane as bfloat16 at base + i * STRIDE (i = 0..3);
// SPLIT also their low parts at + PLANE.
template <bool SPLIT, int STRIDE>
__device__ inline void put4(unsigned char* lds, int base, const f32x4& d)
{
  const uint32_t h01 = pack_bf16(d[0], d[1]);
  const uint32_t h23 = pack_bf16(d[2], d[3]);
  st16(lds, base, h01);
  st16hi(lds, base + STRIDE, h01);
  st16(lds, base + 2 * STRIDE, h23);
  st16hi(lds, base + 3 * STRIDE, h23);
  if constexpr (SPLIT)
  {
    const uint32_t l01 = pack_bf16(d[0] - u2f(h01 << 16),
                                   d[1] - u2f(h01 & 0xffff0000u));
    const uint32_t l23 = pack_bf16(d[2] - u2f(h23 << 16),
                                   d[3] - u2f(h23 & 0xffff0000u));
    st16(lds, base + PLANE, l01);
    st16hi(lds, base + PLANE + STRIDE, l01);
    st16(lds, base + PLANE + 2 * STRIDE, l23);
    st16hi(lds, base + PLANE + 3 * STRIDE, l23);
  }
}

// Two values (of two tensors at offsets o0, o1 from base) and, SPLIT, their
// low parts.
template <bool SPLIT>
__device__ inline void put2(unsigned char* lds, int base, int o0, int o1,
                            float v0, float v1)
{
  const uint32_t h = pack_bf16(v0, v1);
  st16(lds, base + o0, h);
  st16hi(lds, base + o1, h);
  if constexpr (SPLIT)
  {
    const uint32_t l = pack_bf16(v0 - u2f(h << 16), v1 - u2f(h & 0xffff0000u));
    st16(lds, base + o0 + PLANE, l);
    st16hi(lds, base + o1 + PLANE, l);
  }
}

__device__ inline bf16x8 getB(const unsigned char* lds, int off)
{
  return *reinterpret_cast<const bf16x8*>(lds + off);
}

// LDS ordering between stages of the one-wave workgroup.
__device__ inline void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N, bool SPLIT>
__global__ void __launch_bounds__(64)
    kpm_mfma_cells(const uint4* __restrict__ amat, const float* __restrict__ wq,
                   const float* __restrict__ x, float* __restrict__ lattice,
                   const float* __restrict__ cg, float ks, float km, int ncx,
                   int ncy, int ncz, int nblk, const int* __restrict__ skip,
                   double* __restrict__ epart)
{
  static_assert(N >= 2 && N <= 8, "MFMA slot groups hold 8");
  using LY = MfmaLayout<SPLIT>;
  constexpr int P = N - 1;
  constexpr int NM = nmat(SPLIT);
  __shared__ __align__(16) unsigned char lds[LY::BYTES];

  if (skip && *skip)
    return;
  // workgroups are dealt to the 8 XCDs round-robin: give each XCD a
  // contiguous run of cells, so neighbouring cells share x lines and lattice
  // lines in one L2
  const int per_xcd = (nblk + 7) / 8;
  const int cell_i = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (cell_i >= nblk)
    return;
  const long long cell = cell_i;
  const int lane = threadIdx.x;
  const int g = lane >> 4;  // lane group: D rows 4g..4g+3, K slots 8g..8g+7
  const int j = lane & 15;  // B column (line within a tile)
  const bool lower = g < 2; // D block 0
  const int r0 = 4 * (g & 1); // first D row of this lane within its block
  const int jl = j & 7, jh = j >> 3; // line 16 t + j = 8 (2t + jh) + jl

  // ---- tables (A operands), weights, geometry, zeroed LDS ----------------
  bf16x8 A[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m)
    A[m] = __builtin_bit_cast(bf16x8, amat[m * 64 + lane]);
  const unsigned cu = (unsigned)cell_i, t1 = cu / (unsigned)ncz;
  const int ck = (int)(cu - t1 * (unsigned)ncz);
  const int ci = (int)(t1 / (unsigned)ncy);
  const int cj = (int)(t1 - (unsigned)ci * (unsigned)ncy);
  float S[6];
#pragma unroll
  for (int k = 0; k < 6; ++k)
    S[k] = ks * cg[cell * CELLGEO + k];
  const float detJ = km * cg[cell * CELLGEO + 6];
  // weights (zero beyond nq): w(qa) w(qc) per quad, w(qb) per tile
  const float wc = wq[jl];
  const float wac0 = wq[g] * wc, wac1 = wq[4 + g] * wc;
  float wb[TILES];
#pragma unroll
  for (int t = 0; t < TILES; ++t)
    wb[t] = wq[2 * t + jh];

  // ---- gather x (closed form of the structured numbering): lane = line
  // (a, b) writes its 8 slots c at once -----------------------------------
  {
    const int a = lane >> 3, b = lane & 7;
    float v[8] = {};
    if (a < N && b < N)
    {
      const long long NY = (long long)ncy * P + 1, NZ = (long long)ncz * P + 1;
      const float* src = x + ((long long)(ci * P + a) * NY + cj * P + b) * NZ +
                         (long long)ck * P;
#pragma unroll
      for (int c = 0; c < N; ++c)
        v[c] = src[c];
    }
    uint32_t h[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      h[k] = pack_bf16(v[2 * k], v[2 * k + 1]);
    *reinterpret_cast<uint4*>(lds + LY::U + lane * 16) =
        make_uint4(h[0], h[1], h[2], h[3]);
    if constexpr (SPLIT)
    {
      uint32_t l[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        l[k] = pack_bf16(v[2 * k] - u2f(h[k] << 16),
                         v[2 * k + 1] - u2f(h[k] & 0xffff0000u));
      *reinterpret_cast<uint4*>(lds + LY::U + PLANE + lane * 16) =
          make_uint4(l[0], l[1], l[2], l[3]);
    }
  }
  wave_sync();

  // B operand offset of a one-input product: SPLIT groups (hi, lo, hi, hi),
  // else all groups the high plane; two inputs (not SPLIT): groups 0,1 read
  // X, 2,3 read Y
  const int b1 = (SPLIT && g == 1 ? PLANE : 0) + j * 16;
  const int b2 = j * 16;
  const int ysel = lower ? 0 : 1; // which of two tensors a lane reads/writes

  // ---- F1: contract c; lines (a,b) -> T1 (phi), T1g (phi') [(a,qc)][b] ---
  {
    const int wb0 = LY::T1(ysel) + (8 * jh + r0) * 16 + jl * 2;
    bf16x8 B[TILES];
    f32x4 D[TILES];
#pragma unroll
    for (int t = 0; t < TILES; ++t)
      B[t] = getB(lds, LY::U + b1 + t * 256);
#pragma unroll
    for (int t = 0; t < TILES; ++t)
      D[t] = mfma(A[0], B[t], f32x4{});
#pragma unroll
    for (int t = 0; t < TILES; ++t)
      put4<SPLIT, 16>(lds, wb0 + t * 256, D[t]);
  }
  wave_sync();

  // ---- F2: contract b; lines (a,qc) -> T2, T2b, T2c [(qb,qc)][a] ---------
  {
    const int wo = (8 * r0 + jl) * 16 + jh * 2;
    bf16x8 B1[TILES], B2[TILES];
    f32x4 D1[TILES], D2[TILES];
#pragma unroll
    for (int t = 0; t < TILES; ++t)
    {
      B1[t] = getB(lds, LY::T1(0) + b1 + t * 256);
      B2[t] = getB(lds, LY::T1(1) + b1 + t * 256);
    }
#pragma unroll
    for (int t = 0; t < TILES; ++t)
    {
      D1[t] = mfma(A[0], B1[t], f32x4{});
      D2[t] = mfma(A[0], B2[t], f32x4{});
    }
    // lower lanes: T2 (D1) and T2c (D2); upper lanes: T2b (D1)
#pragma unroll
    for (int t = 0; t < TILES; ++t)
      put4<SPLIT, 128>(lds, LY::T2(ysel) + wo + t * 4, D1[t]);
    if (lower)
    {
#pragma unroll
      for (int t = 0; t < TILES; ++t)
        put4<SPLIT, 128>(lds, LY::T2(2) + wo + t * 4, D2[t]);
    }
  }
  wave_sync();

  // ---- F3 + pointwise: lane holds (u, d_a, d_b, d_c) at qa = g and 4+g ----
  double energy = 0.0;
  {
#pragma unroll
    for (int t = 0; t < TILES; ++t)
    {
      f32x4 D0, D1;
      if constexpr (SPLIT)
      {
        const bf16x8 bu = getB(lds, LY::T2(0) + b1 + t * 256);
        const bf16x8 bb = getB(lds, LY::T2(1) + b1 + t * 256);
        const bf16x8 bc = getB(lds, LY::T2(2) + b1 + t * 256);
        D0 = mfma(A[3], bc, mfma(A[2], bb, mfma(A[1], bu, f32x4{})));
        D1 = mfma(A[6], bc, mfma(A[5], bb, mfma(A[4], bu, f32x4{})));
      }
      else
      {
        const bf16x8 bx = getB(lds, LY::T2(ysel) + b2 + t * 256);
        const bf16x8 by = getB(lds, LY::T2(2) + b2 + t * 256);
        D0 = mfma(A[2], by, mfma(A[1], bx, f32x4{}));
        D1 = mfma(A[4], by, mfma(A[3], bx, f32x4{}));
      }
      float e = 0.f;
      float va[2], vb[2], vc[2], vm[2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
      {
        const f32x4 U = h ? D1 : D0;
        const float w = (h ? wac1 : wac0) * wb[t];
        const float u = U[0], da = U[1], db = U[2], dc = U[3];
        va[h] = w * (S[5] * da + S[4] * db + S[2] * dc);
        vb[h] = w * (S[4] * da + S[3] * db + S[1] * dc);
        vc[h] = w * (S[2] * da + S[1] * db + S[0] * dc);
        vm[h] = w * detJ * u;
        e += da * va[h] + db * vb[h] + dc * vc[h] + u * vm[h];
      }
      energy += (double)e;
      // [line (qb,qc)][slot qa]: qa = g (h = 0) and 4 + g (h = 1)
      const int o = (16 * t + j) * 16 + 2 * g;
#pragma unroll
      for (int h = 0; h < 2; ++h)
      {
        put2<SPLIT>(lds, o + 8 * h, LY::VP(0), LY::VP(1), va[h], vb[h]);
        put2<SPLIT>(lds, o + 8 * h, LY::VP(2), LY::VP(3), vc[h], vm[h]);
      }
    }
  }
  wave_sync();

  // ---- B1: contract qa; lines (qb,qc) -> V1, V2, V3 [(a,qc)][qb] ---------
  {
    const int wo = (8 * r0 + jl) * 16 + jh * 2;
    f32x4 D1[TILES], D2[TILES];
    if constexpr (SPLIT)
    {
#pragma unroll
      for (int t = 0; t < TILES; ++t)
      {
        const bf16x8 bm = getB(lds, LY::VP(3) + b1 + t * 256);
        const bf16x8 ba = getB(lds, LY::VP(0) + b1 + t * 256);
        const bf16x8 bb = getB(lds, LY::VP(1) + b1 + t * 256);
        const bf16x8 bc = getB(lds, LY::VP(2) + b1 + t * 256);
        D1[t] = mfma(A[8], ba, mfma(A[7], bm, f32x4{}));
        D2[t] = mfma(A[9], bc, mfma(A[7], bb, f32x4{}));
      }
    }
    else
    {
      // groups 0,1: v_m / v_b; groups 2,3: v_a / v_c
      bf16x8 Bma[TILES], Bbc[TILES];
      const int oma = (lower ? LY::VP(3) : LY::VP(0)) + b2;
      const int obc = (lower ? LY::VP(1) : LY::VP(2)) + b2;
#pragma unroll
      for (int t = 0; t < TILES; ++t)
      {
        Bma[t] = getB(lds, oma + t * 256);
        Bbc[t] = getB(lds, obc + t * 256);
      }
#pragma unroll
      for (int t = 0; t < TILES; ++t)
      {
        D1[t] = mfma(A[5], Bma[t], f32x4{});
        D2[t] = mfma(A[6], Bbc[t], f32x4{});
      }
    }
    // lower lanes: V1 (D1), V2 (D2 block 0); upper lanes: V3 (D2)
#pragma unroll
    for (int t = 0; t < TILES; ++t)
      put4<SPLIT, 128>(lds, LY::V(lower ? 0 : 2) + wo + t * 4,
                       lower ? D1[t] : D2[t]);
    if (lower)
    {
#pragma unroll
      for (int t = 0; t < TILES; ++t)
        put4<SPLIT, 128>(lds, LY::V(1) + wo + t * 4, D2[t]);
    }
  }
  wave_sync();

  // ---- B2: contract qb; lines (a,qc) -> X1, X2 [(a,b)][qc] ---------------
  {
    const int wo = (8 * jh + r0) * 16 + jl * 2;
    f32x4 D1[TILES], D2[TILES];
#pragma unroll
    for (int t = 0; t < TILES; ++t)
    {
      if constexpr (SPLIT)
      {
        const bf16x8 v1 = getB(lds, LY::V(0) + b1 + t * 256);
        const bf16x8 v2 = getB(lds, LY::V(1) + b1 + t * 256);
        const bf16x8 v3 = getB(lds, LY::V(2) + b1 + t * 256);
        D1[t] = mfma(A[8], v2, mfma(A[7], v1, f32x4{}));
        D2[t] = mfma(A[10], v3, f32x4{});
      }
      else
      {
        const bf16x8 v12 = getB(lds, LY::V(ysel) + b2 + t * 256);
        const bf16x8 v3 = getB(lds, LY::V(2) + b2 + t * 256);
        D1[t] = mfma(A[5], v12, f32x4{});
        D2[t] = mfma(A[6], v3, f32x4{});
      }
    }
    // lower lanes: X1 (D1); upper lanes: X2 (D2, both blocks)
#pragma unroll
    for (int t = 0; t < TILES; ++t)
      put4<SPLIT, 16>(lds, LY::X(ysel) + wo + t * 256, lower ? D1[t] : D2[t]);
  }
  wave_sync();

  // ---- B3: contract qc; lines (a,b) -> W, stored into the lattice --------
  {
    const long long wrow = (long long)ncz * N;
    const int b = jl;
    float* dst0 = lattice + ((((long long)ci * N) * ncy + cj) * N + b) * wrow +
                  (long long)ck * N + r0;
    const long long astep = (long long)ncy * N * wrow; // a -> a + 1
#pragma unroll
    for (int t = 0; t < TILES; ++t)
    {
      f32x4 d;
      if constexpr (SPLIT)
      {
        const bf16x8 x1 = getB(lds, LY::X(0) + b1 + t * 256);
        const bf16x8 x2 = getB(lds, LY::X(1) + b1 + t * 256);
        d = mfma(A[8], x2, mfma(A[7], x1, f32x4{}));
      }
      else
        d = mfma(A[5], getB(lds, LY::X(ysel) + b2 + t * 256), f32x4{});
      const int a = 2 * t + jh;
      if (lower && a < N && b < N)
      {
        float* dst = dst0 + a * astep;
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (r0 + i < N)
            dst[i] = d[i];
      }
    }
  }

  // ---- energy partial of the wave ----------------------------------------
  if (epart)
  {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      energy += __shfl_xor(energy, o, 64);
    if (lane == 0)
      epart[cell] = energy;
  }
}

// ---------------------------------------------------------------- host ----

uint16_t bf16_rne(float v)
{
  uint32_t b;
  std::memcpy(&b, &v, 4);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}

float bf16_to_float(uint16_t h)
{
  const uint32_t b = (uint32_t)h << 16;
  float v;
  std::memcpy(&v, &b, 4);
  return v;
}

// hi and lo bfloat16 parts of a table entry
void split2(double v, uint16_t& hi, uint16_t& lo)
{
  const float f = (float)v;
  hi = bf16_rne(f);
  lo = bf16_rne((float)(v - (double)bf16_to_float(hi)));
}

} // namespace

int mfma_table_matrices(int split)
{
  return nmat(split != 0);
}

// The A operands of the table set: for every matrix, lane l holds row l % 16
// and K slots 8 (l / 16) .. 8 (l / 16) + 7, 8 bfloat16 each; out has
// nmat * 64 * 8 values.  tab = [phi | phi'], each [N][Q], FP64 on the host.
int mfma_build_tables(const double* tab, int N, int Q, int split,
                      uint16_t* out)
{
  if (N < 2 || N > 8 || Q < 2 || Q > 8)
    return FEMTC_ERR_DEGREE_NOT_SUPPORTED;
  // part[t][p][d][q]: table t (0 phi, 1 phi'), part p (0 hi, 1 lo)
  uint16_t part[2][2][8][8] = {};
  for (int t = 0; t < 2; ++t)
    for (int d = 0; d < N; ++d)
      for (int q = 0; q < Q; ++q)
        split2(tab[t * N * Q + d * Q + q], part[t][0][d][q], part[t][1][d][q]);
  const int NM = nmat(split != 0);
  std::vector<uint16_t> Am(16 * 32);
  // entry (table t, part p, forward?) at row output index o, slot s
  auto fwd = [&](int t, int p, int o, int s) -> uint16_t {
    return (o < Q && s < N) ? part[t][p][s][o] : 0; // out q = o, in dof s
  };
  auto bwd = [&](int t, int p, int o, int s) -> uint16_t {
    return (o < N && s < Q) ? part[t][p][o][s] : 0; // out dof o, in q s
  };
  for (int m = 0; m < NM; ++m)
  {
    std::fill(Am.begin(), Am.end(), 0);
    for (int r = 0; r < 16; ++r)
      for (int grp = 0; grp < 4; ++grp)
        for (int s = 0; s < 8; ++s)
        {
          uint16_t v = 0;
          const int blk = r / 8, o = r % 8;
          // F3 rows: quantity r % 4 at qa = r / 4 (+ 4 for the second quad)
          const int fq = r % 4, fqa = r / 4;
          if (!split)
          {
            const int pp = grp & 1; // hi, lo, hi, lo
            switch (m)
            {
            case 0: // F
              if (grp < 2)
                v = fwd(blk, pp, o, s);
              break;
            case 1: // X0: groups 0,1 = T2, 2,3 = T2b
            case 3:
            {
              const int qa = fqa + (m == 3 ? 4 : 0);
              if (grp < 2 && fq == 0)
                v = fwd(0, pp, qa, s);
              else if (grp < 2 && fq == 1)
                v = fwd(1, pp, qa, s);
              else if (grp >= 2 && fq == 2)
                v = fwd(0, pp, qa, s);
              break;
            }
            case 2: // Y0: groups 0,1 = T2c
            case 4:
            {
              const int qa = fqa + (m == 4 ? 4 : 0);
              if (grp < 2 && fq == 3)
                v = fwd(0, pp, qa, s);
              break;
            }
            case 5: // BPG
              if (blk == 0)
                v = bwd(grp < 2 ? 0 : 1, pp, o, s);
              break;
            case 6: // BPP
              if ((blk == 0 && grp < 2) || (blk == 1 && grp >= 2))
                v = bwd(0, pp, o, s);
              break;
            }
          }
          else
          {
            if (grp == 3)
              continue;
            const int pp = grp == 2 ? 1 : 0; // hi, hi, lo
            switch (m)
            {
            case 0:
              v = fwd(blk, pp, o, s);
              break;
            case 1:
            case 4:
            {
              const int qa = fqa + (m == 4 ? 4 : 0);
              if (fq == 0)
                v = fwd(0, pp, qa, s);
              else if (fq == 1)
                v = fwd(1, pp, qa, s);
              break;
            }
            case 2:
            case 5:
            {
              const int qa = fqa + (m == 5 ? 4 : 0);
              if (fq == 2)
                v = fwd(0, pp, qa, s);
              break;
            }
            case 3:
            case 6:
            {
              const int qa = fqa + (m == 6 ? 4 : 0);
              if (fq == 3)
                v = fwd(0, pp, qa, s);
              break;
            }
            case 7: // Bp0
              if (blk == 0)
                v = bwd(0, pp, o, s);
              break;
            case 8: // Bg0
              if (blk == 0)
                v = bwd(1, pp, o, s);
              break;
            case 9: // Bp1
              if (blk == 1)
                v = bwd(0, pp, o, s);
              break;
            case 10: // Bpp
              v = bwd(0, pp, o, s);
              break;
            }
          }
          Am[r * 32 + grp * 8 + s] = v;
        }
    for (int l = 0; l < 64; ++l)
      for (int s = 0; s < 8; ++s)
        out[(m * 64 + l) * 8 + s] = Am[(l % 16) * 32 + (l / 16) * 8 + s];
  }
  return FEMTC_OK;
}

namespace
{
template <int N, bool SPLIT>
int launch_mfma_n(const MfmaReq& r)
{
  const long long nblk = (long long)r.ncx * r.ncy * r.ncz;
  const long long grid = 8 * ((nblk + 7) / 8);
  hipLaunchKernelGGL((kpm_mfma_cells<N, SPLIT>), dim3((unsigned)grid),
                     dim3(64), 0, r.stream, (const uint4*)r.amat, r.w, r.x,
                     r.lattice, r.cg, r.ks, r.km, r.ncx, r.ncy, r.ncz,
                     (int)nblk, r.skip,
                     r.energy);
  return hip_status(hipGetLastError());
}
} // namespace

int launch_mfma(const MfmaReq& r)
{
  const long long C = (long long)r.ncx * r.ncy * r.ncz;
  if (r.ncx < 1 || r.ncy < 1 || r.ncz < 1 || C > 0x7ffffff0LL || !r.amat ||
      !r.w || !r.x || !r.lattice || !r.cg || r.Q != r.N)
    return FEMTC_ERR_INVALID_ARGUMENT;
  switch (r.N * 2 + (r.split ? 1 : 0))
  {
#define FEMTC_MFMA_CASE(NN)                                                   \
  case 2 * NN:                                                                \
    return launch_mfma_n<NN, false>(r);                                       \
  case 2 * NN + 1:                                                            \
    return launch_mfma_n<NN, true>(r);
    FEMTC_MFMA_CASE(2)
    FEMTC_MFMA_CASE(3)
    FEMTC_MFMA_CASE(4)
    FEMTC_MFMA_CASE(5)
    FEMTC_MFMA_CASE(6)
    FEMTC_MFMA_CASE(7)
    FEMTC_MFMA_CASE(8)
#undef FEMTC_MFMA_CASE
  default:
    return FEMTC_ERR_DEGREE_NOT_SUPPORTED;
  }
}

} // namespace femtc

using namespace femtc;

namespace
{
int mfma_bytes(int split) { return nmat(split != 0) * 64 * 16 + 8 * 4; }

// y = sum of the lattice (assembled scatter, second pass) at degree P
int assemble_f32(const float* lat, float* y, const LatticeGrid& L,
                 int accumulate, hipStream_t s)
{
  switch (L.P)
  {
#define FEMTC_MFMA_ASM(PP)                                                    \
  case PP:                                                                    \
    return launch_assemble<float, PP>(lat, y, L, accumulate, nullptr, s);
    FEMTC_MFMA_ASM(1)
    FEMTC_MFMA_ASM(2)
    FEMTC_MFMA_ASM(3)
    FEMTC_MFMA_ASM(4)
    FEMTC_MFMA_ASM(5)
    FEMTC_MFMA_ASM(6)
    FEMTC_MFMA_ASM(7)
#undef FEMTC_MFMA_ASM
  default:
    return FEMTC_ERR_DEGREE_NOT_SUPPORTED;
  }
}
} // namespace

extern "C" {

int femtc_mfma_tables_bytes(int split, size_t* bytes)
{
  if (!bytes || (split != 0 && split != 1))
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "split must be 0 or 1");
  *bytes = (size_t)mfma_bytes(split);
  return FEMTC_OK;
}

int femtc_mfma_tables_host(int degree, int nq, int split,
                           const double* tables, void* operands)
{
  if (split != 0 && split != 1)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "split must be 0 or 1");
  if (!tables || !operands)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null tables or operands");
  if (degree < 1 || degree > 7 || nq != degree + 1)
    return fail(FEMTC_ERR_DEGREE_NOT_SUPPORTED,
                "MFMA operator: degree 1..7 with nq = degree + 1");
  uint16_t* h = static_cast<uint16_t*>(operands);
  std::memset(h, 0, (size_t)mfma_bytes(split));
  int st = mfma_build_tables(tables, degree + 1, nq, split, h);
  if (st != FEMTC_OK)
    return fail(st, "MFMA operand set");
  std::vector<double> xq(nq), wq(nq);
  if ((st = femtc_gauss_rule(nq, xq.data(), wq.data())) != FEMTC_OK)
    return st;
  float w[8] = {};
  for (int q = 0; q < nq; ++q)
    w[q] = (float)wq[q];
  std::memcpy(h + nmat(split != 0) * 64 * 8, w, sizeof(w));
  return FEMTC_OK;
}

int femtc_mfma_tables(int degree, int nq, int split, const double* tables,
                      void* operands)
{
  if (split != 0 && split != 1)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "split must be 0 or 1");
  if (!operands)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null operands");
  std::vector<uint16_t> h((size_t)mfma_bytes(split) / 2, 0);
  const int st = femtc_mfma_tables_host(degree, nq, split, tables, h.data());
  if (st != FEMTC_OK)
    return st;
  if (hipMemcpy(operands, h.data(), h.size() * 2, hipMemcpyHostToDevice) !=
      hipSuccess)
    return fail(FEMTC_ERR_HIP, std::string("MFMA operand upload: ") +
                                   hipGetErrorString(hipGetLastError()));
  return FEMTC_OK;
}

} // extern "C"

namespace
{
// Arguments of femtc_global_apply_mfma[_noalloc] other than the lattice.
int check_mfma_args(int op, int degree, int nq, int split,
                    const void* operands, const float* x, const float* y,
                    int64_t ndofs, int nx, int ny, int nz,
                    const float* cellgeo)
{
  if (op != FEMTC_OP_MASS && op != FEMTC_OP_STIFFNESS && op != FEMTC_OP_KPM)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "op must be a femtc_op");
  if (split != 0 && split != 1)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "split must be 0 or 1");
  if (degree < 1 || degree > 7 || nq != degree + 1)
    return fail(FEMTC_ERR_DEGREE_NOT_SUPPORTED,
                "MFMA operator: degree 1..7 with nq = degree + 1");
  if (nx < 1 || ny < 1 || nz < 1)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "grid dimensions must be >= 1");
  const long long C = (long long)nx * ny * nz;
  const long long nd = ((long long)nx * degree + 1) *
                       ((long long)ny * degree + 1) *
                       ((long long)nz * degree + 1);
  if (C > 0x7ffffff0LL || nd > 0x7fffffffLL)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "grid too large");
  if (ndofs != nd)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "ndofs is not the dof count of the structured grid");
  if (!operands || !x || !y || !cellgeo)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null argument");
  return FEMTC_OK;
}
} // namespace

extern "C" {

int femtc_global_apply_mfma(int op, int degree, int nq, int split,
                            const void* operands, const float* x, float* y,
                            int64_t ndofs, int nx, int ny, int nz,
                            const float* cellgeo, float* lattice,
                            int accumulate, void* stream)
{
  int st = check_mfma_args(op, degree, nq, split, operands, x, y, ndofs, nx,
                           ny, nz, cellgeo);
  if (st != FEMTC_OK)
    return st;
  if (!lattice)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "null lattice workspace");
  hipStream_t s = (hipStream_t)stream;
  MfmaReq r;
  r.N = r.Q = degree + 1;
  r.split = split;
  r.amat = operands;
  r.w = reinterpret_cast<const float*>(
      static_cast<const unsigned char*>(operands) +
      (size_t)nmat(split != 0) * 64 * 16);
  r.x = x;
  r.lattice = lattice;
  r.cg = cellgeo;
  r.ks = op == FEMTC_OP_MASS ? 0.f : 1.f;
  r.km = op == FEMTC_OP_STIFFNESS ? 0.f : 1.f;
  r.ncx = nx;
  r.ncy = ny;
  r.ncz = nz;
  r.stream = s;
  st = launch_mfma(r);
  if (st != FEMTC_OK)
    return fail(st, "MFMA cell pass launch failed");
  st = assemble_f32(lattice, y, LatticeGrid{nx, ny, nz, degree}, accumulate, s);
  return st ? fail(st, "assembly launch failed") : FEMTC_OK;
}

} // extern "C"

// ---- the same apply without allocation: a static scratch lattice --------
// The lattice of femtc_global_apply_mfma_noalloc lives in this file-scope
// device array, one copy per GPU the library runs on, held until the
// process ends.  Host code never uses its name as an address: the device
// address comes from hipGetSymbolAddress, once per device.
namespace
{
constexpr long long SCRATCH_FLOATS = 16LL << 20; // 64 MiB
static __device__ __align__(256) float scratch[SCRATCH_FLOATS];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
float* scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device (resolved at the
// first call on that device).
int scratch_address(float** out)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
    return fail(FEMTC_ERR_NO_DEVICE, "no current HIP device");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev])
  {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FEMTC_ERR_HIP, std::string("scratch symbol address: ") +
                                     hipGetErrorString(e));
    scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = scratch_ptr[dev];
  return FEMTC_OK;
}
} // namespace

extern "C" int femtc_global_apply_mfma_noalloc(
    int op, int degree, int nq, int split, const void* operands,
    const float* x, float* y, int64_t ndofs, int nx, int ny, int nz,
    const float* cellgeo, int accumulate, void* stream)
{
  int st = check_mfma_args(op, degree, nq, split, operands, x, y, ndofs, nx,
                           ny, nz, cellgeo);
  if (st != FEMTC_OK)
    return st;
  const long long need = (long long)nx * ny * nz * (degree + 1) *
                         (degree + 1) * (degree + 1);
  if (need > SCRATCH_FLOATS)
    return fail(FEMTC_ERR_INVALID_ARGUMENT,
                "femtc_global_apply_mfma_noalloc: the grid needs " +
                    std::to_string(need) +
                    " lattice values; the static scratch holds " +
                    std::to_string(SCRATCH_FLOATS));
  float* lat = nullptr;
  if ((st = scratch_address(&lat)) != FEMTC_OK)
    return st;
  // every lattice value is written by the cell pass before the assembly
  // reads it, so the scratch needs no clearing between calls
  return femtc_global_apply_mfma(op, degree, nq, split, operands, x, y, ndofs,
                                 nx, ny, nz, cellgeo, lat, accumulate, stream);
}
