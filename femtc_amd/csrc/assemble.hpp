// Assembled scatter, second pass: the sum over cells of each global dof of
// the structured numbering, read from the lattice-ordered cell values the
// cell pass stored (apply_kernel.hpp, SCATTER_ASSEMBLED).
//
// Lattice layout.  Cell (ci, cj, ck) of an ncx x ncy x ncz grid, local dof
// (a, b, c) with a slowest, sits at
//     (((ci N + a) ncy + cj) N + b) (ncz N) + ck N + c,          N = P + 1,
// i.e. the global dof lattice with every point shared by two cells along an
// axis stored twice.  Global dof (gi, gj, gk) (lexicographic, gk fastest, as
// femtc_dofmap_structured numbers them) gathers from at most 2 x 2 x 2 such
// places: up to 4 lattice rows (from gi and gj), and in each row the place
// gk + floor(gk / P) and, at a point shared along k, the one before it.
//
// Work decomposition.  One wave per global row (gi, gj): the row's lattice
// rows are found once, and the wave's lanes walk gk with unit stride, so
// every load and store of the wave is contiguous; the only per-dof integer
// work is a division by the compile-time degree.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace femtc
{

// Values of the lattice buffer for an ncx x ncy x ncz grid at degree P.
inline long long lattice_size(int ncx, int ncy, int ncz, int P)
{
  const long long N = P + 1;
  return (long long)ncx * N * ncy * N * ncz * N;
}

struct LatticeGrid
{
  int ncx, ncy, ncz, P;
};

// Cells (c) and local indices (l) of global index g along one axis of n
// cells: one pair inside a cell, two at a shared point, one at the ends.
template <int P>
__device__ inline int lattice_axis(int g, int n, int (&c)[2], int (&l)[2])
{
  const int q = g / P;
  const int r = g - q * P;
  if (r != 0)
  {
    c[0] = q;
    l[0] = r;
    return 1;
  }
  int k = 0;
  if (q > 0)
  {
    c[k] = q - 1;
    l[k] = P;
    ++k;
  }
  if (q < n)
  {
    c[k] = q;
    l[k] = 0;
    ++k;
  }
  return k;
}

// The lattice rows of global row `row` = gi * (ncy P + 1) + gj: rows[2u+v]
// for the u-th place along i and the v-th along j; returns the mask of the
// rows that exist (1, 2 or 4 of them; the others repeat an existing row).
template <typename T, int P>
__device__ inline int lattice_rows(const T* __restrict__ W, int row,
                                   const LatticeGrid& L, const T* (&rows)[4])
{
  constexpr int N = P + 1;
  const int NY = L.ncy * P + 1;
  const int gi = row / NY;
  const int gj = row - gi * NY;
  int ci[2], ai[2], cj[2], aj[2];
  const int ni = lattice_axis<P>(gi, L.ncx, ci, ai);
  const int nj = lattice_axis<P>(gj, L.ncy, cj, aj);
  if (ni == 1)
  {
    ci[1] = ci[0];
    ai[1] = ai[0];
  }
  if (nj == 1)
  {
    cj[1] = cj[0];
    aj[1] = aj[0];
  }
  const long long rz = (long long)L.ncz * N;
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
      rows[2 * u + v] =
          W + ((((long long)ci[u] * N + ai[u]) * L.ncy + cj[v]) * N + aj[v]) *
                  rz;
  // bit 2u+v set when place u along i and place v along j both exist
  const int mj = nj == 2 ? 3 : 1;
  return ni == 2 ? mj | (mj << 2) : mj;
}

// p[i] for 0 <= i < 2^32 / sizeof(T), addressed as base + 32-bit unsigned
// byte offset: with a wave-uniform p the load takes p in scalar registers
// and only the offset in a vector register.
template <typename T>
__device__ inline const T& at(const T* p, int i)
{
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) +
                                     (unsigned)i * (unsigned)sizeof(T));
}
template <typename T>
__device__ inline T& at(T* p, int i)
{
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) +
                               (unsigned)i * (unsigned)sizeof(T));
}

// The rows that exist moved to the front (rows[0..n)), n returned: the mask
// of lattice_rows is 1, 3, 5 or 15, so only mask 5 moves a row.
template <typename T>
__device__ inline int lattice_rows_front(const T* (&rows)[4], int mask)
{
  if (mask == 5)
    rows[1] = rows[2];
  return mask == 15 ? 4 : mask == 1 ? 1 : 2;
}

// Sum of the cell values of global dof gk of a row from its NR existing
// lattice rows (lattice_rows_front), in the order place by place, the
// shared point before the main one.  Both offsets stay inside the row
// whatever the flags (has_prev or has_main always holds), so every load is
// issued unconditionally and all are in flight before the first is waited
// for; absent terms are then dropped by a select.
template <typename T, int P, int NR>
__device__ inline T lattice_row_sum(const T* const (&rows)[4], int gk, int ncz)
{
  const int kq = gk / P;
  const int main = gk + kq;
  const bool has_main = kq < ncz;
  const bool has_prev = gk == kq * P && kq > 0;
  const int ip = has_prev ? main - 1 : main;
  const int im = has_main ? main : ip;
  T vp[NR], vm[NR];
#pragma unroll
  for (int u = 0; u < NR; ++u)
  {
    vp[u] = at(rows[u], ip);
    vm[u] = at(rows[u], im);
  }
  T s = T(0);
#pragma unroll
  for (int u = 0; u < NR; ++u)
  {
    s += has_prev ? vp[u] : T(0);
    s += has_main ? vm[u] : T(0);
  }
  return s;
}

// f(integral_constant<int, NR>) for the row count n of lattice_rows_front
// (wave-uniform, so the branch is a scalar one).
template <typename F>
__device__ inline void with_row_count(int n, F&& f)
{
  if (n == 1)
    f(std::integral_constant<int, 1>{});
  else if (n == 2)
    f(std::integral_constant<int, 2>{});
  else
    f(std::integral_constant<int, 4>{});
}

constexpr int ASM_THREADS = 256;

// Global rows of the structured dof lattice, and dofs per row.
__host__ __device__ inline long long lattice_global_rows(const LatticeGrid& L)
{
  return ((long long)L.ncx * L.P + 1) * ((long long)L.ncy * L.P + 1);
}

// y = sum (accumulate = 0) or y += sum, one wave per global row.
template <typename T, int P>
__global__ void __launch_bounds__(ASM_THREADS)
    k_assemble(const T* __restrict__ W, T* __restrict__ y, LatticeGrid L,
               int accumulate, const int* __restrict__ skip)
{
  if (skip && *skip)
    return;
  // one row per wave, made wave-uniform: its index arithmetic and the
  // lattice row pointers live in scalar registers
  const long long row =
      (long long)blockIdx.x * (ASM_THREADS / 64) +
      __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (row >= lattice_global_rows(L))
    return; // uniform over the wave
  const int lane = threadIdx.x & 63;
  const T* rows[4];
  const int mask = lattice_rows<T, P>(W, (int)row, L, rows);
  const int nr = lattice_rows_front(rows, mask);
  const int NZ = L.ncz * P + 1;
  T* yr = y + row * NZ;
  with_row_count(nr, [&](auto NRc) {
    constexpr int NR = decltype(NRc)::value;
#pragma unroll 2
    for (int gk = lane; gk < NZ; gk += 64)
    {
      const T s = lattice_row_sum<T, P, NR>(rows, gk, L.ncz);
      at(yr, gk) = accumulate ? at(yr, gk) + s : s;
    }
  });
}

template <typename T, int P>
int launch_assemble(const T* W, T* y, const LatticeGrid& L, int accumulate,
                    const int* skip, hipStream_t s)
{
  const long long waves = lattice_global_rows(L);
  const long long grid = (waves * 64 + ASM_THREADS - 1) / ASM_THREADS;
  if (grid > 0x7fffffffLL)
    return FEMTC_ERR_INVALID_ARGUMENT;
  hipLaunchKernelGGL((k_assemble<T, P>), dim3((unsigned)grid),
                     dim3(ASM_THREADS), 0, s, W, y, L, accumulate, skip);
  return hip_status(hipGetLastError());
}

} // namespace femtc
