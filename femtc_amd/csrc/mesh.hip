// Structured dofmap, affine geometry factors and small fill kernels.
// None of these exist in the reference (its U/W are cell-local and detJ / G
// are rand(), bench_mass.cpp:86-88, bench_stiffness.cpp:86-88); they are the
// data formats on either side of the operator path.
#include <cmath>
#include <vector>

#include "common.hpp"

namespace femtc
{

void set_last_error(const char* msg);

namespace
{
__global__ void __launch_bounds__(256)
    dofmap_kernel(int nx, int ny, int nz, int P, int32_t* __restrict__ dofmap,
                  long long total)
{
  const int N = P + 1;
  const int nd = N * N * N;
  const long long Gy = (long long)ny * P + 1, Gz = (long long)nz * P + 1;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       i < total; i += (long long)gridDim.x * blockDim.x)
  {
    const long long cell = i / nd;
    const int l = (int)(i - cell * nd);
    const int a = l / (N * N), b = (l / N) % N, c = l % N;
    const long long ck = cell % nz;
    const long long cj = (cell / nz) % ny;
    const long long ci = cell / ((long long)ny * nz);
    dofmap[i] = (int32_t)(((ci * P + a) * Gy + cj * P + b) * Gz + ck * P + c);
  }
}

struct Weights1D
{
  double w[16];
};

// J^-1 of a 3x3 row-major J by cofactors (FP64); returns det J.
__device__ inline double inverse3(const double* j, double (&inv)[9])
{
  const double c00 = j[4] * j[8] - j[5] * j[7];
  const double c01 = j[5] * j[6] - j[3] * j[8];
  const double c02 = j[3] * j[7] - j[4] * j[6];
  const double det = j[0] * c00 + j[1] * c01 + j[2] * c02;
  const double id = 1.0 / det;
  inv[0] = c00 * id;
  inv[1] = (j[2] * j[7] - j[1] * j[8]) * id;
  inv[2] = (j[1] * j[5] - j[2] * j[4]) * id;
  inv[3] = c01 * id;
  inv[4] = (j[0] * j[8] - j[2] * j[6]) * id;
  inv[5] = (j[2] * j[3] - j[0] * j[5]) * id;
  inv[6] = c02 * id;
  inv[7] = (j[1] * j[6] - j[0] * j[7]) * id;
  inv[8] = (j[0] * j[4] - j[1] * j[3]) * id;
  return det;
}

// s J^-1 J^-T in the packed order (S22, S12, S02, S11, S01, S00).
template <typename T>
__device__ inline void store_packed(const double (&inv)[9], double s, T* g)
{
  auto K = [&](int r, int t) {
    return inv[3 * r] * inv[3 * t] + inv[3 * r + 1] * inv[3 * t + 1] +
           inv[3 * r + 2] * inv[3 * t + 2];
  };
  g[0] = (T)(s * K(2, 2));
  g[1] = (T)(s * K(1, 2));
  g[2] = (T)(s * K(0, 2));
  g[3] = (T)(s * K(1, 1));
  g[4] = (T)(s * K(0, 1));
  g[5] = (T)(s * K(0, 0));
}

// One thread per (cell, point): invert the per-cell constant Jacobian
// (cofactors, FP64) and scale J^-1 J^-T by |det J| w_q.
template <typename T>
__global__ void __launch_bounds__(256)
    geometry_kernel(const double* __restrict__ J, T* __restrict__ G,
                    T* __restrict__ detJw, int Q, Weights1D w1,
                    long long total)
{
  const int Q3 = Q * Q * Q;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       i < total; i += (long long)gridDim.x * blockDim.x)
  {
    const long long cell = i / Q3;
    const int q = (int)(i - cell * Q3);
    const int qa = q / (Q * Q), qb = (q / Q) % Q, qc = q % Q;
    double inv[9];
    const double det = inverse3(J + 9 * cell, inv);
    const double s = fabs(det) * w1.w[qa] * w1.w[qb] * w1.w[qc];
    if (G)
      store_packed(inv, s, G + i * 6);
    if (detJw)
      detJw[i] = (T)s;
  }
}

// One thread per cell: |det J| J^-1 J^-T (packed) and |det J|, the 7
// values the affine operator variants scale by the quadrature weights.
template <typename T>
__global__ void __launch_bounds__(256)
    cellgeo_kernel(const double* __restrict__ J, T* __restrict__ out,
                   long long num_cells)
{
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       c < num_cells; c += (long long)gridDim.x * blockDim.x)
  {
    double inv[9];
    const double det = fabs(inverse3(J + 9 * c, inv));
    store_packed(inv, det, out + c * CELLGEO);
    out[c * CELLGEO + 6] = (T)det;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    fill_kernel(T* __restrict__ y, T v, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = v;
}

unsigned grid_for(long long n)
{
  long long b = (n + 255) / 256;
  if (b > 65536)
    b = 65536;
  if (b < 1)
    b = 1;
  return (unsigned)b;
}
} // namespace

// y[0..n) = 0 with a kernel (safe inside captured graphs).
int fill_zero(void* y, int64_t n, int dtype, hipStream_t s)
{
  if (n <= 0)
    return FEMTC_OK;
  if (dtype == FEMTC_F64)
    hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n)), dim3(256), 0,
                       s, (double*)y, 0.0, (long long)n);
  else if (dtype == FEMTC_F32)
    hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(256), 0, s,
                       (float*)y, 0.0f, (long long)n);
  else
    return FEMTC_ERR_INVALID_ARGUMENT;
  return hip_status(hipGetLastError());
}

} // namespace femtc

extern "C" {

int femtc_dofmap_structured(int nx, int ny, int nz, int degree,
                            int32_t* dofmap, void* stream)
{
  if (nx < 1 || ny < 1 || nz < 1 || degree < 1 || !dofmap)
  {
    femtc::set_last_error("dofmap_structured: invalid mesh size or degree");
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  const long long gx = (long long)nx * degree + 1,
                  gy = (long long)ny * degree + 1,
                  gz = (long long)nz * degree + 1;
  if (gx * gy * gz > 0x7fffffffLL)
  {
    femtc::set_last_error("dofmap_structured: global dof count exceeds int32");
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  const long long N = degree + 1;
  const long long total = (long long)nx * ny * nz * N * N * N;
  hipLaunchKernelGGL(femtc::dofmap_kernel, dim3(femtc::grid_for(total)),
                     dim3(256), 0, (hipStream_t)stream, nx, ny, nz, degree,
                     dofmap, total);
  return femtc::hip_status(hipGetLastError());
}

int femtc_geometry_affine(int64_t num_cells, int nq, const double* J, void* G,
                          void* detJw, int dtype, void* stream)
{
  if (num_cells < 0 || nq < 1 || nq > 16 || (num_cells > 0 && !J))
  {
    femtc::set_last_error("geometry_affine: invalid arguments");
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  if (num_cells == 0 || (!G && !detJw))
    return FEMTC_OK;
  femtc::Weights1D w1{};
  std::vector<double> x(nq), w(nq);
  femtc_gauss_rule(nq, x.data(), w.data());
  for (int i = 0; i < nq; ++i)
    w1.w[i] = w[i];
  const long long total = (long long)num_cells * nq * nq * nq;
  if (dtype == FEMTC_F64)
    hipLaunchKernelGGL(femtc::geometry_kernel<double>,
                       dim3(femtc::grid_for(total)), dim3(256), 0,
                       (hipStream_t)stream, J, (double*)G, (double*)detJw, nq,
                       w1, total);
  else if (dtype == FEMTC_F32)
    hipLaunchKernelGGL(femtc::geometry_kernel<float>,
                       dim3(femtc::grid_for(total)), dim3(256), 0,
                       (hipStream_t)stream, J, (float*)G, (float*)detJw, nq,
                       w1, total);
  else
  {
    femtc::set_last_error("geometry_affine: dtype must be F64 or F32");
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  return femtc::hip_status(hipGetLastError());
}

int femtc_geometry_cell(int64_t num_cells, const double* J, void* cellgeo,
                        int dtype, void* stream)
{
  if (num_cells < 0 || (num_cells > 0 && (!J || !cellgeo)))
  {
    femtc::set_last_error("geometry_cell: invalid arguments");
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  if (num_cells == 0)
    return FEMTC_OK;
  if (dtype == FEMTC_F64)
    hipLaunchKernelGGL(femtc::cellgeo_kernel<double>,
                       dim3(femtc::grid_for(num_cells)), dim3(256), 0,
                       (hipStream_t)stream, J, (double*)cellgeo,
                       (long long)num_cells);
  else if (dtype == FEMTC_F32)
    hipLaunchKernelGGL(femtc::cellgeo_kernel<float>,
                       dim3(femtc::grid_for(num_cells)), dim3(256), 0,
                       (hipStream_t)stream, J, (float*)cellgeo,
                       (long long)num_cells);
  else
  {
    femtc::set_last_error("geometry_cell: dtype must be F64 or F32");
    return FEMTC_ERR_INVALID_ARGUMENT;
  }
  return femtc::hip_status(hipGetLastError());
}

} // extern "C"
