// Streaming probe: the practical HBM bandwidth of a read-dominated stream
// shaped like the operator apply (about five bytes read per byte written),
// the yardstick beside the 8 TB/s peak.  dst[w][i] = sum_r src[r][i] for
// nread source arrays and nwrite (0 or 1) destination arrays of n doubles,
// read as lane-contiguous 16-byte non-temporal loads, nread in flight per
// lane, grid-stride over a grid of wg_per_cu workgroups per compute unit;
// with nwrite = 0 the sums go to one value per workgroup so that the loads
// are not dead code.  On MI355X a sweep of grid sizes (2..32 workgroups per
// CU) and of normal against non-temporal loads put the best five-read,
// one-write mix at 2 workgroups per CU with non-temporal loads.
#include "common.hpp"

namespace femtc
{
namespace
{
typedef double d2 __attribute__((ext_vector_type(2)));
constexpr int SB = 256;
constexpr int MAXR = 8;

template <int NR>
__global__ void __launch_bounds__(SB)
    k_stream(const d2* __restrict__ src, d2* __restrict__ dst, int64_t n2,
             int nwrite, double* __restrict__ sink)
{
  d2 acc = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * SB;
  for (int64_t i = (int64_t)blockIdx.x * SB + threadIdx.x; i < n2; i += stride)
  {
    d2 s = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < NR; ++r)
      s += __builtin_nontemporal_load(src + r * n2 + i);
    if (nwrite)
      __builtin_nontemporal_store(s, dst + i);
    else
      acc += s;
  }
  if (!nwrite)
  {
    __shared__ double red[SB];
    red[threadIdx.x] = acc.x + acc.y;
    __syncthreads();
    for (int o = SB / 2; o > 0; o >>= 1)
    {
      if ((int)threadIdx.x < o)
        red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0)
      sink[blockIdx.x] = red[0];
  }
}
} // namespace
} // namespace femtc

using namespace femtc;

extern "C" int femtc_stream_probe(int nread, int nwrite, const double* src,
                                  double* dst, int64_t n, double* sink,
                                  int nsink, int wg_per_cu, void* stream)
{
  if (nread < 1 || nread > MAXR || (nwrite != 0 && nwrite != 1) || n < 2 ||
      n % 2 || !src || (nwrite && !dst) || (!nwrite && (!sink || nsink < 1)) ||
      wg_per_cu < 0 || wg_per_cu > 64)
    return fail(FEMTC_ERR_INVALID_ARGUMENT, "bad stream probe arguments");
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                dev);
  int grid = (wg_per_cu ? wg_per_cu : 2) * cus;
  if (!nwrite && grid > nsink)
    grid = nsink;
  const int64_t n2 = n / 2;
  hipStream_t s = (hipStream_t)stream;
  switch (nread)
  {
#define FEMTC_STREAM(R)                                                       \
  case R:                                                                     \
    hipLaunchKernelGGL(k_stream<R>, dim3(grid), dim3(SB), 0, s,               \
                       (const d2*)src, (d2*)dst, n2, nwrite, sink);           \
    break;
    FEMTC_STREAM(1)
    FEMTC_STREAM(2)
    FEMTC_STREAM(3)
    FEMTC_STREAM(4)
    FEMTC_STREAM(5)
    FEMTC_STREAM(6)
    FEMTC_STREAM(7)
    FEMTC_STREAM(8)
#undef FEMTC_STREAM
  }
  return hip_status(hipGetLastError()) ? fail(FEMTC_ERR_HIP, "probe launch")
                                       : FEMTC_OK;
}
