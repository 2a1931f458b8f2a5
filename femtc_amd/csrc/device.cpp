// Minimal device-memory helpers of the C ABI, so C/C++ callers (the drivers,
// ctypes users) can keep operator inputs resident without their own HIP code.
#include "common.hpp"

namespace femtc
{
void set_last_error(const char* msg);
}

extern "C" {

int femtc_device_count(int* n)
{
  if (!n)
    return FEMTC_ERR_INVALID_ARGUMENT;
  *n = 0;
  if (hipGetDeviceCount(n) != hipSuccess || *n == 0)
  {
    *n = 0;
    femtc::set_last_error("no HIP device visible");
    return FEMTC_ERR_NO_DEVICE;
  }
  return FEMTC_OK;
}

int femtc_device_alloc(void** ptr, size_t bytes)
{
  if (!ptr)
    return FEMTC_ERR_INVALID_ARGUMENT;
  *ptr = nullptr;
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess)
  {
    *ptr = nullptr;
    femtc::set_last_error("hipMalloc failed");
    return FEMTC_ERR_OUT_OF_MEMORY;
  }
  return FEMTC_OK;
}

int femtc_device_free(void* ptr)
{
  if (!ptr)
    return FEMTC_OK;
  return femtc::hip_status(hipFree(ptr));
}

int femtc_memcpy_htod(void* dst, const void* src, size_t bytes)
{
  if (bytes == 0)
    return FEMTC_OK;
  return femtc::hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

int femtc_memcpy_dtoh(void* dst, const void* src, size_t bytes)
{
  if (bytes == 0)
    return FEMTC_OK;
  return femtc::hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

int femtc_device_synchronize(void)
{
  return femtc::hip_status(hipDeviceSynchronize());
}

} // extern "C"
