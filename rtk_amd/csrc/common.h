// Shared internals of librtkhip: status handling, HIP error mapping, device
// buffers.  Nothing here throws across the C ABI.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/rtkhip.h"

namespace rtkhip
{

// FP32 value of a stored projection pixel (FP32, FP16 or BF16 stacks).
__device__ __forceinline__ float
LoadPix(const float * p, long i)
{
  return p[i];
}
__device__ __forceinline__ float
LoadPix(const __half * p, long i)
{
  return __half2float(p[i]);
}
__device__ __forceinline__ float
LoadPix(const __hip_bfloat16 * p, long i)
{
  return __bfloat162float(p[i]);
}

// Exception carrying an rtkhip_status; converted to a return code at the ABI.
struct Error : std::runtime_error
{
  int code;
  Error(int c, const std::string & m)
    : std::runtime_error(m)
    , code(c)
  {}
};

void SetLastError(const std::string & msg);

inline hipStream_t
Stream(void * s)
{
  return reinterpret_cast<hipStream_t>(s);
}

#define RTKHIP_CHECK(expr)                                                                                           \
  do                                                                                                                 \
  {                                                                                                                  \
    hipError_t e_ = (expr);                                                                                          \
    if (e_ != hipSuccess)                                                                                            \
      throw ::rtkhip::Error(e_ == hipErrorOutOfMemory ? RTKHIP_ERR_OUT_OF_MEMORY : RTKHIP_ERR_HIP,                  \
                            std::string(#expr) + ": " + hipGetErrorString(e_));                                      \
  } while (0)

// Checks the launch that just happened on this thread.
#define RTKHIP_CHECK_LAUNCH() RTKHIP_CHECK(hipGetLastError())

#define RTKHIP_REQUIRE(cond, code, msg)                                                                              \
  do                                                                                                                 \
  {                                                                                                                  \
    if (!(cond))                                                                                                     \
      throw ::rtkhip::Error((code), (msg));                                                                          \
  } while (0)

// Runs `f` and maps exceptions to status codes + last-error text.
template <class F>
int
Guard(F && f)
{
  try
  {
    f();
    return RTKHIP_OK;
  }
  catch (const Error & e)
  {
    SetLastError(e.what());
    return e.code;
  }
  catch (const std::bad_alloc &)
  {
    SetLastError("host allocation failed");
    return RTKHIP_ERR_OUT_OF_MEMORY;
  }
  catch (const std::exception & e)
  {
    SetLastError(e.what());
    return RTKHIP_ERR_INVALID_ARGUMENT;
  }
}

// Owned device allocation (RAII).
template <class T>
class DeviceBuffer
{
public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { Resize(n); }
  ~DeviceBuffer() { Release(); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer & operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer && o) noexcept
  {
    m_Ptr = o.m_Ptr;
    m_Size = o.m_Size;
    o.m_Ptr = nullptr;
    o.m_Size = 0;
  }

  void Resize(size_t n)
  {
    if (n <= m_Size && m_Ptr)
      return;
    Release();
    if (n == 0)
      return;
    void *     p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess)
    {
      (void)hipGetLastError();
      throw Error(RTKHIP_ERR_OUT_OF_MEMORY,
                  "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
    }
    m_Ptr = static_cast<T *>(p);
    m_Size = n;
  }
  void Release()
  {
    if (m_Ptr)
      (void)hipFree(m_Ptr);
    m_Ptr = nullptr;
    m_Size = 0;
  }
  T *    get() const { return m_Ptr; }
  size_t size() const { return m_Size; }

private:
  T *    m_Ptr = nullptr;
  size_t m_Size = 0;
};

// Makes sure a device is usable before the first launch; maps "no device" to a status.
void RequireDevice();

} // namespace rtkhip
