// Device-resident image buffers with a host mirror (the image data manager of the C ABI).
//
// Reference behaviour: RTK's CUDA filters keep their images on the device through
// itk::CudaImage and its CudaDataManager (ITKCudaCommon, an unpinned dependency that the
// reference does not vendor); they fetch device pointers with
// GetCudaDataManager()->GetGPUBufferPointer() (e.g. src/rtkCudaFDKBackProjectionImageFilter.cxx:71-73,
// include/rtkCudaForwardProjectionImageFilter.hxx:95-97,
// include/rtkCudaConjugateGradientImageFilter.hxx:58-72), and the manager copies between
// host and device only when the side being accessed is stale.  This is the same
// contract in plain C: one device allocation per buffer, an optional attached host
// array (the CPU image's pixels), two dirty flags and transfer counters, so that a
// chain of filters (forward projection, weighting, backprojection, CG update) keeps
// its images in HBM and crosses PCIe only where the caller reads or writes host memory.
#include "common.h"

#include <cstring>
#include <mutex>

struct rtkhip_buffer
{
  size_t                              bytes = 0;
  rtkhip::DeviceBuffer<unsigned char> dev;
  void *                              host = nullptr;
  bool                                hostStale = false; // the device copy holds the image
  bool                                devStale = false;  // the host copy holds the image
  long long                           uploads = 0, downloads = 0;
  std::mutex                          lock;
};

extern "C" rtkhip_buffer *
rtkhip_buffer_create(size_t bytes)
{
  rtkhip_buffer * b = nullptr;
  const int       rc = rtkhip::Guard([&] {
    RTKHIP_REQUIRE(bytes > 0, RTKHIP_ERR_INVALID_ARGUMENT, "buffer_create: empty buffer");
    rtkhip::RequireDevice();
    auto * nb = new rtkhip_buffer;
    try
    {
      nb->dev.Resize(bytes);
    }
    catch (...)
    {
      delete nb;
      throw;
    }
    nb->bytes = bytes;
    b = nb;
  });
  return rc == RTKHIP_OK ? b : nullptr;
}

extern "C" void
rtkhip_buffer_destroy(rtkhip_buffer * b)
{
  delete b;
}

extern "C" int
rtkhip_buffer_set_host(rtkhip_buffer * b, void * host)
{
  return rtkhip::Guard([&] {
    RTKHIP_REQUIRE(b, RTKHIP_ERR_INVALID_ARGUMENT, "buffer_set_host: null buffer");
    std::lock_guard<std::mutex> g(b->lock);
    b->host = host;
    // newly attached host pixels are the image; detaching leaves the device copy current
    b->devStale = host != nullptr;
    b->hostStale = false;
  });
}

extern "C" int
rtkhip_buffer_device(rtkhip_buffer * b, int write, void ** dev, void * stream)
{
  return rtkhip::Guard([&] {
    RTKHIP_REQUIRE(b && dev, RTKHIP_ERR_INVALID_ARGUMENT, "buffer_device: null argument");
    std::lock_guard<std::mutex> g(b->lock);
    hipStream_t                 s = rtkhip::Stream(stream);
    if (b->devStale && b->host)
    {
      RTKHIP_CHECK(hipMemcpyAsync(b->dev.get(), b->host, b->bytes, hipMemcpyHostToDevice, s));
      b->uploads++;
    }
    b->devStale = false;
    if (write)
      b->hostStale = b->host != nullptr;
    *dev = b->dev.get();
  });
}

extern "C" int
rtkhip_buffer_host(rtkhip_buffer * b, int write, void ** host, void * stream)
{
  return rtkhip::Guard([&] {
    RTKHIP_REQUIRE(b && host, RTKHIP_ERR_INVALID_ARGUMENT, "buffer_host: null argument");
    std::lock_guard<std::mutex> g(b->lock);
    RTKHIP_REQUIRE(b->host, RTKHIP_ERR_INVALID_ARGUMENT, "buffer_host: no host memory attached");
    hipStream_t s = rtkhip::Stream(stream);
    if (b->hostStale)
    {
      RTKHIP_CHECK(hipMemcpyAsync(b->host, b->dev.get(), b->bytes, hipMemcpyDeviceToHost, s));
      RTKHIP_CHECK(hipStreamSynchronize(s));
      b->downloads++;
    }
    b->hostStale = false;
    if (write)
      b->devStale = true;
    *host = b->host;
  });
}

extern "C" int
rtkhip_buffer_transfers(const rtkhip_buffer * b, long long * uploads, long long * downloads)
{
  return rtkhip::Guard([&] {
    RTKHIP_REQUIRE(b, RTKHIP_ERR_INVALID_ARGUMENT, "buffer_transfers: null buffer");
    if (uploads)
      *uploads = b->uploads;
    if (downloads)
      *downloads = b->downloads;
  });
}

namespace
{
__global__ void
multiply_pixels_kernel(size_t n, const float * a, const float * b, float * out)
{
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    out[i] = a[i] * b[i];
}
} // namespace

extern "C" int
rtkhip_multiply(size_t n, const float * dev_a, const float * dev_b, float * dev_out, void * stream)
{
  return rtkhip::Guard([&] {
    RTKHIP_REQUIRE(dev_a && dev_b && dev_out, RTKHIP_ERR_INVALID_ARGUMENT, "multiply: null argument");
    rtkhip::RequireDevice();
    if (n == 0)
      return;
    const unsigned g = unsigned(std::min<size_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(multiply_pixels_kernel, dim3(g), dim3(256), 0, rtkhip::Stream(stream), n, dev_a, dev_b,
                       dev_out);
    RTKHIP_CHECK_LAUNCH();
  });
}
