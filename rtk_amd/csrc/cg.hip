// Conjugate-gradient vector kernels for gfx950 with FP64 scalars kept on the
// device.
//
// Reference behaviour: include/rtkConjugateGradientImageFilter.hxx:139-250 (CPU
// loop) and src/rtkCudaConjugateGradientImageFilter.cu:80-122 (one iteration
// given A*P).  The reference computes the scalars in float (cuBLAS Sdot); here
// every dot product, alpha, beta and ||R||^2 is FP64 and reduced in a fixed
// order (per-block partial sums, then one block), so results are bitwise
// reproducible run to run.  Each kernel is a single streaming pass over the
// vectors (HBM-bound); the host never reads the scalars back.
#include "common.h"

#include <cfloat>
#include <map>
#include <mutex>

namespace rtkhip
{
namespace
{

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024; // partial sums per reduction

int
Blocks(size_t n)
{
  size_t b = (n + kThreads - 1) / kThreads;
  return int(b < kMaxBlocks ? (b == 0 ? 1 : b) : kMaxBlocks);
}

__device__ __forceinline__ double
BlockSum(double v)
{
  __shared__ double sh[kThreads / 64];
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    sh[wid] = v;
  __syncthreads();
  double s = 0.;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; w++)
      s += sh[w];
  return s; // valid in thread 0
}

__global__ __launch_bounds__(kThreads) void dot_kernel(size_t n, const float * __restrict__ a,
                                                        const float * __restrict__ b, double * partial)
{
  double acc = 0.;
  for (size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += size_t(gridDim.x) * kThreads)
    acc += double(a[i]) * double(b[i]);
  const double s = BlockSum(acc);
  if (threadIdx.x == 0)
    partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void cg_init_kernel(size_t n, const float * __restrict__ b,
                                                            const float * ax, float * r, float * __restrict__ p,
                                                            double * partial)
{
  double acc = 0.;
  for (size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += size_t(gridDim.x) * kThreads)
  {
    const float ri = b[i] - ax[i];
    r[i] = ri;
    p[i] = ri;
    acc += double(ri) * double(ri);
  }
  const double s = BlockSum(acc);
  if (threadIdx.x == 0)
    partial[blockIdx.x] = s;
}

// X += alpha P; R -= alpha AP; partial ||R||^2.  alpha = state[2].
__global__ __launch_bounds__(kThreads) void cg_xr_kernel(size_t n, float * __restrict__ x, float * __restrict__ r,
                                                          const float * __restrict__ p,
                                                          const float * __restrict__ ap,
                                                          const double * __restrict__ state, double * partial)
{
  const double alpha = state[2];
  double       acc = 0.;
  for (size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += size_t(gridDim.x) * kThreads)
  {
    x[i] = float(double(x[i]) + alpha * double(p[i]));
    const float ri = float(double(r[i]) - alpha * double(ap[i]));
    r[i] = ri;
    acc += double(ri) * double(ri);
  }
  const double s = BlockSum(acc);
  if (threadIdx.x == 0)
    partial[blockIdx.x] = s;
}

// P = R + beta P.  beta = state[3].
__global__ __launch_bounds__(kThreads) void cg_p_kernel(size_t n, const float * __restrict__ r,
                                                         float * __restrict__ p, const double * __restrict__ state)
{
  const double beta = state[3];
  for (size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += size_t(gridDim.x) * kThreads)
    p[i] = float(double(r[i]) + beta * double(p[i]));
}

// Sums the partials in a fixed order.  mode 0: dst[0] = sum.
// mode 1 (alpha): state[1] = sum (pAp); state[2] = state[0] / (pAp + eps).
// mode 2 (beta):  state[3] = sum / (state[0] + eps); state[0] = sum.
__global__ __launch_bounds__(kThreads) void finalize_kernel(const double * __restrict__ partial, int np,
                                                             double * dst, int mode)
{
  double acc = 0.;
  for (int i = threadIdx.x; i < np; i += kThreads)
    acc += partial[i];
  const double s = BlockSum(acc);
  if (threadIdx.x == 0)
  {
    const double eps = double(FLT_MIN); // include/rtkConjugateGradientImageFilter.hxx:132
    if (mode == 0)
      dst[0] = s;
    else if (mode == 1)
    {
      dst[1] = s;
      dst[2] = dst[0] / (s + eps);
    }
    else
    {
      dst[3] = s / (dst[0] + eps);
      dst[0] = s;
    }
  }
}

thread_local std::map<hipStream_t, DeviceBuffer<double>> g_partials;

double *
Partials(hipStream_t s)
{
  auto & b = g_partials[s];
  b.Resize(kMaxBlocks);
  return b.get();
}

} // namespace
} // namespace rtkhip

// Partial sums of rtkhip_cg_step_noalloc: one FP64 per workgroup of the reductions,
// at most kMaxBlocks of them (8 KiB).  Each device that runs the entry point holds its
// own copy until the process ends.  Host code never passes this name to a kernel: it
// takes the device address from hipGetSymbolAddress (NoAllocScratch below).
static __device__ __align__(256) double scratch[rtkhip::kMaxBlocks];

namespace rtkhip
{
namespace
{

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device, looked up once per device with
// that device current and kept.
double *
NoAllocScratch()
{
  static std::mutex lock;
  static double *   ptr[kMaxDevices] = {};
  int               dev = 0;
  RTKHIP_CHECK(hipGetDevice(&dev));
  RTKHIP_REQUIRE(dev >= 0 && dev < kMaxDevices, RTKHIP_ERR_UNSUPPORTED, "cg_step_noalloc: device index above 63");
  std::lock_guard<std::mutex> g(lock);
  if (!ptr[dev])
  {
    void * p = nullptr;
    RTKHIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    RTKHIP_REQUIRE(p != nullptr, RTKHIP_ERR_HIP, "cg_step_noalloc: no device address for the scratch array");
    ptr[dev] = static_cast<double *>(p);
  }
  return ptr[dev];
}

// The five launches of one CG step with the partial sums in `part` (at least Blocks(n)
// doubles).
void
CgStepLaunches(size_t n, float * x, float * r, float * p, const float * ap, double * state, double * part,
               hipStream_t s)
{
  const int g = Blocks(n);
  hipLaunchKernelGGL(dot_kernel, dim3(g), dim3(kThreads), 0, s, n, p, ap, part);
  RTKHIP_CHECK_LAUNCH();
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, s, part, g, state, 1);
  RTKHIP_CHECK_LAUNCH();
  hipLaunchKernelGGL(cg_xr_kernel, dim3(g), dim3(kThreads), 0, s, n, x, r, p, ap, state, part);
  RTKHIP_CHECK_LAUNCH();
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, s, part, g, state, 2);
  RTKHIP_CHECK_LAUNCH();
  hipLaunchKernelGGL(cg_p_kernel, dim3(g), dim3(kThreads), 0, s, n, r, p, state);
  RTKHIP_CHECK_LAUNCH();
}

} // namespace
} // namespace rtkhip

extern "C" int
rtkhip_dot(size_t n, const float * dev_a, const float * dev_b, double * dev_result, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(dev_a && dev_b && dev_result, RTKHIP_ERR_INVALID_ARGUMENT, "dot: null argument");
    RequireDevice();
    hipStream_t s = Stream(stream);
    double *    part = Partials(s);
    const int   g = Blocks(n);
    hipLaunchKernelGGL(dot_kernel, dim3(g), dim3(kThreads), 0, s, n, dev_a, dev_b, part);
    RTKHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, s, part, g, dev_result, 0);
    RTKHIP_CHECK_LAUNCH();
  });
}

extern "C" int
rtkhip_cg_init(size_t n, const float * dev_b, const float * dev_ax, float * dev_r, float * dev_p,
               double * dev_state, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(dev_b && dev_ax && dev_r && dev_p && dev_state, RTKHIP_ERR_INVALID_ARGUMENT,
                   "cg_init: null argument");
    RequireDevice();
    hipStream_t s = Stream(stream);
    double *    part = Partials(s);
    const int   g = Blocks(n);
    hipLaunchKernelGGL(cg_init_kernel, dim3(g), dim3(kThreads), 0, s, n, dev_b, dev_ax, dev_r, dev_p, part);
    RTKHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, s, part, g, dev_state, 0);
    RTKHIP_CHECK_LAUNCH();
  });
}

extern "C" int
rtkhip_cg_step(size_t n, float * dev_x, float * dev_r, float * dev_p, const float * dev_ap, double * dev_state,
               void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(dev_x && dev_r && dev_p && dev_ap && dev_state, RTKHIP_ERR_INVALID_ARGUMENT,
                   "cg_step: null argument");
    RequireDevice();
    hipStream_t s = Stream(stream);
    CgStepLaunches(n, dev_x, dev_r, dev_p, dev_ap, dev_state, Partials(s), s);
  });
}

extern "C" int
rtkhip_cg_step_noalloc(size_t n, float * dev_x, float * dev_r, float * dev_p, const float * dev_ap,
                       double * dev_state, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(dev_x && dev_r && dev_p && dev_ap && dev_state, RTKHIP_ERR_INVALID_ARGUMENT,
                   "cg_step_noalloc: null argument");
    // The reductions keep one partial sum per workgroup and their grid strides over the
    // vectors with at most kMaxBlocks workgroups, so every n fits the scratch array; this
    // check guards that invariant.
    RTKHIP_REQUIRE(Blocks(n) <= kMaxBlocks, RTKHIP_ERR_UNSUPPORTED,
                   "cg_step_noalloc: " + std::to_string(Blocks(n)) + " partial sums exceed the scratch array's " +
                     std::to_string(kMaxBlocks));
    hipStream_t s = Stream(stream);
    CgStepLaunches(n, dev_x, dev_r, dev_p, dev_ap, dev_state, NoAllocScratch(), s);
  });
}
