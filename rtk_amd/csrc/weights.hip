// Short-scan (Parker) and displaced-detector weighting of projection stacks.
//
// Both filters multiply every detector row of a view by the same row of weights,
// which depends only on the column's physical position u and on the view
// (include/rtkParkerShortScanImageFilter.hxx:112-209,
//  include/rtkDisplacedDetectorImageFilter.hxx:175-323).  The kernel is a stream
// over the stack (HBM-bound: 4 B read + 4 B written per pixel): each thread owns
// four adjacent output columns of one view, evaluates their weights once in FP64
// (atan/sin, as the reference's CPU filters do) and then walks a share of the
// view's rows with 16-byte loads and stores.
#include "common.h"

#include <algorithm>
#include <cstdint>
#include <map>
#include <vector>

namespace rtkhip
{
namespace
{

constexpr double kPiD = 3.141592653589793238462643383279502884;

// Per-view parameters (the CUDA host wrappers upload sdd, source offset x,
// projection offset x, sid and, for Parker, the gantry angle:
// src/rtkCudaParkerShortScanImageFilter.cxx:89-97,
// src/rtkCudaDisplacedDetectorImageFilter.cxx:112-119).
struct RowView
{
  double sdd, sx, px, sid;
  double beta; // Parker: gantry angle - first angle, in [0, 2 pi)
};

// src/rtkThreeDCircularProjectionGeometry.cxx:641-663
__device__ __forceinline__ double
Untilted(const RowView & v, double tilted)
{
  const double sidu = sqrt(v.sid * v.sid + v.sx * v.sx);
  const double l = (tilted + v.px - v.sx) * v.sid / v.sdd + v.sx;
  const double cosa = v.sx / sidu;
  return l * fabs(v.sid) / (sidu - l * cosa);
}

enum WeightMode
{
  kParker = 0,
  kDisplacedPositive = 1, // truncated side at -u: weights rise from 0 to 2
  kDisplacedNegative = 2
};

// Weight of the column at physical coordinate u for one view.  Parker weights are
// rounded to the pixel type (the reference keeps them in an image of the output
// pixel type); displaced-detector weights stay FP64 (an itk::Image<double, 1>).
template <int MODE>
__device__ __forceinline__ double
ColumnWeight(const RowView & v, double u, double p0)
{
  const double l = Untilted(v, u);
  if constexpr (MODE == kParker)
  {
    const double delta = p0;
    const double invsid = 1. / sqrt(v.sid * v.sid + v.sx * v.sx);
    const double alpha = atan(-1. * l * invsid);
    const double beta = v.beta;
    double       w;
    if (beta <= 2 * delta - 2 * alpha)
    {
      const double s = sin((kPiD * beta) / (4 * (delta - alpha)));
      w = 2. * (s * s);
    }
    else if (beta <= kPiD - 2 * alpha)
      w = 2.;
    else if (beta <= kPiD + 2 * delta)
    {
      const double s = sin((kPiD * (kPiD + 2 * delta - beta)) / (4 * (delta + alpha)));
      w = 2. * (s * s);
    }
    else
      w = 0.;
    return double(float(w));
  }
  else
  {
    const double theta = p0;
    const double sidu = sqrt(v.sid * v.sid + v.sx * v.sx);
    double       invsid = 0., invden = 0.;
    if (sidu != 0.)
    {
      invsid = 1. / sidu;
      invden = 1. / (2. * atan(theta * invsid));
    }
    if (l <= -1. * theta)
      return MODE == kDisplacedPositive ? 0. : 2.;
    if (l >= theta)
      return MODE == kDisplacedPositive ? 2. : 0.;
    const double s = sin(kPiD * atan(l * invsid) * invden);
    return MODE == kDisplacedPositive ? s + 1. : 1. - s;
  }
}

// Streaming (non-temporal) 16-byte accesses: the stack is touched once.
typedef float Vec4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ Vec4
LoadStream(const float * p)
{
  return __builtin_nontemporal_load(reinterpret_cast<const Vec4 *>(p));
}

__device__ __forceinline__ void
StoreStream(float * p, Vec4 v)
{
  __builtin_nontemporal_store(v, reinterpret_cast<Vec4 *>(p));
}

template <int MODE>
__device__ __forceinline__ float
Apply(float x, double w)
{
  if constexpr (MODE == kParker)
    return x * float(w); // float image times float weight
  else
    return float(double(x) * w); // float pixel times double weight, rounded once
}

// grid.x = views * colBlocks * rowSplit; 256 threads; 4 output columns per thread.
// Output column j holds input column j - offset (zero where that is outside
// [0, nuIn)).  VEC: nuIn, nuOut and offset are multiples of 4 and both buffers are
// 16-byte aligned, so each thread's four columns are all inside or all outside.
template <int MODE, bool VEC>
__global__ __launch_bounds__(256) void row_weight_kernel(const RowView * __restrict__ views, int nuIn, int nuOut,
                                                         int offset, int nv, int colBlocks, int rowSplit,
                                                         double uOrigin, double uSpacing, double p0,
                                                         const float * in, float * out)
{
  const int b = blockIdx.x;
  const int cb = b % colBlocks;
  const int rs = (b / colBlocks) % rowSplit;
  const int k = b / (colBlocks * rowSplit);
  const int j0 = cb * 1024 + int(threadIdx.x) * 4;
  if (j0 >= nuOut)
    return;
  const RowView v = views[k];
  double        w[4];
  bool          inside[4];
#pragma unroll
  for (int c = 0; c < 4; c++)
  {
    const int i = j0 + c - offset;
    inside[c] = (j0 + c < nuOut) && i >= 0 && i < nuIn;
    w[c] = inside[c] ? ColumnWeight<MODE>(v, uOrigin + double(i) * uSpacing, p0) : 0.;
  }
  const size_t inView = size_t(k) * nv * nuIn, outView = size_t(k) * nv * nuOut;
  if constexpr (VEC)
  {
    const bool    any = inside[0];
    const float * src = in + inView + (j0 - offset);
    float *       dst = out + outView + j0;
    // Four rows per step: all loads issue before the stores (rows are distinct, so
    // this is safe when in == out), keeping 64 B per lane in flight.
    constexpr int U = 4;
    int           r = rs;
    for (; r + (U - 1) * rowSplit < nv; r += U * rowSplit)
    {
      Vec4 x[U];
#pragma unroll
      for (int q = 0; q < U; q++)
        x[q] = any ? LoadStream(src + size_t(r + q * rowSplit) * nuIn) : Vec4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
      for (int q = 0; q < U; q++)
        StoreStream(dst + size_t(r + q * rowSplit) * nuOut,
                    Vec4{ Apply<MODE>(x[q].x, w[0]), Apply<MODE>(x[q].y, w[1]), Apply<MODE>(x[q].z, w[2]),
                          Apply<MODE>(x[q].w, w[3]) });
    }
    for (; r < nv; r += rowSplit)
    {
      const Vec4 x = any ? LoadStream(src + size_t(r) * nuIn) : Vec4{ 0.f, 0.f, 0.f, 0.f };
      StoreStream(dst + size_t(r) * nuOut,
                  Vec4{ Apply<MODE>(x.x, w[0]), Apply<MODE>(x.y, w[1]), Apply<MODE>(x.z, w[2]), Apply<MODE>(x.w, w[3]) });
    }
  }
  else
  {
    for (int r = rs; r < nv; r += rowSplit)
    {
#pragma unroll
      for (int c = 0; c < 4; c++)
      {
        const int j = j0 + c;
        if (j >= nuOut)
          break;
        const float y = inside[c] ? Apply<MODE>(in[inView + size_t(r) * nuIn + (j - offset)], w[c]) : 0.f;
        out[outView + size_t(r) * nuOut + j] = y;
      }
    }
  }
}

struct WeightCache
{
  std::map<hipStream_t, DeviceBuffer<RowView>> views;
};
thread_local WeightCache g_weights;

void
LaunchRowWeights(int mode, const std::vector<RowView> & hv, int nuIn, int nuOut, int offset, int nv, double uOrigin,
                 double uSpacing, double p0, const float * in, float * out, hipStream_t s)
{
  const int np = int(hv.size());
  auto &    dv = g_weights.views[s];
  dv.Resize(np);
  RTKHIP_CHECK(hipMemcpyAsync(dv.get(), hv.data(), sizeof(RowView) * np, hipMemcpyHostToDevice, s));
  const int  colBlocks = (nuOut + 1023) / 1024;
  const long base = long(np) * colBlocks;
  // enough workgroups to fill 256 CUs several times over, at most one per row
  const int rowSplit = int(std::max(1L, std::min<long>(nv, (8192 + base - 1) / base)));
  const long blocks = base * rowSplit;
  RTKHIP_REQUIRE(blocks <= 2147483647L, RTKHIP_ERR_INVALID_ARGUMENT, "row weights: too many views");
  const bool vec = (nuIn % 4 == 0) && (nuOut % 4 == 0) && (offset % 4 == 0) &&
                   (reinterpret_cast<uintptr_t>(in) % 16 == 0) && (reinterpret_cast<uintptr_t>(out) % 16 == 0);
  const dim3 grid{ unsigned(blocks) }, block{ 256 };
#define RTKHIP_ROWW(M)                                                                                               \
  if (vec)                                                                                                           \
    hipLaunchKernelGGL((row_weight_kernel<M, true>), grid, block, 0, s, dv.get(), nuIn, nuOut, offset, nv, colBlocks, \
                       rowSplit, uOrigin, uSpacing, p0, in, out);                                                    \
  else                                                                                                               \
    hipLaunchKernelGGL((row_weight_kernel<M, false>), grid, block, 0, s, dv.get(), nuIn, nuOut, offset, nv,         \
                       colBlocks, rowSplit, uOrigin, uSpacing, p0, in, out);
  if (mode == kParker)
  {
    RTKHIP_ROWW(kParker)
  }
  else if (mode == kDisplacedPositive)
  {
    RTKHIP_ROWW(kDisplacedPositive)
  }
  else
  {
    RTKHIP_ROWW(kDisplacedNegative)
  }
#undef RTKHIP_ROWW
  RTKHIP_CHECK_LAUNCH();
}

} // namespace
} // namespace rtkhip

extern "C" int
rtkhip_parker_weight(const int proj_dim[3], const double * geom5, double delta, double first_angle,
                     double proj_origin_u, double proj_spacing_u, const float * dev_in, float * dev_out, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(proj_dim && geom5 && dev_in && dev_out, RTKHIP_ERR_INVALID_ARGUMENT, "parker_weight: null argument");
    const int nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
    RTKHIP_REQUIRE(nu > 0 && nv > 0 && np >= 0, RTKHIP_ERR_INVALID_ARGUMENT, "parker_weight: bad sizes");
    RequireDevice();
    if (np == 0)
      return;
    std::vector<RowView> hv(np);
    for (int p = 0; p < np; p++)
    {
      const double * g = geom5 + 5 * p;
      double         beta = g[4] - first_angle;
      if (beta < 0)
        beta += 2 * kPiD;
      hv[p] = RowView{ g[0], g[1], g[2], g[3], beta };
    }
    LaunchRowWeights(kParker, hv, nu, nu, 0, nv, proj_origin_u, proj_spacing_u, delta, dev_in, dev_out,
                     Stream(stream));
  });
}

extern "C" int
rtkhip_displaced_detector_weight(const int in_dim[3], int out_nu, int in_offset, const double * geom4, double theta,
                                 int positive_case, double proj_origin_u, double proj_spacing_u,
                                 const float * dev_in, float * dev_out, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(in_dim && geom4 && dev_in && dev_out, RTKHIP_ERR_INVALID_ARGUMENT,
                   "displaced_detector_weight: null argument");
    const int nu = in_dim[0], nv = in_dim[1], np = in_dim[2];
    RTKHIP_REQUIRE(nu > 0 && nv > 0 && np >= 0 && in_offset >= 0 && out_nu >= nu + in_offset,
                   RTKHIP_ERR_INVALID_ARGUMENT, "displaced_detector_weight: bad sizes");
    RTKHIP_REQUIRE(dev_in != dev_out || out_nu == nu, RTKHIP_ERR_INVALID_ARGUMENT,
                   "displaced_detector_weight: a padded output cannot alias the input");
    RequireDevice();
    if (np == 0)
      return;
    std::vector<RowView> hv(np);
    for (int p = 0; p < np; p++)
    {
      const double * g = geom4 + 4 * p;
      hv[p] = RowView{ g[0], g[1], g[2], g[3], 0. };
    }
    LaunchRowWeights(positive_case ? kDisplacedPositive : kDisplacedNegative, hv, nu, out_nu, in_offset, nv,
                     proj_origin_u, proj_spacing_u, theta, dev_in, dev_out, Stream(stream));
  });
}
