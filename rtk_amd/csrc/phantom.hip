// Analytic phantom generators (inputs of tests and benchmarks), FP64.
//
// Reference behaviour: QuadricShape (src/rtkQuadricShape.cxx:30-186),
// DrawConvexImageFilter (include/rtkDrawConvexImageFilter.hxx:41-61) and the
// RayConvexIntersectionImageFilter chain built by ProjectGeometricPhantomImageFilter
// (include/rtkRayConvexIntersectionImageFilter.hxx:57-84,
//  include/rtkProjectGeometricPhantomImageFilter.hxx:51-101).  Each shape adds
// density * chord length in mm along the whole source->pixel line; the running
// sum is rounded to the image pixel type after every shape, like the chain of
// float filters.
#include "common.h"

#include <cfloat>
#include <cmath>
#include <vector>

namespace rtkhip
{
namespace
{

struct Shape
{
  double q[11]; // A..J, density
};

__device__ __forceinline__ bool
Intersect(const double * q, const double o[3], const double d[3], double & tin, double & tout)
{
  const double Aq =
    q[0] * d[0] * d[0] + q[1] * d[1] * d[1] + q[2] * d[2] * d[2] + q[3] * d[0] * d[1] + q[4] * d[0] * d[2] +
    q[5] * d[1] * d[2];
  const double Bq = 2 * (q[0] * o[0] * d[0] + q[1] * o[1] * d[1] + q[2] * o[2] * d[2]) +
                    q[3] * (o[0] * d[1] + o[1] * d[0]) + q[4] * (o[2] * d[0] + o[0] * d[2]) +
                    q[5] * (o[1] * d[2] + o[2] * d[1]) + q[6] * d[0] + q[7] * d[1] + q[8] * d[2];
  const double Cq = q[0] * o[0] * o[0] + q[1] * o[1] * o[1] + q[2] * o[2] * o[2] + q[3] * o[0] * o[1] +
                    q[4] * o[0] * o[2] + q[5] * o[1] * o[2] + q[6] * o[0] + q[7] * o[1] + q[8] * o[2] + q[9];
  if (Aq == 0.)
    return false;
  const double disc = Bq * Bq - 4 * Aq * Cq;
  if (fabs(disc) <= 1e5 * DBL_EPSILON)
  {
    tin = tout = -Bq / (2 * Aq);
    return true;
  }
  if (disc < 0.)
    return false;
  const double sq = sqrt(disc);
  tin = (-Bq - sq) / (2 * Aq);
  tout = (-Bq + sq) / (2 * Aq);
  if (tin > tout)
  {
    const double t = tin;
    tin = tout;
    tout = t;
  }
  return true;
}

__global__ __launch_bounds__(256) void draw_kernel(const Shape * __restrict__ shapes, int ns, const float * in,
                                                    float * out, int nx, int ny, int nz, double ox, double oy,
                                                    double oz, double sx, double sy, double sz)
{
  const long idx = long(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= long(nx) * ny * nz)
    return;
  const int    i = int(idx % nx), j = int((idx / nx) % ny), k = int(idx / (long(nx) * ny));
  const double p[3] = { ox + i * sx, oy + j * sy, oz + k * sz };
  float        v = in ? in[idx] : 0.f;
  for (int s = 0; s < ns; s++)
  {
    const double * q = shapes[s].q;
    const double   f = q[0] * p[0] * p[0] + q[1] * p[1] * p[1] + q[2] * p[2] * p[2] + q[3] * p[0] * p[1] +
                     q[4] * p[0] * p[2] + q[5] * p[1] * p[2] + q[6] * p[0] + q[7] * p[1] + q[8] * p[2] + q[9];
    if (f <= 0.)
      v = float(double(v) + q[10]);
  }
  out[idx] = v;
}

struct RayView
{
  double m[12]; // projection index -> physical point (rows 0..2 of the 4x4)
  double src[4];
};

__global__ __launch_bounds__(256) void project_kernel(const RayView * __restrict__ views,
                                                       const Shape * __restrict__ shapes, int ns, const float * in,
                                                       float * out, int nu, int nv)
{
  const int iu = blockIdx.x * 64 + (threadIdx.x & 63);
  const int jv = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int p = blockIdx.z;
  if (iu >= nu || jv >= nv)
    return;
  const RayView & V = views[p];
  double          d[3];
  for (int r = 0; r < 3; r++)
    d[r] = V.m[4 * r + 3] + V.m[4 * r + 0] * iu + V.m[4 * r + 1] * jv - V.src[r];
  const double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int r = 0; r < 3; r++)
    d[r] /= n;
  const long pix = (long(p) * nv + jv) * nu + iu;
  float      acc = in ? in[pix] : 0.f;
  for (int s = 0; s < ns; s++)
  {
    double a, b;
    if (Intersect(shapes[s].q, V.src, d, a, b))
      acc = float(double(acc) + shapes[s].q[10] * (b - a));
  }
  out[pix] = acc;
}

} // namespace

// Used by the filter layer (filters.cpp), which builds the per-view rays.
int
ProjectQuadricsLaunch(int nu, int nv, int np, const double * rays16, int nshapes, const double * shapes,
                      const float * in, float * out, hipStream_t s)
{
  DeviceBuffer<Shape>   ds(nshapes > 0 ? nshapes : 1);
  DeviceBuffer<RayView> dv(np);
  std::vector<Shape>    hs(nshapes);
  for (int i = 0; i < nshapes; i++)
    for (int c = 0; c < 11; c++)
      hs[i].q[c] = shapes[11 * i + c];
  std::vector<RayView> hv(np);
  for (int p = 0; p < np; p++)
    for (int c = 0; c < 16; c++)
      (c < 12 ? hv[p].m[c] : hv[p].src[c - 12]) = rays16[16 * p + c];
  if (nshapes)
    RTKHIP_CHECK(hipMemcpyAsync(ds.get(), hs.data(), sizeof(Shape) * nshapes, hipMemcpyHostToDevice, s));
  RTKHIP_CHECK(hipMemcpyAsync(dv.get(), hv.data(), sizeof(RayView) * np, hipMemcpyHostToDevice, s));
  dim3 grid((nu + 63) / 64, (nv + 3) / 4, np);
  hipLaunchKernelGGL(project_kernel, grid, dim3(256), 0, s, dv.get(), ds.get(), nshapes, in, out, nu, nv);
  RTKHIP_CHECK_LAUNCH();
  RTKHIP_CHECK(hipStreamSynchronize(s)); // the shape/ray buffers are freed on return
  return RTKHIP_OK;
}

} // namespace rtkhip

extern "C" int
rtkhip_draw_quadrics(const rtkhip_image * vol, const float * dev_in, float * dev_out, int nshapes,
                     const double * shapes, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(vol && dev_out && (nshapes == 0 || shapes) && nshapes >= 0, RTKHIP_ERR_INVALID_ARGUMENT,
                   "draw_quadrics: bad argument");
    RTKHIP_REQUIRE(vol->size[0] > 0 && vol->size[1] > 0 && vol->size[2] > 0, RTKHIP_ERR_INVALID_ARGUMENT,
                   "draw_quadrics: sizes must be positive");
    RequireDevice();
    hipStream_t        s = Stream(stream);
    std::vector<Shape> hs(nshapes);
    for (int i = 0; i < nshapes; i++)
      for (int c = 0; c < 11; c++)
        hs[i].q[c] = shapes[11 * i + c];
    DeviceBuffer<Shape> ds(nshapes > 0 ? nshapes : 1);
    if (nshapes)
      RTKHIP_CHECK(hipMemcpyAsync(ds.get(), hs.data(), sizeof(Shape) * nshapes, hipMemcpyHostToDevice, s));
    const long n = long(vol->size[0]) * vol->size[1] * vol->size[2];
    double     o[3];
    for (int r = 0; r < 3; r++)
      o[r] = vol->origin[r] + vol->index[r] * vol->spacing[r];
    hipLaunchKernelGGL(draw_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, ds.get(), nshapes, dev_in,
                       dev_out, vol->size[0], vol->size[1], vol->size[2], o[0], o[1], o[2], vol->spacing[0],
                       vol->spacing[1], vol->spacing[2]);
    RTKHIP_CHECK_LAUNCH();
    RTKHIP_CHECK(hipStreamSynchronize(s));
  });
}
