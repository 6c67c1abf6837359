// Joseph forward projection (rtk::JosephForwardProjectionImageFilter) and its
// splat transpose, the Joseph backprojection (rtk::JosephBackProjectionImageFilter),
// for gfx950.
//
// Reference behaviour: include/rtkJosephForwardProjectionImageFilter.hxx:182-543,
// include/rtkJosephBackProjectionImageFilter.hxx:58-412, default functors
// (include/rtkJosephForwardProjectionImageFilter.h:42-146,
// include/rtkJosephBackProjectionImageFilter.h:67-136), BoxShape::IsIntersectedByRay
// (src/rtkBoxShape.cxx:48-97) and the flat-panel ray iterator
// (include/rtkProjectionsRegionConstIteratorRayBasedWithFlatPanel.hxx:40-71).
//
// Design (CDNA4):
//  * One lane per detector pixel; a 64-lane wavefront is a 16 (u) x 4 (v) patch
//    of one view, so at every step the 64 rays sit in neighbouring voxels of
//    neighbouring slices and their taps share L1 lines.
//  * Ray set-up and the in-slice position are FP64, as RTK computes them
//    (CoordinateType = double): the slab/rounding logic and the partial first and
//    last steps match the CPU filters.  The in-slice position is an integer index
//    plus an FP64 fraction advanced by one add per step (|step| <= 1 along the
//    dominant axis); interpolation weights and sums are FP32 like the float pixels.
//  * Rays whose dominant axis is x use a z-fastest copy of the volume (gather) or a
//    z-fastest accumulator (scatter), so their in-slice (y, z) taps are neighbours
//    in memory, as they are for z-dominant rays in the x-fastest layout.
//  * The backprojector is the exact transpose of the forward kernel's sampling (same
//    set-up code, same taps and weights) with hardware FP32 atomics, which makes
//    BP_Joseph W FP_Joseph symmetric for conjugate gradient.
#include "common.h"

#include <hip/amd_detail/amd_hip_unsafe_atomics.h>

#include <cfloat>
#include <cmath>
#include <map>
#include <vector>

namespace rtkhip
{
namespace
{

struct JosephView
{
  double m[12];  // rows 0..2 of the projection-index -> volume-index matrix
  double src[4]; // source position in volume index coordinates (w unused)
};

// std::max / std::min semantics (NaN in the first argument propagates).
__device__ __forceinline__ double
StdMax(double a, double b)
{
  return (a < b) ? b : a;
}
__device__ __forceinline__ double
StdMin(double a, double b)
{
  return (b < a) ? b : a;
}

// Copy of the volume with z fastest (index z + nz (y + ny x)).
__global__ __launch_bounds__(256) void transpose_xz_kernel(const float * __restrict__ in, float * __restrict__ out,
                                                            int nx, int ny, int nz)
{
  __shared__ float tile[32][33];
  const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32, y = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 32 x 8
  for (int r = ty; r < 32; r += 8)
  {
    const int x = x0 + tx, z = z0 + r;
    if (x < nx && z < nz)
      tile[r][tx] = in[(size_t(z) * ny + y) * nx + x];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8)
  {
    const int z = z0 + tx, x = x0 + r;
    if (x < nx && z < nz)
      out[(size_t(x) * ny + y) * nz + z] = tile[tx][r];
  }
}

// out[z][y][x] += accT[x][y][z]
__global__ __launch_bounds__(256) void add_transposed_kernel(const float * __restrict__ accT, float * __restrict__ out,
                                                              int nx, int ny, int nz)
{
  __shared__ float tile[32][33];
  const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32, y = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8)
  {
    const int z = z0 + tx, x = x0 + r;
    if (x < nx && z < nz)
      tile[r][tx] = accT[(size_t(x) * ny + y) * nz + z];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8)
  {
    const int x = x0 + tx, z = z0 + r;
    if (x < nx && z < nz)
      out[(size_t(z) * ny + y) * nx + x] += tile[tx][r];
  }
}

// Everything a ray needs to walk its samples.
struct Ray
{
  int    ns, fs, nmid;   // first/last slice, number of middle steps
  int    ox, oy, oz;     // in-slice and slice offsets (signed slice offset)
  int    start;          // offset of slice ns (ns times the unsigned slice offset)
  int    nxm2, nym2;     // in-slice index limits (N - 2)
  bool   transposed;     // dominant axis x: z-fastest buffer
  double cx, cy;         // in-slice position at slice ns
  double stepx, stepy;   // in-slice advance per slice
  double maxx, maxy;     // box limits used by the border clamp
  double wB, wE, wOne;   // first/last step lengths, single-step length when fs == ns
  double stepNorm;       // |stepMM|
};

// BACK selects RTK's backprojector entry test (before clipping, supDist >= 0),
// otherwise the forward projector's (after clipping, supDist > 0, supDist > infDist).
template <bool BACK>
__device__ __forceinline__ bool
SetupRay(const JosephView & V, int iu, int jv, int p, int nx, int ny, int nz, double spx, double spy, double spz,
         double infClip, double supClip, Ray & R)
{
  double pos[3], dir[3];
#pragma unroll
  for (int r = 0; r < 3; r++)
  {
    pos[r] = V.m[4 * r + 3] + V.m[4 * r + 0] * iu + V.m[4 * r + 1] * jv + V.m[4 * r + 2] * p;
    dir[r] = -(pos[r] - V.src[r]);
  }
  int mainDir = 0;
  if (fabs(dir[1]) > fabs(dir[mainDir]))
    mainDir = 1;
  if (fabs(dir[2]) > fabs(dir[mainDir]))
    mainDir = 2;
  const int  n3[3] = { nx, ny, nz };
  const bool useT = (mainDir == 0);
  const int  off3[3] = { useT ? ny * nz : 1, useT ? nz : nx, useT ? 1 : nx * ny };
  const double sp3[3] = { spx, spy, spz };
  double       bmax[3];
#pragma unroll
  for (int r = 0; r < 3; r++)
    bmax[r] = (n3[r] - 1) * (1. - DBL_EPSILON);
  double infDist = -DBL_MAX, supDist = DBL_MAX;
  bool   hit = true;
#pragma unroll
  for (int r = 0; r < 3; r++)
  {
    if (!hit)
      break;
    if (dir[r] == 0. && (pos[r] < 0. || pos[r] > bmax[r]))
    {
      hit = false;
      break;
    }
    const double inv = 1 / dir[r];
    double       T1 = (0. - pos[r]) * inv, T2 = (bmax[r] - pos[r]) * inv;
    if (T1 > T2)
    {
      const double t = T1;
      T1 = T2;
      T2 = t;
    }
    infDist = StdMax(T1, infDist);
    supDist = StdMin(T2, supDist);
    if (infDist > supDist || supDist < 0.)
      hit = false;
  }
  if (BACK)
  {
    if (!(hit && supDist >= 0. && infDist <= 1.))
      return false;
    infDist = StdMax(infDist, infClip);
    supDist = StdMin(supDist, supClip);
  }
  else
  {
    infDist = StdMax(infDist, infClip);
    supDist = StdMin(supDist, supClip);
    if (!(hit && supDist > 0. && infDist <= 1. && supDist > infDist))
      return false;
  }
  double npt[3], fpt[3];
#pragma unroll
  for (int r = 0; r < 3; r++)
  {
    npt[r] = pos[r] + infDist * dir[r];
    fpt[r] = pos[r] + supDist * dir[r];
  }
  R.ns = int(round(npt[mainDir])); // itk::Math::rnd, half away from zero
  R.fs = int(round(fpt[mainDir]));
  R.nmid = abs(R.fs - R.ns) - 1;
  int dInf = (mainDir + 1) % 3, dSup = (mainDir + 2) % 3;
  if (dInf > dSup)
  {
    const int t = dInf;
    dInf = dSup;
    dSup = t;
  }
  R.transposed = useT;
  R.maxx = bmax[dInf];
  R.maxy = bmax[dSup];
  R.nxm2 = n3[dInf] - 2;
  R.nym2 = n3[dSup] - 2;
  R.ox = off3[dInf];
  R.oy = off3[dSup];
  R.oz = off3[mainDir];
  R.start = R.ns * R.oz; // before the sign of the step is applied
  double       residualB = R.ns - npt[mainDir], residualE = fpt[mainDir] - R.fs;
  const double norm = 1. / dir[mainDir];
  R.stepx = dir[dInf] * norm;
  R.stepy = dir[dSup] * norm;
  if (npt[mainDir] > fpt[mainDir])
  {
    residualB = -residualB;
    residualE = -residualE;
    R.oz = -R.oz;
    R.stepx = -R.stepx;
    R.stepy = -R.stepy;
  }
  R.cx = npt[dInf] + residualB * R.stepx;
  R.cy = npt[dSup] + residualB * R.stepy;
  R.wB = residualB + 0.5;
  R.wE = residualE + 0.5;
  R.wOne = fabs(fpt[mainDir] - npt[mainDir]);
  const double smx = sp3[dInf] * R.stepx, smy = sp3[dSup] * R.stepy, smz = sp3[mainDir];
  R.stepNorm = sqrt(smx * smx + smy * smy + smz * smz);
  return true;
}

// Tap offsets and FP32 weights of a first/last step (BilinearInterpolationOnBorders,
// .hxx:493-543: taps outside the box are clamped inward).
__device__ __forceinline__ void
BorderTaps(const Ray & R, double x, double y, int t[4], float w[4])
{
  // |residual * step| <= 0.5 keeps floor in [-1, N-1]; the clamp guards rounding.
  const int    ix = min(max(int(floor(x)), -1), R.nxm2 + 1), iy = min(max(int(floor(y)), -1), R.nym2 + 1);
  const int    idx = ix * R.ox + iy * R.oy;
  const double lx = x - ix, ly = y - iy, lxc = 1. - lx, lyc = 1. - ly;
  const int    xi = (ix < 0) ? R.ox : 0, yi = (iy < 0) ? R.oy : 0;
  const int    xs = (ix >= R.maxx) ? -R.ox : 0, ys = (iy >= R.maxy) ? -R.oy : 0;
  t[0] = idx + xi + yi;
  w[0] = float(lxc * lyc);
  t[1] = idx + R.oy + xi + ys;
  w[1] = float(lxc * ly);
  t[2] = idx + R.ox + xs + yi;
  w[2] = float(lx * lyc);
  t[3] = idx + R.ox + R.oy + xs + ys;
  w[3] = float(lx * ly);
}

// Middle steps: the in-slice position as integer index + FP64 fraction.  Middle
// samples lie on the segment inside the box; the clamp only guards grazing rays
// whose rounded position leaves [0, N-2].
struct Walker
{
  int    ix, iy;
  double fx, fy;
  __device__ __forceinline__ void Init(double cx, double cy)
  {
    ix = int(floor(cx));
    iy = int(floor(cy));
    fx = cx - ix;
    fy = cy - iy;
  }
  __device__ __forceinline__ int Stage(const Ray & R, float & lx, float & ly)
  {
    const int cix = min(max(ix, 0), R.nxm2), ciy = min(max(iy, 0), R.nym2);
    const int idx = int(__umul24(unsigned(cix), unsigned(R.ox)) + __umul24(unsigned(ciy), unsigned(R.oy)));
    lx = float(fx);
    ly = float(fy);
    fx += R.stepx;
    fy += R.stepy;
    const bool xu = fx >= 1., xd = fx < 0., yu = fy >= 1., yd = fy < 0.;
    fx = xu ? fx - 1. : (xd ? fx + 1. : fx);
    fy = yu ? fy - 1. : (yd ? fy + 1. : fy);
    ix += xu ? 1 : (xd ? -1 : 0);
    iy += yu ? 1 : (yd ? -1 : 0);
    return idx;
  }
};

template <bool COUNT>
__global__ __launch_bounds__(256) void joseph_kernel(const JosephView * __restrict__ views, int nu, int nv,
                                                      const float * pin, float * pout,
                                                      const float * __restrict__ vol,
                                                      const float * __restrict__ volT, int nx, int ny, int nz,
                                                      double spx, double spy, double spz, double infClip,
                                                      double supClip, unsigned long long * samples)
{
  const int iu = blockIdx.x * 16 + (threadIdx.x & 15);
  const int jv = blockIdx.y * 16 + (threadIdx.x >> 4);
  const int p = blockIdx.z;
  long long cnt = 0;
  if (iu < nu && jv < nv)
  {
    const long pix = (long(p) * nv + jv) * nu + iu;
    Ray        R;
    if (SetupRay<false>(views[p], iu, jv, p, nx, ny, nz, spx, spy, spz, infClip, supClip, R))
    {
      const float * b = (R.transposed ? volT : vol) + R.start;
      auto          border = [&](double step, const float * bb, double x, double y) -> float {
        int   t[4];
        float w[4];
        BorderTaps(R, x, y, t, w);
        float r = 0.f;
        r += w[0] * bb[t[0]];
        r += w[1] * bb[t[1]];
        r += w[2] * bb[t[2]];
        r += w[3] * bb[t[3]];
        return float(double(r) * step);
      };
      float sum = 0.f;
      if (R.fs == R.ns)
        sum += border(R.wOne, b, R.cx, R.cy);
      else
      {
        sum += border(R.wB, b, R.cx, R.cy);
        b += R.oz;
        Walker wk;
        wk.Init(R.cx + R.stepx, R.cy + R.stepy);
        for (int t = 0; t < R.nmid; ++t)
        {
          float       lx, ly;
          const int   idx = wk.Stage(R, lx, ly);
          const float lxc = 1.f - lx, lyc = 1.f - ly;
          sum += (lxc * lyc) * b[idx] + (lx * lyc) * b[idx + R.ox] + (lxc * ly) * b[idx + R.oy] +
                 (lx * ly) * b[idx + R.ox + R.oy];
          b += R.oz;
        }
        sum += border(R.wE, b, wk.ix + wk.fx, wk.iy + wk.fy);
      }
      cnt = abs(R.fs - R.ns) + 1;
      pout[pix] = float(double(pin ? pin[pix] : 0.f) + double(sum) * R.stepNorm);
    }
    else
      pout[pix] = pin ? pin[pix] : 0.f;
  }
  if (COUNT)
  {
    // Wavefront reduction, then one atomic per wavefront.
    for (int o = 32; o > 0; o >>= 1)
      cnt += __shfl_down(cnt, o, 64);
    if ((threadIdx.x & 63) == 0 && cnt)
      atomicAdd(samples, (unsigned long long)cnt);
  }
}

// Splat transpose: every sample adds value * weight * |stepMM| * stepLength to its
// four taps (BilinearSplat / BilinearSplatOnBorders, .hxx:296-370).
__global__ __launch_bounds__(256) void joseph_bp_kernel(const JosephView * __restrict__ views, int nu, int nv,
                                                         const float * __restrict__ proj, float * acc,
                                                         float * accT, int nx, int ny, int nz, double spx,
                                                         double spy, double spz, double infClip, double supClip)
{
  const int iu = blockIdx.x * 16 + (threadIdx.x & 15);
  const int jv = blockIdx.y * 16 + (threadIdx.x >> 4);
  const int p = blockIdx.z;
  if (iu >= nu || jv >= nv)
    return;
  const float value = proj[(long(p) * nv + jv) * nu + iu];
  Ray         R;
  if (value == 0.f || !SetupRay<true>(views[p], iu, jv, p, nx, ny, nz, spx, spy, spz, infClip, supClip, R))
    return;
  float *      b = (R.transposed ? accT : acc) + R.start;
  const double vn = double(value) * R.stepNorm;
  auto         border = [&](double step, float * bb, double x, double y) {
    int   t[4];
    float w[4];
    BorderTaps(R, x, y, t, w);
    const float s = float(vn * step);
#pragma unroll
    for (int q = 0; q < 4; q++)
      unsafeAtomicAdd(bb + t[q], s * w[q]);
  };
  if (R.fs == R.ns)
  {
    border(R.wOne, b, R.cx, R.cy);
    return;
  }
  border(R.wB, b, R.cx, R.cy);
  b += R.oz;
  Walker wk;
  wk.Init(R.cx + R.stepx, R.cy + R.stepy);
  const float s = float(vn);
  for (int t = 0; t < R.nmid; ++t)
  {
    float       lx, ly;
    const int   idx = wk.Stage(R, lx, ly);
    const float lxc = 1.f - lx, lyc = 1.f - ly;
    unsafeAtomicAdd(b + idx, s * (lxc * lyc));
    unsafeAtomicAdd(b + idx + R.ox, s * (lx * lyc));
    unsafeAtomicAdd(b + idx + R.oy, s * (lxc * ly));
    unsafeAtomicAdd(b + idx + R.ox + R.oy, s * (lx * ly));
    b += R.oz;
  }
  border(R.wE, b, wk.ix + wk.fx, wk.iy + wk.fy);
}

struct JosephCache
{
  std::map<hipStream_t, DeviceBuffer<JosephView>>          views;
  std::map<hipStream_t, DeviceBuffer<float>>               volT;
  std::map<hipStream_t, DeviceBuffer<unsigned long long>> count;
};
thread_local JosephCache g_joseph;

void
CheckArgs(const int proj_dim[3], const int vol_dim[3], const double * proj_to_vol, const double * source,
          const double vol_spacing[3])
{
  RTKHIP_REQUIRE(proj_dim && vol_dim && proj_to_vol && source && vol_spacing, RTKHIP_ERR_INVALID_ARGUMENT,
                 "joseph: null argument");
  RTKHIP_REQUIRE(proj_dim[0] > 0 && proj_dim[1] > 0 && proj_dim[2] >= 0 && vol_dim[0] > 0 && vol_dim[1] > 0 &&
                   vol_dim[2] > 0,
                 RTKHIP_ERR_INVALID_ARGUMENT, "joseph: sizes must be positive");
  RTKHIP_REQUIRE(proj_dim[2] <= 65535, RTKHIP_ERR_UNSUPPORTED, "joseph: at most 65535 views per call");
  RTKHIP_REQUIRE(vol_dim[1] <= 65535, RTKHIP_ERR_UNSUPPORTED, "joseph: at most 65535 volume rows");
  const long long nvox = (long long)vol_dim[0] * vol_dim[1] * vol_dim[2];
  RTKHIP_REQUIRE(nvox < (1LL << 31), RTKHIP_ERR_UNSUPPORTED, "joseph: volumes are limited to 2^31 voxels");
}

JosephView *
UploadViews(int np, const double * proj_to_vol, const double * source, hipStream_t s)
{
  std::vector<JosephView> hv(np);
  for (int p = 0; p < np; p++)
  {
    for (int j = 0; j < 12; j++)
      hv[p].m[j] = proj_to_vol[12 * p + j];
    for (int j = 0; j < 3; j++)
      hv[p].src[j] = source[3 * p + j];
    hv[p].src[3] = 1.;
  }
  auto & dv = g_joseph.views[s];
  dv.Resize(np);
  RTKHIP_CHECK(hipMemcpyAsync(dv.get(), hv.data(), sizeof(JosephView) * np, hipMemcpyHostToDevice, s));
  return dv.get();
}

} // namespace
} // namespace rtkhip

extern "C" int
rtkhip_joseph_forward_project(const int proj_dim[3], const int vol_dim[3], const double * proj_to_vol,
                              const double * source, const double vol_spacing[3], double inferior_clip,
                              double superior_clip, const float * dev_proj_in, float * dev_proj_out,
                              const float * dev_vol, long long * samples, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    CheckArgs(proj_dim, vol_dim, proj_to_vol, source, vol_spacing);
    RTKHIP_REQUIRE(dev_proj_out && dev_vol, RTKHIP_ERR_INVALID_ARGUMENT, "joseph_forward_project: null buffer");
    RequireDevice();
    hipStream_t s = Stream(stream);
    if (samples)
      *samples = 0;
    const int nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
    if (np == 0)
      return;
    JosephView * dv = UploadViews(np, proj_to_vol, source, s);
    const size_t nvox = size_t(vol_dim[0]) * vol_dim[1] * vol_dim[2];
    auto &       dT = g_joseph.volT[s];
    dT.Resize(nvox);
    dim3 tg((vol_dim[0] + 31) / 32, (vol_dim[2] + 31) / 32, vol_dim[1]);
    hipLaunchKernelGGL(transpose_xz_kernel, tg, dim3(256), 0, s, dev_vol, dT.get(), vol_dim[0], vol_dim[1],
                       vol_dim[2]);
    RTKHIP_CHECK_LAUNCH();
    // Ray clip (.hxx:216-218): understood pixel->source, hence inverted.
    const double infClip = 1. - superior_clip, supClip = 1. - inferior_clip;
    dim3         grid((nu + 15) / 16, (nv + 15) / 16, np);
    if (samples)
    {
      auto & dc = g_joseph.count[s];
      dc.Resize(1);
      RTKHIP_CHECK(hipMemsetAsync(dc.get(), 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(joseph_kernel<true>, grid, dim3(256), 0, s, dv, nu, nv, dev_proj_in, dev_proj_out, dev_vol,
                         dT.get(), vol_dim[0], vol_dim[1], vol_dim[2], vol_spacing[0], vol_spacing[1], vol_spacing[2],
                         infClip, supClip, dc.get());
      RTKHIP_CHECK_LAUNCH();
      unsigned long long h = 0;
      RTKHIP_CHECK(hipMemcpyAsync(&h, dc.get(), sizeof(h), hipMemcpyDeviceToHost, s));
      RTKHIP_CHECK(hipStreamSynchronize(s));
      *samples = (long long)h;
    }
    else
    {
      hipLaunchKernelGGL(joseph_kernel<false>, grid, dim3(256), 0, s, dv, nu, nv, dev_proj_in, dev_proj_out, dev_vol,
                         dT.get(), vol_dim[0], vol_dim[1], vol_dim[2], vol_spacing[0], vol_spacing[1], vol_spacing[2],
                         infClip, supClip, nullptr);
      RTKHIP_CHECK_LAUNCH();
    }
  });
}

extern "C" int
rtkhip_joseph_backproject(const int proj_dim[3], const int vol_dim[3], const double * proj_to_vol,
                          const double * source, const double vol_spacing[3], double inferior_clip,
                          double superior_clip, const float * dev_vol_in, float * dev_vol_out, const float * dev_proj,
                          void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    CheckArgs(proj_dim, vol_dim, proj_to_vol, source, vol_spacing);
    RTKHIP_REQUIRE(dev_vol_out && dev_proj, RTKHIP_ERR_INVALID_ARGUMENT, "joseph_backproject: null buffer");
    RequireDevice();
    hipStream_t  s = Stream(stream);
    const int    nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
    const size_t nvox = size_t(vol_dim[0]) * vol_dim[1] * vol_dim[2];
    if (dev_vol_in != dev_vol_out)
    {
      if (dev_vol_in)
        RTKHIP_CHECK(hipMemcpyAsync(dev_vol_out, dev_vol_in, nvox * sizeof(float), hipMemcpyDeviceToDevice, s));
      else
        RTKHIP_CHECK(hipMemsetAsync(dev_vol_out, 0, nvox * sizeof(float), s));
    }
    if (np == 0)
      return;
    JosephView * dv = UploadViews(np, proj_to_vol, source, s);
    auto &       dT = g_joseph.volT[s];
    dT.Resize(nvox);
    RTKHIP_CHECK(hipMemsetAsync(dT.get(), 0, nvox * sizeof(float), s));
    const double infClip = 1. - superior_clip, supClip = 1. - inferior_clip;
    dim3         grid((nu + 15) / 16, (nv + 15) / 16, np);
    hipLaunchKernelGGL(joseph_bp_kernel, grid, dim3(256), 0, s, dv, nu, nv, dev_proj, dev_vol_out, dT.get(),
                       vol_dim[0], vol_dim[1], vol_dim[2], vol_spacing[0], vol_spacing[1], vol_spacing[2], infClip,
                       supClip);
    RTKHIP_CHECK_LAUNCH();
    dim3 tg((vol_dim[0] + 31) / 32, (vol_dim[2] + 31) / 32, vol_dim[1]);
    hipLaunchKernelGGL(add_transposed_kernel, tg, dim3(256), 0, s, dT.get(), dev_vol_out, vol_dim[0], vol_dim[1],
                       vol_dim[2]);
    RTKHIP_CHECK_LAUNCH();
  });
}
