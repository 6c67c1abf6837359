// Joseph forward projection for gfx950 (rtk::JosephForwardProjectionImageFilter).
//
// Reference behaviour: include/rtkJosephForwardProjectionImageFilter.hxx:182-543
// with the default functors (include/rtkJosephForwardProjectionImageFilter.h:42-146),
// BoxShape::IsIntersectedByRay (src/rtkBoxShape.cxx:48-97) and the flat-panel ray
// iterator (include/rtkProjectionsRegionConstIteratorRayBasedWithFlatPanel.hxx:40-71).
//
// Design (CDNA4):
//  * One lane per detector pixel; a 64-lane wavefront is a 16 (u) x 4 (v) patch
//    of one view, so at every step the 64 rays sit in neighbouring voxels of
//    neighbouring slices and the bilinear gathers share L1 lines.
//  * Ray set-up, slice positions and interpolation weights are FP64, as RTK
//    computes them (CoordinateType = double): the per-ray slab/rounding logic
//    and the partial first/last steps then match the CPU filter; the samples
//    and the sum along the ray are FP32 like the reference's float pixels.
//    FP64 VALU is half the FP32 rate on CDNA4 and the loop is gather-bound.
#include "common.h"

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

// Copy of the volume with z fastest (index z + nz (y + ny x)), so that rays whose
// dominant direction is x read their in-slice (y, z) taps from neighbouring
// addresses, as z-dominant rays do from the original x-fastest layout.
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
    const JosephView & V = views[p];
    const long         pix = (long(p) * nv + jv) * nu + iu;
    double             pos[3], dir[3];
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
    const int n3[3] = { nx, ny, nz };
    // x-dominant rays use the z-fastest copy; offsets fit in 32 bits (checked on the host).
    const bool  useT = (mainDir == 0);
    const int   off3[3] = { useT ? ny * nz : 1, useT ? nz : nx, useT ? 1 : nx * ny };
    const float * volume = useT ? volT : vol;
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
    infDist = StdMax(infDist, infClip);
    supDist = StdMin(supDist, supClip);
    if (hit && supDist > 0. && infDist <= 1. && supDist > infDist)
    {
      double npt[3], fpt[3];
#pragma unroll
      for (int r = 0; r < 3; r++)
      {
        npt[r] = pos[r] + infDist * dir[r];
        fpt[r] = pos[r] + supDist * dir[r];
      }
      const int ns = int(round(npt[mainDir])); // itk::Math::rnd, half away from zero
      const int fs = int(round(fpt[mainDir]));
      int       dInf = (mainDir + 1) % 3, dSup = (mainDir + 2) % 3;
      if (dInf > dSup)
      {
        const int t = dInf;
        dInf = dSup;
        dSup = t;
      }
      const double maxx = bmax[dInf], maxy = bmax[dSup];
      const int    nxm2 = n3[dInf] - 2, nym2 = n3[dSup] - 2;
      const int     ox = off3[dInf], oy = off3[dSup];
      int           oz = off3[mainDir];
      const float * b = volume + ns * oz;
      double        residualB = ns - npt[mainDir], residualE = fpt[mainDir] - fs;
      const double  norm = 1. / dir[mainDir];
      double        stepx = dir[dInf] * norm, stepy = dir[dSup] * norm;
      if (npt[mainDir] > fpt[mainDir])
      {
        residualB = -residualB;
        residualE = -residualE;
        oz = -oz;
        stepx = -stepx;
        stepy = -stepy;
      }
      double       cx = npt[dInf] + residualB * stepx;
      double       cy = npt[dSup] + residualB * stepy;
      const double smx = sp3[dInf] * stepx, smy = sp3[dSup] * stepy, smz = sp3[mainDir];
      const double stepNorm = sqrt(smx * smx + smy * smy + smz * smz);

      // BilinearInterpolationOnBorders (.hxx:493-543): inward clamp at the box faces.
      auto border = [&](double step, const float * bb, double x, double y) -> float {
        // |residual * step| <= 0.5 keeps floor in [-1, N-1]; the clamp guards rounding.
        const int    ix = min(max(int(floor(x)), -1), nxm2 + 1), iy = min(max(int(floor(y)), -1), nym2 + 1);
        const int    idx = ix * ox + iy * oy;
        const double lx = x - ix, ly = y - iy, lxc = 1. - lx, lyc = 1. - ly;
        const int    xi = (ix < 0) ? ox : 0, yi = (iy < 0) ? oy : 0;
        const int    xs = (ix >= maxx) ? -ox : 0, ys = (iy >= maxy) ? -oy : 0;
        float        r = 0.f;
        r += float(lxc * lyc) * bb[idx + xi + yi];
        r += float(lxc * ly) * bb[idx + oy + xi + ys];
        r += float(lx * lyc) * bb[idx + ox + xs + yi];
        r += float(lx * ly) * bb[idx + ox + oy + xs + ys];
        return float(double(r) * step);
      };
      float sum = 0.f;
      if (fs == ns)
        sum += border(fabs(fpt[mainDir] - npt[mainDir]), b, cx, cy);
      else
      {
        sum += border(residualB + 0.5, b, cx, cy);
        b += oz;
        cx += stepx;
        cy += stepy;
        const int nmid = abs(fs - ns) - 1;
        // Middle steps, BilinearInterpolation (.hxx:443-477).  The in-slice
        // position is kept as an integer index plus an FP64 fraction advanced
        // by one add per step (|step| <= 1 along the dominant axis, so the index
        // moves by at most one); the weights are formed in FP32 from the FP64
        // fraction.  Middle samples lie on the segment inside the box; the clamp
        // only guards grazing rays whose rounded position leaves [0, N-2].
        int    ix = int(floor(cx)), iy = int(floor(cy));
        double fx = cx - ix, fy = cy - iy;
        const unsigned uox = unsigned(ox), uoy = unsigned(oy);
        // One step: tap offset (relative to the slice) and FP32 weights, then advance.
        auto stage = [&](int & idx, float & lx, float & ly) {
          const int cix = min(max(ix, 0), nxm2), ciy = min(max(iy, 0), nym2);
          idx = int(__umul24(unsigned(cix), uox) + __umul24(unsigned(ciy), uoy));
          lx = float(fx);
          ly = float(fy);
          fx += stepx;
          fy += stepy;
          const bool xu = fx >= 1., xd = fx < 0., yu = fy >= 1., yd = fy < 0.;
          fx = xu ? fx - 1. : (xd ? fx + 1. : fx);
          fy = yu ? fy - 1. : (yd ? fy + 1. : fy);
          ix += xu ? 1 : (xd ? -1 : 0);
          iy += yu ? 1 : (yd ? -1 : 0);
        };
        auto tap4 = [&](const float * bb, int idx, float lx, float ly) {
          const float lxc = 1.f - lx, lyc = 1.f - ly;
          return (lxc * lyc) * bb[idx] + (lx * lyc) * bb[idx + ox] + (lxc * ly) * bb[idx + oy] +
                 (lx * ly) * bb[idx + ox + oy];
        };
        int t = 0;
        for (; t < nmid; ++t)
        {
          int   id;
          float wx, wy;
          stage(id, wx, wy);
          sum += tap4(b, id, wx, wy);
          b += oz;
        }
        cx = ix + fx;
        cy = iy + fy;
        sum += border(residualE + 0.5, b, cx, cy);
      }
      cnt = abs(fs - ns) + 1;
      pout[pix] = float(double(pin ? pin[pix] : 0.f) + double(sum) * stepNorm);
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

thread_local std::map<hipStream_t, DeviceBuffer<JosephView>>          g_views;
thread_local std::map<hipStream_t, DeviceBuffer<float>>               g_volT;
thread_local std::map<hipStream_t, DeviceBuffer<unsigned long long>> g_count;

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
    RTKHIP_REQUIRE(proj_dim && vol_dim && proj_to_vol && source && vol_spacing && dev_proj_out && dev_vol,
                   RTKHIP_ERR_INVALID_ARGUMENT, "joseph_forward_project: null argument");
    const int nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
    RTKHIP_REQUIRE(nu > 0 && nv > 0 && np >= 0 && vol_dim[0] > 0 && vol_dim[1] > 0 && vol_dim[2] > 0,
                   RTKHIP_ERR_INVALID_ARGUMENT, "joseph_forward_project: sizes must be positive");
    RTKHIP_REQUIRE(np <= 65535, RTKHIP_ERR_UNSUPPORTED, "joseph_forward_project: at most 65535 views per call");
    const long long nvox = (long long)vol_dim[0] * vol_dim[1] * vol_dim[2];
    RTKHIP_REQUIRE(nvox < (1LL << 31), RTKHIP_ERR_UNSUPPORTED,
                   "joseph_forward_project: volumes are limited to 2^31 voxels (32-bit offsets)");
    RequireDevice();
    hipStream_t s = Stream(stream);
    if (samples)
      *samples = 0;
    if (np == 0)
      return;
    std::vector<JosephView> hv(np);
    for (int p = 0; p < np; p++)
    {
      for (int j = 0; j < 12; j++)
        hv[p].m[j] = proj_to_vol[12 * p + j];
      for (int j = 0; j < 3; j++)
        hv[p].src[j] = source[3 * p + j];
      hv[p].src[3] = 1.;
    }
    auto & dv = g_views[s];
    dv.Resize(np);
    RTKHIP_CHECK(hipMemcpyAsync(dv.get(), hv.data(), sizeof(JosephView) * np, hipMemcpyHostToDevice, s));
    auto & dT = g_volT[s];
    dT.Resize(size_t(nvox));
    {
      dim3 tg((vol_dim[0] + 31) / 32, (vol_dim[2] + 31) / 32, vol_dim[1]);
      hipLaunchKernelGGL(transpose_xz_kernel, tg, dim3(256), 0, s, dev_vol, dT.get(), vol_dim[0], vol_dim[1],
                         vol_dim[2]);
      RTKHIP_CHECK_LAUNCH();
    }
    // Ray clip (.hxx:216-218): understood pixel->source, hence inverted.
    const double infClip = 1. - superior_clip, supClip = 1. - inferior_clip;
    dim3         grid((nu + 15) / 16, (nv + 15) / 16, np);
    if (samples)
    {
      auto & dc = g_count[s];
      dc.Resize(1);
      RTKHIP_CHECK(hipMemsetAsync(dc.get(), 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(joseph_kernel<true>, grid, dim3(256), 0, s, dv.get(), nu, nv, dev_proj_in, dev_proj_out,
                         dev_vol, dT.get(), vol_dim[0], vol_dim[1], vol_dim[2], vol_spacing[0], vol_spacing[1],
                         vol_spacing[2], infClip, supClip, dc.get());
      RTKHIP_CHECK_LAUNCH();
      unsigned long long h = 0;
      RTKHIP_CHECK(hipMemcpyAsync(&h, dc.get(), sizeof(h), hipMemcpyDeviceToHost, s));
      RTKHIP_CHECK(hipStreamSynchronize(s));
      *samples = (long long)h;
    }
    else
    {
      hipLaunchKernelGGL(joseph_kernel<false>, grid, dim3(256), 0, s, dv.get(), nu, nv, dev_proj_in, dev_proj_out,
                         dev_vol, dT.get(), vol_dim[0], vol_dim[1], vol_dim[2], vol_spacing[0], vol_spacing[1],
                         vol_spacing[2], infClip, supClip, nullptr);
      RTKHIP_CHECK_LAUNCH();
    }
  });
}
