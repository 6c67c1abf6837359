// Voxel-driven backprojection for gfx950: plain (rtk::BackProjectionImageFilter)
// and FDK-weighted (rtk::FDKBackProjectionImageFilter).
//
// Reference behaviour: include/rtkBackProjectionImageFilter.hxx:185-475 and
// include/rtkFDKBackProjectionImageFilter.hxx:52-306 (bilinear interpolation with
// the CPU border rule, FDK weight w^2 after normalisation at the rotation centre).
//
// Design (CDNA4):
//  * Circular geometries make the detector column u and the perspective w
//    independent of the volume y index (RTK's "OptimizedBackprojectionY" case).
//    Each thread owns a column of NJ voxels along y at fixed (x, z) and keeps
//    them in registers for the whole launch; per view it computes u, w and the
//    row step once (one reciprocal) and then only steps v along the column.
//  * A wavefront covers 64 consecutive x at one z: its bilinear taps fall in
//    a narrow band of detector columns, so gathers coalesce into few L1 lines;
//    the 4 wavefronts of a workgroup cover neighbouring z and share those lines.
//  * Many views per launch: the volume is read and written once per launch,
//    the per-view matrices are wave-uniform (scalar loads).
//  * Everything else (X-path views, ITK generic interpolation, non-zero
//    y-dependence of u or w) runs in a per-voxel general kernel.
#include "common.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <map>
#include <mutex>
#include <vector>

namespace rtkhip
{
namespace
{

// Per-view coefficients, centred on the volume: with x = i - ci, y = j - cj,
// z = k - ck:  U = c[0] + c[1] x + c[2] y + c[3] z, V = c[4..7], W = c[8..11].
struct alignas(16) ViewCoef
{
  float c[12];
};

constexpr int kBlockX = 64; // one wavefront along x
constexpr int kBlockZ = 4;  // 4 wavefronts along z

// Fast column kernel: requires c[2] == c[10] == 0 (u and w independent of y)
// and the OPTIMIZED border rule for every view of the launch.
template <typename TP, bool FDK, int NJ>
__global__ __launch_bounds__(kBlockX * kBlockZ) void bp_column_kernel(const ViewCoef * __restrict__ views,
                                                                        int nviews, const TP * __restrict__ proj,
                                                                        int nu, int nv, const float * vol_in,
                                                                        float * vol_out, int nx, int ny, int nz,
                                                                        float ci, float cj, float ck)
{
  const int i = blockIdx.x * kBlockX + threadIdx.x;
  const int k = blockIdx.y * kBlockZ + threadIdx.y;
  const int j0 = blockIdx.z * NJ;
  if (i >= nx || k >= nz)
    return;
  const long   sliceStride = long(nx) * ny;
  const long   base = long(k) * sliceStride + long(j0) * nx + i;
  const int    jn = min(NJ, ny - j0);
  float        acc[NJ];
#pragma unroll
  for (int jj = 0; jj < NJ; jj++)
    acc[jj] = (jj < jn && vol_in) ? vol_in[base + long(jj) * nx] : 0.f;

  const float  x = float(i) - ci;
  const float  z = float(k) - ck;
  const float  y0 = float(j0) - cj;
  const long   projStride = long(nu) * nv;
  const float  numax = float(nu - 1);
  for (int p = 0; p < nviews; p++)
  {
    const ViewCoef & c = views[p];
    const float      w = c.c[8] + c.c[9] * x + c.c[11] * z;
    const float      rw = 1.f / w;
    const float      u = (c.c[0] + c.c[1] * x + c.c[3] * z) * rw;
    const float      uf = floorf(u);
    if (!(uf >= 0.f && uf < numax))
      continue;
    const int   ui = int(uf);
    const float fu = u - uf;
    const float vb = (c.c[4] + c.c[5] * x + c.c[7] * z + c.c[6] * y0) * rw;
    const float dv = c.c[6] * rw;
    const float wgt = FDK ? rw * rw : 1.f;
    const TP *  pp = proj + long(p) * projStride + ui;
#pragma unroll
    for (int jj = 0; jj < NJ; jj++)
    {
      const float vv = vb + float(jj) * dv;
      const float vf = floorf(vv);
      const int   vi = int(vf);
      if (vf >= 0.f && vi < nv - 1)
      {
        const float fv = vv - vf;
        const long  o = long(vi) * nu;
        const float a = LoadPix(pp, o), b = LoadPix(pp, o + nu);
        const float cc = LoadPix(pp, o + 1), d = LoadPix(pp, o + nu + 1);
        const float t0 = (1.f - fv) * a + fv * b;
        const float t1 = (1.f - fv) * cc + fv * d;
        acc[jj] += wgt * ((1.f - fu) * t0 + fu * t1);
      }
    }
  }
#pragma unroll
  for (int jj = 0; jj < NJ; jj++)
    if (jj < jn)
      vol_out[base + long(jj) * nx] = acc[jj];
}

// itk::LinearInterpolateImageFunction (2-D) on the projection, in the
// transposed frame RTK uses (first index = original v, second = original u).
template <typename TP>
__device__ __forceinline__ float
ItkLinear(const TP * pp, int nu, int nv, float tu, float tv)
{
  // tu indexes original v (size nv), tv indexes original u (size nu).
  const int n0 = nv, n1 = nu;
  int       b0 = int(floorf(tu)), b1 = int(floorf(tv));
  b0 = b0 < 0 ? 0 : b0;
  b1 = b1 < 0 ? 0 : b1;
  const float d0 = tu - float(b0), d1 = tv - float(b1);
  auto        at = [&](int a, int b) { return LoadPix(pp, long(a) * nu + b); }; // a: v row, b: u col
  const float v00 = at(b0, b1);
  if (d0 <= 0.f && d1 <= 0.f)
    return v00;
  if (d1 <= 0.f)
    return (b0 + 1 > n0 - 1) ? v00 : v00 + (at(b0 + 1, b1) - v00) * d0;
  if (d0 <= 0.f)
    return (b1 + 1 > n1 - 1) ? v00 : v00 + (at(b0, b1 + 1) - v00) * d1;
  if (b0 + 1 > n0 - 1)
    return (b1 + 1 > n1 - 1) ? v00 : v00 + (at(b0, b1 + 1) - v00) * d1;
  const float v10 = at(b0 + 1, b1);
  const float vx0 = v00 + (v10 - v00) * d0;
  if (b1 + 1 > n1 - 1)
    return vx0;
  const float v11 = at(b0 + 1, b1 + 1), v01 = at(b0, b1 + 1);
  const float vx1 = v01 + (v11 - v01) * d0;
  return vx0 + (vx1 - vx0) * d1;
}

// General kernel: per-voxel perspective, per-view border rule.
template <typename TP, bool FDK>
__global__ __launch_bounds__(256) void bp_general_kernel(const ViewCoef * __restrict__ views,
                                                          const int * __restrict__ border, int nviews,
                                                          const TP * __restrict__ proj, int nu, int nv,
                                                          const float * vol_in, float * vol_out, int nx, int ny,
                                                          int nz, float ci, float cj, float ck)
{
  const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
  const long n = long(nx) * ny * nz;
  if (idx >= n)
    return;
  const int   i = int(idx % nx);
  const int   j = int((idx / nx) % ny);
  const int   k = int(idx / (long(nx) * ny));
  const float x = float(i) - ci, y = float(j) - cj, z = float(k) - ck;
  float       acc = vol_in ? vol_in[idx] : 0.f;
  const long  projStride = long(nu) * nv;
  for (int p = 0; p < nviews; p++)
  {
    const ViewCoef & c = views[p];
    const float      w = c.c[8] + c.c[9] * x + c.c[10] * y + c.c[11] * z;
    const float      rw = 1.f / w;
    const float      u = (c.c[0] + c.c[1] * x + c.c[2] * y + c.c[3] * z) * rw;
    const float      v = (c.c[4] + c.c[5] * x + c.c[6] * y + c.c[7] * z) * rw;
    const TP *       pp = proj + long(p) * projStride;
    if (border && border[p] == RTKHIP_BORDER_ITK_LINEAR)
    {
      // Inside iff -0.5 <= idx < n - 0.5 in both directions.
      if (v >= -0.5f && v < float(nv) - 0.5f && u >= -0.5f && u < float(nu) - 0.5f)
        acc += (FDK ? rw * rw : 1.f) * ItkLinear(pp, nu, nv, v, u);
      continue;
    }
    const float uf = floorf(u), vf = floorf(v);
    if (uf >= 0.f && int(uf) < nu - 1 && vf >= 0.f && int(vf) < nv - 1)
    {
      const int   ui = int(uf), vi = int(vf);
      const float fu = u - uf, fv = v - vf;
      const long  o = long(vi) * nu + ui;
      const float a = LoadPix(pp, o), b = LoadPix(pp, o + nu), cc = LoadPix(pp, o + 1), d = LoadPix(pp, o + nu + 1);
      const float t0 = (1.f - fv) * a + fv * b;
      const float t1 = (1.f - fv) * cc + fv * d;
      acc += (FDK ? rw * rw : 1.f) * ((1.f - fu) * t0 + fu * t1);
    }
  }
  vol_out[idx] = acc;
}

// Per-(thread, stream) cache of the uploaded view coefficients.
struct CoefCache
{
  std::map<hipStream_t, DeviceBuffer<ViewCoef>> coef;
  std::map<hipStream_t, DeviceBuffer<int>>      border;
};
thread_local CoefCache g_cache;

constexpr int kNJ = 16;

template <typename TP, bool FDK>
void
Launch(bool column, const ViewCoef * dcoef, const int * dborder, int nviews, const void * proj, int nu, int nv,
       const float * vin, float * vout, int nx, int ny, int nz, float ci, float cj, float ck, hipStream_t s)
{
  const TP * p = static_cast<const TP *>(proj);
  if (column)
  {
    dim3 block(kBlockX, kBlockZ, 1);
    dim3 grid((nx + kBlockX - 1) / kBlockX, (nz + kBlockZ - 1) / kBlockZ, (ny + kNJ - 1) / kNJ);
    hipLaunchKernelGGL((bp_column_kernel<TP, FDK, kNJ>), grid, block, 0, s, dcoef, nviews, p, nu, nv, vin, vout, nx,
                       ny, nz, ci, cj, ck);
  }
  else
  {
    const long n = long(nx) * ny * nz;
    dim3       grid(unsigned((n + 255) / 256));
    hipLaunchKernelGGL((bp_general_kernel<TP, FDK>), grid, dim3(256), 0, s, dcoef, dborder, nviews, p, nu, nv, vin,
                       vout, nx, ny, nz, ci, cj, ck);
  }
  RTKHIP_CHECK_LAUNCH();
}

int
Backproject(bool fdk, const int proj_dim[3], const int vol_dim[3], const double * matrices, const int * border,
            const float * vin, float * vout, const void * proj, int dtype, void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(proj_dim && vol_dim && matrices && vout && proj, RTKHIP_ERR_INVALID_ARGUMENT,
                   "backproject: null argument");
    const int nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
    const int nx = vol_dim[0], ny = vol_dim[1], nz = vol_dim[2];
    RTKHIP_REQUIRE(nu > 0 && nv > 0 && np >= 0 && nx > 0 && ny > 0 && nz > 0, RTKHIP_ERR_INVALID_ARGUMENT,
                   "backproject: sizes must be positive");
    RTKHIP_REQUIRE(dtype == RTKHIP_F32 || dtype == RTKHIP_F16 || dtype == RTKHIP_BF16, RTKHIP_ERR_INVALID_ARGUMENT,
                   "backproject: unknown projection dtype");
    RequireDevice();
    hipStream_t s = Stream(stream);
    if (np == 0)
    {
      if (vin != vout)
        RTKHIP_CHECK(hipMemcpyAsync(vout, vin, sizeof(float) * size_t(nx) * ny * nz, hipMemcpyDeviceToDevice, s));
      return;
    }
    // Centre the matrices on the volume (smaller FP32 magnitudes) and convert.
    const double          ci = nx / 2, cj = ny / 2, ck = nz / 2;
    std::vector<ViewCoef> coef(np);
    bool                  column = true;
    for (int p = 0; p < np; p++)
    {
      const double * m = matrices + 12 * p;
      for (int r = 0; r < 3; r++)
      {
        const double t = m[4 * r + 3] + m[4 * r + 0] * ci + m[4 * r + 1] * cj + m[4 * r + 2] * ck;
        coef[p].c[4 * r + 0] = float(t);
        coef[p].c[4 * r + 1] = float(m[4 * r + 0]);
        coef[p].c[4 * r + 2] = float(m[4 * r + 1]);
        coef[p].c[4 * r + 3] = float(m[4 * r + 2]);
      }
      // RTK's Y-path test (include/rtkBackProjectionImageFilter.hxx:239-243 on
      // the transposed matrix): u and w independent of y.
      const bool yInvariant = std::abs(m[1]) < 1e-10 && std::abs(m[9]) < 1e-10;
      const bool optimized = !border || border[p] == RTKHIP_BORDER_OPTIMIZED;
      if (yInvariant)
      {
        coef[p].c[2] = 0.f;
        coef[p].c[10] = 0.f;
      }
      column = column && yInvariant && optimized;
    }
    auto & dc = g_cache.coef[s];
    dc.Resize(np);
    RTKHIP_CHECK(hipMemcpyAsync(dc.get(), coef.data(), sizeof(ViewCoef) * np, hipMemcpyHostToDevice, s));
    const int * dborder = nullptr;
    if (border && !column)
    {
      auto & db = g_cache.border[s];
      db.Resize(np);
      RTKHIP_CHECK(hipMemcpyAsync(db.get(), border, sizeof(int) * np, hipMemcpyHostToDevice, s));
      dborder = db.get();
    }
    const float fci = float(ci), fcj = float(cj), fck = float(ck);
#define RTKHIP_BP_DISPATCH(TP)                                                                                       \
  if (fdk)                                                                                                           \
    Launch<TP, true>(column, dc.get(), dborder, np, proj, nu, nv, vin, vout, nx, ny, nz, fci, fcj, fck, s);         \
  else                                                                                                               \
    Launch<TP, false>(column, dc.get(), dborder, np, proj, nu, nv, vin, vout, nx, ny, nz, fci, fcj, fck, s);
    if (dtype == RTKHIP_F32)
    {
      RTKHIP_BP_DISPATCH(float)
    }
    else if (dtype == RTKHIP_F16)
    {
      RTKHIP_BP_DISPATCH(__half)
    }
    else
    {
      RTKHIP_BP_DISPATCH(__hip_bfloat16)
    }
#undef RTKHIP_BP_DISPATCH
  });
}

} // namespace
} // namespace rtkhip

extern "C" int
rtkhip_fdk_backproject(const int proj_dim[3], const int vol_dim[3], const double * matrices, const int * border,
                       const float * dev_vol_in, float * dev_vol_out, const void * dev_proj, int proj_dtype,
                       void * stream)
{
  return rtkhip::Backproject(
    true, proj_dim, vol_dim, matrices, border, dev_vol_in, dev_vol_out, dev_proj, proj_dtype, stream);
}

extern "C" int
rtkhip_voxel_backproject(const int proj_dim[3], const int vol_dim[3], const double * matrices, const int * border,
                         const float * dev_vol_in, float * dev_vol_out, const void * dev_proj, int proj_dtype,
                         void * stream)
{
  return rtkhip::Backproject(
    false, proj_dim, vol_dim, matrices, border, dev_vol_in, dev_vol_out, dev_proj, proj_dtype, stream);
}
