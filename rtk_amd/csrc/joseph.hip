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
//    in memory, as they are for z-dominant rays in the x-fastest layout.  The
//    unit-stride in-slice axis is always walked as the first one, so each pair of
//    taps along it is one 8-byte load.
//  * The backprojector is the exact transpose of the forward kernel's sampling (same
//    set-up code, same taps and weights), summed per ray tile in LDS and added with
//    FP32 atomics, which makes BP_Joseph W FP_Joseph symmetric for conjugate
//    gradient.
#include "common.h"

#include <hip/amd_detail/amd_hip_unsafe_atomics.h>

#include <array>
#include <cfloat>
#include <climits>
#include <cmath>
#include <map>
#include <type_traits>
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

// In-slice axes of a ray whose dominant axis is mainDir, the unit-stride one first:
// x for y- and z-dominant rays (x-fastest volume), z for x-dominant rays (z-fastest
// copy).  Tap pairs along the first axis are then adjacent floats, which the
// compiler loads as one 8-byte access.  Each axis is clamped and interpolated on
// its own, so the order only changes the FP32 rounding of the 4-tap sum.
__host__ __device__ __forceinline__ void
InSliceAxes(int mainDir, int & dA, int & dB)
{
  dA = mainDir == 0 ? 2 : 0;
  dB = mainDir == 1 ? 2 : 1;
}

// Everything a ray needs to walk its samples.
struct Ray
{
  int    ns, fs, nmid;   // first/last slice, number of middle steps
  int    oy, oz;         // second in-slice offset and signed slice offset (the first is 1)
  int    start;          // offset of slice ns (ns times the unsigned slice offset)
  int    nxm2, nym2;     // in-slice index limits (N - 2)
  bool   transposed;     // dominant axis x: z-fastest buffer
  int    axis;           // dominant axis (0 x, 1 y, 2 z)
  double cx, cy;         // in-slice position at slice ns
  double stepx, stepy;   // in-slice advance per slice
  long long sx32, sy32;  // the same in 32.32 fixed point (|step| <= 1)
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
  int dInf, dSup;
  InSliceAxes(mainDir, dInf, dSup);
  R.transposed = useT;
  R.axis = mainDir;
  R.maxx = bmax[dInf];
  R.maxy = bmax[dSup];
  R.nxm2 = n3[dInf] - 2;
  R.nym2 = n3[dSup] - 2;
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
  R.sx32 = __double2ll_rn(R.stepx * 4294967296.0);
  R.sy32 = __double2ll_rn(R.stepy * 4294967296.0);
  R.cx = npt[dInf] + residualB * R.stepx;
  R.cy = npt[dSup] + residualB * R.stepy;
  R.wB = residualB + 0.5;
  R.wE = residualE + 0.5;
  R.wOne = fabs(fpt[mainDir] - npt[mainDir]);
  const double smx = sp3[dInf] * R.stepx, smy = sp3[dSup] * R.stepy, smz = sp3[mainDir];
  R.stepNorm = sqrt(smx * smx + smy * smy + smz * smz);
  return true;
}

// In-slice tap coordinates and FP32 weights of a first/last step
// (BilinearInterpolationOnBorders, .hxx:493-543: taps outside the box are clamped
// inward).
__device__ __forceinline__ void
BorderTaps2D(const Ray & R, double x, double y, int tx[4], int ty[4], float w[4])
{
  // |residual * step| <= 0.5 keeps floor in [-1, N-1]; the clamp guards rounding.
  const int    ix = min(max(int(floor(x)), -1), R.nxm2 + 1), iy = min(max(int(floor(y)), -1), R.nym2 + 1);
  const double lx = x - ix, ly = y - iy, lxc = 1. - lx, lyc = 1. - ly;
  const int    x0 = ix + ((ix < 0) ? 1 : 0), y0 = iy + ((iy < 0) ? 1 : 0);
  const int    x1 = ix + 1 - ((ix >= R.maxx) ? 1 : 0), y1 = iy + 1 - ((iy >= R.maxy) ? 1 : 0);
  tx[0] = x0;
  ty[0] = y0;
  w[0] = float(lxc * lyc);
  tx[1] = x0;
  ty[1] = y1;
  w[1] = float(lxc * ly);
  tx[2] = x1;
  ty[2] = y0;
  w[2] = float(lx * lyc);
  tx[3] = x1;
  ty[3] = y1;
  w[3] = float(lx * ly);
}

// Same taps as buffer offsets relative to the slice.
__device__ __forceinline__ void
BorderTaps(const Ray & R, double x, double y, int t[4], float w[4])
{
  int tx[4], ty[4];
  BorderTaps2D(R, x, y, tx, ty, w);
#pragma unroll
  for (int q = 0; q < 4; q++)
    t[q] = tx[q] + ty[q] * R.oy;
}

// Middle steps: the in-slice position in 32.32 fixed point, the voxel index in the
// high word and the fraction in the low word, advanced by one 64-bit integer add per
// slice (no FP64 arithmetic on the walk).  Rounding the start and the step to 2^-32
// moves a sample by at most n * 2^-33 voxel after n steps, far below the FP32
// interpolation weights' resolution.  Middle samples lie on the segment inside the
// box; the index clamp only guards grazing rays whose rounded position leaves
// [0, N-2].  The forward and back projectors share this walk, so they stay exact
// transposes of each other.
struct Walker
{
  long long X, Y;
  __device__ __forceinline__ void Init(double cx, double cy)
  {
    X = __double2ll_rn(cx * 4294967296.0);
    Y = __double2ll_rn(cy * 4294967296.0);
  }
  __device__ __forceinline__ double PosX() const { return double(X) * (1. / 4294967296.); }
  __device__ __forceinline__ double PosY() const { return double(Y) * (1. / 4294967296.); }
  __device__ __forceinline__ int Stage(const Ray & R, float & lx, float & ly)
  {
    int cix, ciy;
    Stage2D(R, cix, ciy, lx, ly);
    return cix + int(__umul24(unsigned(ciy), unsigned(R.oy)));
  }
  // Lower-left tap (cix, ciy) of the current sample, then advance one slice.
  __device__ __forceinline__ void Stage2D(const Ray & R, int & cix, int & ciy, float & lx, float & ly)
  {
    cix = min(max(int(X >> 32), 0), R.nxm2);
    ciy = min(max(int(Y >> 32), 0), R.nym2);
    lx = float(unsigned(X)) * 0x1p-32f;
    ly = float(unsigned(Y)) * 0x1p-32f;
    X += R.sx32;
    Y += R.sy32;
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
  // grid (u tiles, views, v bands): workgroups in flight share a band of detector
  // rows, whose rays stay in one slab of the volume, so the slab stays in L2
  const int  iu = blockIdx.x * 16 + (threadIdx.x & 15);
  const int  jv = blockIdx.z * 16 + (threadIdx.x >> 4);
  const int  p = blockIdx.y;
  const bool inside = iu < nu && jv < nv;
  const long pix = (long(p) * nv + jv) * nu + iu;
  long long  cnt = 0;
  Ray        R;
  const bool hit = inside && SetupRay<false>(views[p], iu, jv, p, nx, ny, nz, spx, spy, spz, infClip, supClip, R);
  auto       border = [&](double step, const float * bb, double x, double y) -> float {
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
  auto middle = [&](Walker & wk, const float * bb) -> float {
    float       lx, ly;
    const int   idx = wk.Stage(R, lx, ly);
    const float lxc = 1.f - lx, lyc = 1.f - ly;
    return (lxc * lyc) * bb[idx] + (lx * lyc) * bb[idx + 1] + (lxc * ly) * bb[idx + R.oy] + (lx * ly) * bb[idx + 1 + R.oy];
  };
  float sum = 0.f;

  // The wavefront's rays that share its most common dominant axis and walking
  // direction step through absolute slice indices together, so at every step the
  // 64 lanes read neighbouring voxels of one slice (a ray-relative step count would
  // put lanes that entered the volume at different slices a slice stride apart).
  // Each ray still sums its own samples in its own order.
  const int key = hit ? 2 * R.axis + (R.oz < 0 ? 1 : 0) : -1;
  int       best = 0, bestCount = -1;
  for (int q = 0; q < 6; q++)
  {
    const int c = __popcll(__ballot(key == q));
    if (c > bestCount)
    {
      best = q;
      bestCount = c;
    }
  }
  const bool lock = hit && key == best;
  int        klo = lock ? min(R.ns, R.fs) : INT_MAX, khi = lock ? max(R.ns, R.fs) : INT_MIN;
  for (int o = 32; o > 0; o >>= 1)
  {
    klo = min(klo, __shfl_xor(klo, o, 64));
    khi = max(khi, __shfl_xor(khi, o, 64));
  }
  if (klo <= khi)
  {
    const int     axis = best >> 1, dirSign = (best & 1) ? -1 : 1;
    const float * buf = axis == 0 ? volT : vol;
    const long    absOz = axis == 0 ? long(ny) * nz : (axis == 1 ? long(nx) : long(nx) * ny);
    const int     kfirst = dirSign > 0 ? klo : khi;
    const int     nsl = khi - klo + 1;
    const int     t0 = lock ? (R.ns - kfirst) * dirSign : INT_MAX;
    const int     tEnd = lock ? t0 + R.nmid + 1 : INT_MIN; // R.nmid = -1 when fs == ns
    Walker        wk;
    for (int s = 0; s < nsl; ++s)
    {
      if (s >= t0 && s <= tEnd)
      {
        const float * bb = buf + (kfirst + s * dirSign) * absOz;
        const int     t = s - t0;
        if (t == 0)
        {
          if (R.fs == R.ns)
            sum += border(R.wOne, bb, R.cx, R.cy);
          else
          {
            sum += border(R.wB, bb, R.cx, R.cy);
            wk.Init(R.cx + R.stepx, R.cy + R.stepy);
          }
        }
        else if (t <= R.nmid)
          sum += middle(wk, bb);
        else
          sum += border(R.wE, bb, wk.PosX(), wk.PosY());
      }
    }
  }
  if (hit && !lock)
  {
    // rays of a minority axis or direction: their own walk
    const float * b = (R.transposed ? volT : vol) + R.start;
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
        sum += middle(wk, b);
        b += R.oz;
      }
      sum += border(R.wE, b, wk.PosX(), wk.PosY());
    }
  }
  if (hit)
  {
    cnt = abs(R.fs - R.ns) + 1;
    pout[pix] = float(double(pin ? pin[pix] : 0.f) + double(sum) * R.stepNorm);
  }
  else if (inside)
    pout[pix] = pin ? pin[pix] : 0.f;
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
// four taps (BilinearSplat / BilinearSplatOnBorders, .hxx:296-370).  One ray,
// straight to the accumulator with FP32 atomics.
__device__ __forceinline__ void
SplatRay(const Ray & R, float value, float * acc, float * accT)
{
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
    unsafeAtomicAdd(b + idx + 1, s * (lx * lyc));
    unsafeAtomicAdd(b + idx + R.oy, s * (lxc * ly));
    unsafeAtomicAdd(b + idx + 1 + R.oy, s * (lx * ly));
    b += R.oz;
  }
  border(R.wE, b, wk.PosX(), wk.PosY());
}

// Tiled splat.  A workgroup is a 16 x 16 tile of detector pixels of one view.  Rays
// of the tile that share the tile's most common dominant axis and walking direction
// step through the slices in lockstep; each slice's taps are summed in an LDS window
// and the window is then added to the accumulator with one global FP32 atomic per
// touched voxel, instead of one per ray tap: a 16 x 16 bundle covers about 10 x 10
// voxels of a slice, so global atomic traffic drops by about 10x.  The window holds
// 64-bit fixed point (ds_add_u64; LDS float atomics run at a fraction of the integer
// rate on gfx950): each tap's FP32 contribution, computed exactly as the per-ray path
// computes it, is scaled by a power of two chosen from the tile's largest |value| *
// |stepMM| (an upper bound of every contribution, with room for 2^11 of them per
// voxel), so the window sums are exact to 2^-50 of that bound and are rounded to FP32
// once, at the flush.  Rows are padded to 33 cells to spread LDS banks.  The window
// is double-buffered so a slice is flushed while the next one is accumulated, with
// one barrier per slice.  The window of slice k is placed from the four corner rays
// of the tile: a central projection maps the pixel rectangle onto a convex
// quadrilateral of the slice plane, so the corners bound every ray's position; its
// width is a convex function of k, so checking the first and last slice bounds all.
// The placements of 256 slices are computed at once, one slice per lane, while the
// previous 256 slices are walked.  Slices whose footprint fits 16 x 16 cells use (and
// flush) only that corner of the window, one cell per lane.  A middle step's four
// taps are a 2 x 2 block of cells: one window test and one LDS address, the other
// three taps at fixed offsets.  Taps that still fall outside (rounding), rays of a
// minority axis or direction, and tiles whose footprint exceeds the window or that
// hold a non-finite value take the per-ray path.  Every tap lies inside the volume,
// so a cell with a non-zero sum is a voxel of the slice.
// round(v * scale) as a two's-complement 64-bit integer, for |v * scale| < 2^51:
// adding 1.5 * 2^52 leaves the rounded integer (round to nearest even, as
// __double2ll_rn) in the low mantissa bits, and subtracting the constant's bit
// pattern recovers it, in one FMA and one 64-bit subtraction.
__device__ __forceinline__ unsigned long long
ToFixed(float v, double scale)
{
  constexpr double kMagic = 6755399441055744.0; // 1.5 * 2^52
  const double     t = fma(double(v), scale, kMagic);
  return (unsigned long long)(__double_as_longlong(t) - __double_as_longlong(kMagic));
}

constexpr int kTileWin = 32;    // LDS window edge, in voxels
constexpr int kTileStride = 33; // LDS row stride, in cells
constexpr int kTileCells = kTileWin * kTileStride;
constexpr int kTileDummy = kTileStride + 2; // cells after the window that absorb outside taps
constexpr int kPlaceChunk = 256;            // slices placed per pass, one per lane

template <typename TP>
__global__ __launch_bounds__(256) void joseph_bp_tile_kernel(const JosephView * __restrict__ views, int nu, int nv,
                                                              const TP * __restrict__ proj, float * acc,
                                                              float * accT, int nx, int ny, int nz, double spx,
                                                              double spy, double spz, double infClip,
                                                              double supClip)
{
  // cells after the window receive the taps that fall outside it (their global
  // atomics are issued separately), so the LDS adds need no branch; they are never
  // flushed
  __shared__ unsigned long long win[2][kTileCells + kTileDummy];
  __shared__ int    vote[6];
  __shared__ int    maxBits; // largest |value * stepNorm| of the participating rays, as FP32 bits
  __shared__ int    krange[2]; // min and max slice over the tile's participating rays
  __shared__ float  line[4][4]; // corner c: in-slice x = line[c][0] + k line[c][1], y likewise
  __shared__ int    cornerOk;
  __shared__ int4   wnd[2][kPlaceChunk]; // window origin and edge per slice, two chunks

  const int tid = threadIdx.x;
  // grid (u tiles, views, v bands), as the forward projector: the band's slab of the
  // accumulator stays in L2 while all views add to it
  const int u0 = blockIdx.x * 16, v0 = blockIdx.z * 16, p = blockIdx.y;
  const int iu = u0 + (tid & 15), jv = v0 + (tid >> 4);
  for (int c = tid; c < 2 * (kTileCells + kTileDummy); c += 256)
    (&win[0][0])[c] = 0ull;
  if (tid < 6)
    vote[tid] = 0;
  if (tid == 0)
  {
    krange[0] = INT_MAX;
    krange[1] = INT_MIN;
    cornerOk = 1;
    maxBits = 0;
  }
  __syncthreads();

  const JosephView & V = views[p];
  Ray                R;
  float              value = 0.f;
  bool               active = false;
  if (iu < nu && jv < nv)
  {
    value = LoadPix(proj, (long(p) * nv + jv) * nu + iu);
    active = value != 0.f && SetupRay<true>(V, iu, jv, p, nx, ny, nz, spx, spy, spz, infClip, supClip, R);
  }
  const int key = active ? 2 * R.axis + (R.oz < 0 ? 1 : 0) : -1;
  if (active)
    atomicAdd(&vote[key], 1);
  __syncthreads();
  int best = 0;
  for (int q = 1; q < 6; q++)
    best = vote[q] > vote[best] ? q : best;
  best = __builtin_amdgcn_readfirstlane(best); // workgroup-uniform: scalar registers
  const int  axis = best >> 1, dirSign = (best & 1) ? -1 : 1;
  const bool part = active && key == best;
  if (part)
  {
    atomicMin(&krange[0], min(R.ns, R.fs));
    atomicMax(&krange[1], max(R.ns, R.fs));
    // non-negative FP32 bit patterns order like the values; NaN and Inf sort above
    atomicMax(&maxBits, __float_as_int(fabsf(float(double(value) * R.stepNorm))));
  }
  int dInf, dSup;
  InSliceAxes(axis, dInf, dSup);
  if (tid < 4)
  {
    // corner pixels of the part of the tile inside the detector
    const int    cu = (tid & 1) ? min(u0 + 15, nu - 1) : u0, cv = (tid & 2) ? min(v0 + 15, nv - 1) : v0;
    double       pos[3], dir[3];
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
      pos[r] = V.m[4 * r + 3] + V.m[4 * r + 0] * cu + V.m[4 * r + 1] * cv + V.m[4 * r + 2] * p;
      dir[r] = V.src[r] - pos[r];
    }
    const double dm = dir[axis];
    if (!(dm * dirSign > 0.)) // walking direction is the sign of the dominant component
      cornerOk = 0;
    else
    {
      const double bx = dir[dInf] / dm, by = dir[dSup] / dm;
      // FP32 is enough to place a window with a one-voxel margin; taps that still
      // fall outside take the global path
      line[tid][0] = float(pos[dInf] - pos[axis] * bx);
      line[tid][1] = float(bx);
      line[tid][2] = float(pos[dSup] - pos[axis] * by);
      line[tid][3] = float(by);
    }
  }
  __syncthreads();
  const int   kmin = __builtin_amdgcn_readfirstlane(krange[0]), kmax = __builtin_amdgcn_readfirstlane(krange[1]);
  const float maxc = __int_as_float(maxBits);
  // tiles whose largest contribution is below 2^-70 keep the per-ray path: the FP32
  // flush factor 2^(emax - 50) would leave the normal range
  bool        tiled = cornerOk && kmin <= kmax && maxc <= FLT_MAX && maxc >= 0x1p-70f;
  auto      window = [&](int k, int & wx0, int & wy0, float & spanx, float & spany) {
    float lox = FLT_MAX, hix = -FLT_MAX, loy = FLT_MAX, hiy = -FLT_MAX;
#pragma unroll
    for (int c = 0; c < 4; c++)
    {
      const float x = line[c][0] + k * line[c][1], y = line[c][2] + k * line[c][3];
      lox = fmin(lox, x);
      hix = fmax(hix, x);
      loy = fmin(loy, y);
      hiy = fmax(hiy, y);
    }
    wx0 = int(floor(lox)) - 1;
    wy0 = int(floor(loy)) - 1;
    spanx = hix - lox;
    spany = hiy - loy;
  };
  if (tiled)
  {
    int   wx, wy;
    float sx0, sy0, sx1, sy1;
    window(kmin, wx, wy, sx0, sy0);
    window(kmax, wx, wy, sx1, sy1);
    // taps reach floor(hi) + 1, the window starts at floor(lo) - 1: span + 4 cells
    tiled = fmaxf(sx0, sx1) + 5.f <= kTileWin && fmaxf(sy0, sy1) + 5.f <= kTileWin;
  }
  if (!tiled)
  {
    if (active)
      SplatRay(R, value, acc, accT);
    return; // uniform across the workgroup
  }
  if (active && !part)
    SplatRay(R, value, acc, accT);

  // Lockstep walk over the tile's slices in the walking direction.
  const int    nsl = kmax - kmin + 1;
  const int    kfirst = dirSign > 0 ? kmin : kmax;
  const bool   useT = axis == 0;
  float *      buf = useT ? accT : acc;
  const int    n3[3] = { nx, ny, nz };
  const int    off3[3] = { useT ? ny * nz : 1, useT ? nz : nx, useT ? 1 : nx * ny };
  const int    oy = off3[dSup], oz = off3[axis]; // off3[dInf] == 1
  const int    NX = n3[dInf], NY = n3[dSup];
  const int    r0 = part ? (R.ns - kfirst) * dirSign : INT_MAX; // this ray's first and last tile step
  const int    r1 = part ? (R.fs - kfirst) * dirSign : INT_MIN;
  const double vn = part ? double(value) * R.stepNorm : 0.;
  // every contribution is at most maxc * (1 + rounding): 2^11 of them stay below 2^61
  int          emax;
  frexp(double(maxc), &emax);
  const double toFixed = ldexp(1., 50 - emax);
  const double vnFixed = vn * toFixed; // |w * vnFixed| < 2^50 for weights w <= 1
  const float  fromFixed32 = ldexpf(1.f, emax - 50);
  Walker       wk;
  int          pwx = 0, pwy = 0; // window origin of the previous slice
  int          plim = kTileWin;  // cells per window edge flushed for the previous slice
  // Lane l places slice c * 256 + l of chunk c in wnd[c & 1][l] while chunk c - 1 is
  // walked (at its first slice): that buffer was last read before the barrier that
  // ended chunk c - 2.  The window edge is 16 when the slice footprint fits.
  auto place = [&](int s1) {
    if (s1 < nsl)
    {
      int   ox, oy2;
      float sx, sy;
      window(kfirst + s1 * dirSign, ox, oy2, sx, sy);
      wnd[(s1 / kPlaceChunk) & 1][s1 % kPlaceChunk] =
        make_int4(ox, oy2, (sx + 5.f <= 16.f && sy + 5.f <= 16.f) ? 16 : kTileWin, 0);
    }
  };
  place(tid);
  __syncthreads();
  for (int s = 0; s <= nsl; ++s)
  {
    const int k = kfirst + s * dirSign;
    int       wx = 0, wy = 0;
    int       lim = kTileWin; // window edge used for this slice
    if (s < nsl)
    {
      const int4 wv = wnd[(s / kPlaceChunk) & 1][s % kPlaceChunk];
      wx = __builtin_amdgcn_readfirstlane(wv.x);
      wy = __builtin_amdgcn_readfirstlane(wv.y);
      lim = __builtin_amdgcn_readfirstlane(wv.z);
      if (s % kPlaceChunk == 0)
        place(s + kPlaceChunk + tid);
      if (s >= r0 && s <= r1)
      {
        unsigned long long * w = win[s & 1];
        float *              g = buf + size_t(k) * oz;
        const int            t = s - r0;
        if (t == 0 || t > R.nmid)
        {
          // first or last step: clamped taps, tested one by one (rare)
          int    tx[4], ty[4];
          float  tw[4];
          double step, x, y;
          if (t == 0)
          {
            step = (R.fs == R.ns) ? R.wOne : R.wB;
            x = R.cx;
            y = R.cy;
            wk.Init(R.cx + R.stepx, R.cy + R.stepy);
          }
          else
          {
            step = R.wE;
            x = wk.PosX();
            y = wk.PosY();
          }
          BorderTaps2D(R, x, y, tx, ty, tw);
          const float sc = float(vn * step);
          float       v[4];
          int         outside = 0;
#pragma unroll
          for (int q = 0; q < 4; q++)
          {
            const int  a = tx[q] - wx, b = ty[q] - wy;
            const bool in = unsigned(a) < unsigned(lim) && unsigned(b) < unsigned(lim);
            v[q] = sc * tw[q];
            atomicAdd(&w[in ? int(__umul24(unsigned(b), unsigned(kTileStride))) + a : kTileCells],
                      ToFixed(v[q], toFixed));
            outside |= in ? 0 : 1 << q;
          }
          if (outside) // rounding at the window edge: rare
          {
#pragma unroll
            for (int q = 0; q < 4; q++)
              if (outside & (1 << q))
                unsafeAtomicAdd(g + tx[q] + ty[q] * oy, v[q]);
          }
        }
        else
        {
          // middle step: taps (cx, cy), (cx, cy + 1), (cx + 1, cy), (cx + 1, cy + 1)
          int   cx, cy;
          float lx, ly;
          wk.Stage2D(R, cx, cy, lx, ly);
          // the weights in FP32, their product with value * |stepMM| rounded once, to
          // the fixed-point grid (the per-ray path rounds it to FP32 first)
          const float lxc = 1.f - lx, lyc = 1.f - ly;
          const float w0 = lxc * lyc, w1 = lxc * ly, w2 = lx * lyc, w3 = lx * ly;
          const int   a = cx - wx, b = cy - wy;
          const bool  in = unsigned(a) < unsigned(lim - 1) && unsigned(b) < unsigned(lim - 1);
          unsigned long long * c = w + (in ? int(__umul24(unsigned(b), unsigned(kTileStride))) + a : kTileCells);
          atomicAdd(c, ToFixed(w0, vnFixed));
          atomicAdd(c + kTileStride, ToFixed(w1, vnFixed));
          atomicAdd(c + 1, ToFixed(w2, vnFixed));
          atomicAdd(c + kTileStride + 1, ToFixed(w3, vnFixed));
          if (!in) // rounding at the window edge: rare
          {
            const float sc = float(vn);
            float *     gc = g + cx + cy * oy;
            unsafeAtomicAdd(gc, sc * w0);
            unsafeAtomicAdd(gc + oy, sc * w1);
            unsafeAtomicAdd(gc + 1, sc * w2);
            unsafeAtomicAdd(gc + oy + 1, sc * w3);
          }
        }
      }
    }
    if (s > 0)
    {
      // flush slice s - 1; lanes run along the unit-stride in-slice axis (window x)
      const int k1 = k - dirSign;
      unsigned long long * w = win[(s - 1) & 1];
      float *              g = buf + size_t(k1) * oz;
      // a 16 x 16 window is one cell per lane; a 32 x 32 one four.  Only taps make a
      // cell non-zero, and taps are voxels of the slice: the range test is a guard.
      // Offsets fit 32 bits (volumes below 2^31 voxels).
      auto flushCell = [&](int i, int j) {
        const int                cell = int(__umul24(unsigned(j), unsigned(kTileStride))) + i;
        const unsigned long long q = w[cell];
        w[cell] = 0ull;
        const unsigned X = unsigned(pwx + i), Y = unsigned(pwy + j);
        if (q != 0ull && X < unsigned(NX) && Y < unsigned(NY))
        {
          const unsigned off = X + Y * unsigned(oy);
          // hi * 2^32 + lo in FP32 (one more rounding than an exact conversion, far
          // below the FP32 atomic's own): no FP64 on the flush
          const float    hi = float(int(q >> 32)), lo = float(unsigned(q));
          unsafeAtomicAdd(g + off, fmaf(hi, fromFixed32 * 4294967296.f, lo * fromFixed32));
        }
      };
      if (plim == 16)
        flushCell(tid & 15, tid >> 4);
      else
      {
#pragma unroll
        for (int c = tid; c < kTileWin * kTileWin; c += 256)
          flushCell(c & (kTileWin - 1), c / kTileWin);
      }
    }
    pwx = wx;
    pwy = wy;
    plim = lim;
    __syncthreads();
  }
}

// ---------------------------------------------------- gather backprojection --
// The same backprojection computed voxel by voxel, without atomics, for geometries
// whose detector columns are vertical lines in the volume (no in-plane or out-of-plane
// tilt: a ray's horizontal position depends on its column only, its height on its row
// only, and every ray of a column has the same dominant axis, x or z).
//
// A ray of column iu walks slices k along its dominant axis A; at slice k its in-slice
// position is (X, y): X along the other horizontal axis, a Moebius function of iu, and
// y, affine in the row jv.  Its Joseph splat adds s_k v |stepMM| T(X - h) T(y - j) to
// voxel (k, j, h) of the slice (T(d) = max(0, 1 - |d|), the bilinear weight of a tap),
// with s_k = 1 on middle steps and the partial weights wB / wE / wOne on its first and
// last slices, whose positions are clamped into the box (BilinearSplatOnBorders).  So a
// voxel receives from the rays passing within one voxel of it in its slice: about
// 2 / (ray spacing) columns times as many rows.  A voxel whose three indices lie in
// [2, N - 3] only ever receives middle steps: a first or last step lies on a slice face
// (k = 0 or N - 1) or within half a voxel of a side face.  Those voxels need only the
// ray values; the two outer layers (the "shell") also read each ray's first and last
// slice and weights.  Rays whose dominant axis is not their column's (y-dominant rays
// of steep cones, ties decided differently in FP64) and rays carrying non-finite values
// are splatted by the per-ray path instead, so the result is RTK's for every ray.

constexpr int kGTileU = 64, kGTileV = 64, kGStride = 65; // staged ray values: columns x rows
constexpr int kGRows = 5;                                // candidate rows per voxel and column
constexpr int kGNJ = 16;                                 // voxels per thread along y

struct GatherView
{
  float  pu[4], pv[4], pw[4]; // centred volume index (x', y', z', 1) -> detector (u w, v w, w)
  double y0, y1, sy;          // y' of row jv at depth t: (y0 + y1 jv)(1 - t) + sy t
  int    hasX, hasZ;          // columns walking along x / along z exist in this view
  int    pad[2];
};

struct GatherColumn
{
  double h0, h1; // in-slice horizontal coordinate X' = h0 + k' h1 at centred slice k'
  double t0, t1; // ray parameter t = t0 + k' t1 (0 at the pixel, 1 at the source)
};

// Per-ray data of the gather: value * |stepMM| (0 for rays the gather must not take);
// the same value beside the first/last slice (16 bits each) and their weights.
// Views [p0, p0 + gridDim.y) of the call; gval/gend/gw and colAxis are indexed by the
// view's position in the chunk.
template <typename TP>
__global__ __launch_bounds__(256) void joseph_gather_prep_kernel(
  const JosephView * __restrict__ views, int p0, const signed char * __restrict__ colAxis, int nu, int nv,
  const TP * __restrict__ proj, float * acc, float * accT, int nx, int ny, int nz, double spx, double spy,
  double spz, double infClip, double supClip, float * __restrict__ gval, uint2 * __restrict__ gend,
  float2 * __restrict__ gw)
{
  const int iu = blockIdx.x * 16 + (threadIdx.x & 15);
  const int jv = blockIdx.z * 16 + (threadIdx.x >> 4);
  const int pl = blockIdx.y, p = p0 + pl;
  if (iu >= nu || jv >= nv)
    return;
  const size_t pix = (size_t(pl) * nv + jv) * nu + iu;
  const float  value = LoadPix(proj, (long(p) * nv + jv) * nu + iu);
  Ray          R;
  float        vn = 0.f;
  unsigned     e = 0u;
  float2       w = make_float2(0.f, 0.f);
  if (value != 0.f && SetupRay<true>(views[p], iu, jv, p, nx, ny, nz, spx, spy, spz, infClip, supClip, R))
  {
    if (R.axis == int(colAxis[size_t(pl) * nu + iu]) && isfinite(value))
    {
      vn = float(double(value) * R.stepNorm);
      e = unsigned(R.ns) | (unsigned(R.fs) << 16);
      w = make_float2(float(R.fs == R.ns ? R.wOne : R.wB), float(R.wE));
    }
    else
      SplatRay(R, value, acc, accT);
  }
  gval[pix] = vn;
  gend[pix] = make_uint2(__float_as_uint(vn), e); // the value again, so the shell reads one 8-byte word
  gw[pix] = w;
}

// Rectangle of rays [u0, u0 + W) x [v0, v0 + H) that can reach the voxel block
// [xlo, xhi] x [ylo, yhi] x [zlo, zhi] (centred, already widened by one voxel): the
// rays crossing the widened block, bounded by the central projection of its corners.
// W <= 0 when the rectangle does not fit the tile (the block then reads its rays from
// global memory for this view) or no ray reaches it.
__device__ __forceinline__ int4
GatherRect(const GatherView & V, float xlo, float xhi, float ylo, float yhi, float zlo, float zhi, int nu, int nv)
{
  float umin = FLT_MAX, umax = -FLT_MAX, vmin = FLT_MAX, vmax = -FLT_MAX, smin = 1.f, smax = -1.f;
#pragma unroll
  for (int q = 0; q < 8; q++)
  {
    const float x = (q & 1) ? xhi : xlo, y = (q & 2) ? yhi : ylo, z = (q & 4) ? zhi : zlo;
    const float w = V.pw[3] + V.pw[0] * x + V.pw[1] * y + V.pw[2] * z;
    const float rw = 1.f / w;
    const float u = (V.pu[3] + V.pu[0] * x + V.pu[1] * y + V.pu[2] * z) * rw;
    const float v = (V.pv[3] + V.pv[0] * x + V.pv[1] * y + V.pv[2] * z) * rw;
    const float sgn = (w > 0.f) ? 1.f : (w < 0.f ? -1.f : 0.f);
    umin = fminf(umin, u);
    umax = fmaxf(umax, u);
    vmin = fminf(vmin, v);
    vmax = fmaxf(vmax, v);
    smin = fminf(smin, sgn);
    smax = fmaxf(smax, sgn);
  }
  if (!(smin == smax && smin != 0.f) || !(umin <= umax) || !(vmin <= vmax) || !(umax - umin < 1e6f) ||
      !(vmax - vmin < 1e6f))
    return make_int4(0, 0, -1, 0);
  // columns outside the detector hold no rays; rows are staged with kGRows rows of
  // margin on both sides, rows outside the detector as zeros, so that every row a
  // voxel reads (up to kGRows from its first candidate) lies in the tile
  const int u0 = max(int(floorf(fmaxf(umin, -8.f))) - 1, 0) & ~3;
  const int u1 = min(int(floorf(fminf(umax, float(nu) + 8.f))) + 2, nu - 1);
  const int v0 = int(floorf(fmaxf(vmin, -1e6f))) - 1 - kGRows;
  const int v1 = int(floorf(fminf(vmax, 1e6f))) + 2 + kGRows;
  if (u1 < u0 || v1 < 0 || v0 > nv - 1)
    return make_int4(0, 0, 0, 0);
  const int W = u1 - u0 + 1, H = v1 - v0 + 1;
  return make_int4(u0, v0, (W > kGTileU || H > kGTileV) ? -1 : W, H);
}

// Step weight of slice k for a ray with first/last slices (ns, fs) packed in e and
// their weights w: 0 outside [ns, fs], 1 inside, wB / wE / wOne on the end slices.
__device__ __forceinline__ float
StepWeight(int k, unsigned e, float2 w)
{
  const int ns = int(e & 0xffffu), fs = int(e >> 16);
  if (k < min(ns, fs) || k > max(ns, fs))
    return 0.f;
  if (k == ns)
    return w.x;
  return k == fs ? w.y : 1.f;
}

// Thread = a column of kGNJ voxels along y at (x, z); workgroup = 16 x 16 columns.
// Per view the block's rays are staged in LDS (double-buffered, one barrier per view,
// the next view's loads in flight during the current view's arithmetic).  Columns
// whose x and z lie in [2, N - 3] are computed here: their voxels with y in [2, N - 3]
// receive middle steps only; the two outer layers along y (a choice uniform across the
// workgroup) also read each candidate ray's first/last slices and weights from global
// memory.  Columns of the two outer layers along x or z are joseph_gather_shell_kernel's.
template <int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void joseph_gather_kernel(
  const GatherView * __restrict__ gviews, const GatherColumn * __restrict__ gcols,
  const signed char * __restrict__ colAxis, int nviews, int nu, int nv, const float * __restrict__ gval,
  const uint2 * __restrict__ gend, const float2 * __restrict__ gw, const int * __restrict__ blocks, float * vol,
  int nx, int ny, int nz)
{
  __shared__ float        tile[2][kGTileV * kGStride];
  __shared__ GatherColumn cols[2][kGTileU];
  __shared__ signed char  caxis[2][kGTileU];
  __shared__ int4         rects[2][64];
  const int               t = threadIdx.x;
  const int               blk = blocks[blockIdx.x]; // packed (bx, bz, by)
  const int               bx0 = (blk & 0x3ff) * 16, bz0 = ((blk >> 10) & 0x3ff) * 16, j0 = (blk >> 20) * kGNJ;
  const int               i = bx0 + (t & 15), k = bz0 + (t >> 4);
  // columns of the two outer layers along x or z hold only shell voxels
  const bool              live = i >= 2 && i <= nx - 3 && k >= 2 && k <= nz - 3;
  const float             cx = 0.5f * float(nx - 1), cy = 0.5f * float(ny - 1), cz = 0.5f * float(nz - 1);
  const float             xc = float(i) - cx, zc = float(k) - cz, j0c = float(j0) - cy;
  // widened block (one voxel on every side), centred
  const float xlo = float(bx0 - 1) - cx, xhi = float(min(bx0 + 16, nx)) - cx;
  const float zlo = float(bz0 - 1) - cz, zhi = float(min(bz0 + 16, nz)) - cz;
  const float ylo = float(j0 - 1) - cy, yhi = float(min(j0 + kGNJ, ny)) - cy;
  float       acc[kGNJ];
#pragma unroll
  for (int jj = 0; jj < kGNJ; jj++)
    acc[jj] = 0.f;
  // the block holds voxels of the outer layers along y (rows 0, 1, ny - 2, ny - 1)
  const bool yshell = j0 < 2 || j0 + kGNJ > ny - 2;
  float      ash[4] = {0.f, 0.f, 0.f, 0.f};
  const size_t viewRays = size_t(nu) * nv;

  // staging of view p's rectangle: lane = 4 consecutive rays of one row
  const int sc = 4 * (t & 15), sr = t >> 4; // 16 rows per pass, 4 passes
  float     sreg[16];
  auto      load = [&](int p, int4 r) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int e = 0; e < 4; e++)
      {
        const int row = sr + 16 * q, c = sc + e, jv = r.y + row;
        sreg[4 * q + e] = (r.z > 0 && row < r.w && c < r.z && jv >= 0 && jv < nv)
                            ? gval[size_t(p) * viewRays + size_t(jv) * nu + (r.x + c)]
                            : 0.f;
      }
  };
  auto store = [&](int b, int p, int4 r) {
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (sr + 16 * q < kGTileV)
#pragma unroll
        for (int e = 0; e < 4; e++)
          tile[b][(sr + 16 * q) * kGStride + sc + e] = sreg[4 * q + e];
    if (t < kGTileU && r.z > 0 && t < r.z)
    {
      cols[b][t] = gcols[size_t(p) * nu + r.x + t];
      caxis[b][t] = colAxis[size_t(p) * nu + r.x + t];
    }
  };
  const int wave = t >> 6;
  auto      place = [&](int chunk) {
    const int q = chunk * 64 + (t & 63);
    if (q < nviews)
      rects[chunk & 1][t & 63] = GatherRect(gviews[q], xlo, xhi, ylo, yhi, zlo, zhi, nu, nv);
  };
  auto rectOf = [&](int q) {
    const int4 r = rects[(q / 64) & 1][q % 64];
    return make_int4(__builtin_amdgcn_readfirstlane(r.x), __builtin_amdgcn_readfirstlane(r.y),
                     __builtin_amdgcn_readfirstlane(r.z), __builtin_amdgcn_readfirstlane(r.w));
  };
  if (wave == 0)
    place(0);
  __syncthreads();
  int4 cur = rectOf(0);
  load(0, cur);
  store(0, 0, cur);
  __syncthreads();

  const double xd = double(xc), zd = double(zc);
  for (int p = 0; p < nviews; p++)
  {
    const int b = p & 1;
    int4      nxt = make_int4(0, 0, 0, 0);
    if (p + 1 < nviews)
    {
      nxt = rectOf(p + 1);
      load(p + 1, nxt);
    }
    if (wave == 0 && p % 64 == 0 && p + 64 < nviews)
      place(p / 64 + 1);
    if (live && cur.z != 0)
    {
      const GatherView & V = gviews[p];
      const bool         staged = cur.z > 0;
#pragma unroll 1
      for (int ax = 0; ax < 2; ax++)
      {
        const int axis = ax == 0 ? 0 : 2;
        if ((axis == 0 && !V.hasX) || (axis == 2 && !V.hasZ))
          continue;
        // slice coordinate k' and in-slice horizontal coordinate h' of this voxel column
        const float  kc = axis == 0 ? xc : zc, hc = axis == 0 ? zc : xc;
        const double kd = axis == 0 ? xd : zd;
        // candidate columns: the rays whose X' lies within one voxel of h' at slice k'
        const float xa = axis == 0 ? kc : hc - 1.f, za = axis == 0 ? hc - 1.f : kc;
        const float xb = axis == 0 ? kc : hc + 1.f, zb = axis == 0 ? hc + 1.f : kc;
        const float ua = (V.pu[3] + V.pu[0] * xa + V.pu[2] * za) * __builtin_amdgcn_rcpf(V.pw[3] + V.pw[0] * xa + V.pw[2] * za);
        const float ub = (V.pu[3] + V.pu[0] * xb + V.pu[2] * zb) * __builtin_amdgcn_rcpf(V.pw[3] + V.pw[0] * xb + V.pw[2] * zb);
        int         iuLo = max(int(ceilf(fminf(ua, ub) - 1e-3f)), 0);
        int         iuHi = min(int(floorf(fmaxf(ua, ub) + 1e-3f)), nu - 1);
        if (staged)
        {
          iuLo = max(iuLo, cur.x);
          iuHi = min(iuHi, cur.x + cur.z - 1);
        }
        // the staged and the global-memory paths are separate loops (the view's choice is
        // uniform), so the column tables are read with LDS or global loads, not flat ones
        auto columns = [&](auto stagedTag) {
          constexpr bool STAGED = decltype(stagedTag)::value;
          for (int iu = iuLo; iu <= iuHi; iu++)
          {
            const int    lc = iu - cur.x;
            GatherColumn C;
            int          cax;
            if constexpr (STAGED)
            {
              C = cols[b][lc];
              cax = int(caxis[b][lc]);
            }
            else
            {
              C = gcols[size_t(p) * nu + iu];
              cax = int(colAxis[size_t(p) * nu + iu]);
            }
            if (cax != axis)
              continue;
            const double X = C.h0 + kd * C.h1;
            const double tt = C.t0 + kd * C.t1;
            const double a = V.y0 + tt * (V.sy - V.y0);
            const double bd = V.y1 * (1. - tt);
            // reference row near the block's centre height; e: y' of that row relative to j0
            const float bf = float(bd), rb = __builtin_amdgcn_rcpf(bf), sg = bf > 0.f ? 1.f : -1.f;
            const int   jref = int(rintf(float(double(j0c) + 0.5 * (kGNJ - 1) - a) * rb));
            const float e = float(a + bd * jref - double(j0c));
            const float wh = fmaxf(0.f, 1.f - fabsf(float(X - double(hc))));
            const int   rowBase = jref - cur.y; // LDS row of jref
            const int   kidx = axis == 0 ? i : k;   // this voxel column's slice index
            // R rows hold every ray within one voxel of a height: 4 where the row step is
            // at least half a voxel (on every lane of the wave), else 5.  The first G rows
            // lie inside the window by construction (|y' - j| < 1), so their weight needs
            // no clamp at zero.
            // first row with |y' - j| < 1 for voxel jj: ceil((-sg - (e - jj)) / b)
            const float q0 = (-sg - e) * rb;
            // the staged rows of the tile (a block's candidate rows lie in it: the tile
            // holds the widened block's footprint with kGRows + 1 rows of margin)
            const float * tcol = &tile[b][rowBase * kGStride + lc];
            // EDGE: the block holds voxels of the outer layers along y, which take their
            // rays below; elsewhere the loop has no branch, so the LDS reads of several
            // voxels are issued before their first use
            auto middle = [&](auto rowsTag, auto edgeTag) {
              constexpr int  R = decltype(rowsTag)::value;
              constexpr bool EDGE = decltype(edgeTag)::value;
#pragma unroll
              for (int jj = 0; jj < kGNJ; jj++)
              {
                if (EDGE && (j0 + jj < 2 || j0 + jj > ny - 3))
                  continue;
                const float d0f = ceilf(fmaf(float(jj), rb, q0));
                const int   d0 = int(d0f);
                const float y0 = fmaf(bf, d0f, e - float(jj)); // y' - j of that row
                float       s = 0.f;
                auto        weight = [&](int r) {
                  const float w = 1.f - fabsf(r == 0 ? y0 : fmaf(bf, float(r), y0));
                  return r < G ? w : fmaxf(0.f, w);
                };
                if constexpr (STAGED)
                {
                  const float * src = tcol + __mul24(d0, kGStride);
#pragma unroll
                  for (int r = 0; r < R; r++)
                    s = fmaf(weight(r), src[r * kGStride], s);
                }
                else
                {
#pragma unroll
                  for (int r = 0; r < R; r++)
                  {
                    const int jv = jref + d0 + r;
                    if (jv >= 0 && jv < nv)
                      s = fmaf(weight(r), gval[size_t(p) * viewRays + size_t(jv) * nu + iu], s);
                  }
                }
                acc[jj] = fmaf(wh, s, acc[jj]);
              }
            };
            const bool four = __all(fabsf(bf) >= 0.5f);
            if (STAGED && !yshell)
            {
              if (four)
                middle(std::integral_constant<int, 4>{}, std::false_type{});
              else
                middle(std::integral_constant<int, kGRows>{}, std::false_type{});
            }
            else if (four)
              middle(std::integral_constant<int, 4>{}, std::true_type{});
            else
              middle(std::integral_constant<int, kGRows>{}, std::true_type{});
            if (yshell)
            {
              // rows 0, 1, ny - 2, ny - 1 (those in this block, each once): the step weight
              // of this slice from the ray's first/last slices; on those, the height clamped
              // into the box (a candidate's horizontal position at this interior column's
              // slice lies inside the box, so it needs no clamp)
#pragma unroll
              for (int js = 0; js < 4; js++)
              {
                const int jy = js < 2 ? js : ny - 4 + js, jj = jy - j0;
                if (jj < 0 || jj >= kGNJ || jy >= ny || (js >= 2 && jy < 2))
                  continue;
                const float base = e - float(jj);
                const int   d0 = int(ceilf((-sg - base) * rb));
                const float y0 = fmaf(bf, float(d0), base);
                const float yv = j0c + float(jj);
                float       s = 0.f;
#pragma unroll 1
                for (int r = 0; r < kGRows; r++)
                {
                  const int jv = jref + d0 + r;
                  if (jv < 0 || jv >= nv)
                    continue;
                  const size_t pix = size_t(p) * viewRays + size_t(jv) * nu + iu;
                  const uint2  ve = gend[pix];
                  const float  val = __uint_as_float(ve.x);
                  const int    ns = int(ve.y & 0xffffu), fs = int(ve.y >> 16);
                  if (val == 0.f || kidx < min(ns, fs) || kidx > max(ns, fs))
                    continue;
                  const float yrel = fmaf(bf, float(r), y0);
                  if (kidx == ns || kidx == fs)
                  {
                    const float2 w2 = gw[pix];
                    const float  yc = fminf(fmaxf(yrel + yv, -cy), cy) - yv;
                    s = fmaf((kidx == ns ? w2.x : w2.y) * fmaxf(0.f, 1.f - fabsf(yc)), val, s);
                  }
                  else
                    s = fmaf(fmaxf(0.f, 1.f - fabsf(yrel)), val, s);
                }
                ash[js] = fmaf(wh, s, ash[js]);
              }
            }
          }
        };
        if (staged)
          columns(std::true_type{});
        else
          columns(std::false_type{});
      }
    }
    if (p + 1 < nviews)
      store(b ^ 1, p + 1, nxt);
    __syncthreads();
    cur = nxt;
  }
  if (live)
  {
#pragma unroll
    for (int jj = 0; jj < kGNJ; jj++)
      if (jj < ny - j0 && j0 + jj >= 2 && j0 + jj <= ny - 3)
        vol[(size_t(k) * ny + (j0 + jj)) * nx + i] += acc[jj];
    if (yshell)
    {
#pragma unroll
      for (int js = 0; js < 4; js++)
      {
        const int jy = js < 2 ? js : ny - 4 + js, jj = jy - j0;
        if (jj >= 0 && jj < kGNJ && jy < ny && !(js >= 2 && jy < 2))
          vol[(size_t(k) * ny + jy) * nx + i] += ash[js];
      }
    }
  }
}

// The voxels of the two outer layers along x and along z, one thread per voxel (voxel
// indices from `shell`), reading the rays from global memory: these voxels also
// receive first and last steps, so each candidate (ray, slice) takes its step weight
// from the ray's first/last slices and, on those, its position clamped into the box.
__global__ __launch_bounds__(256) void joseph_gather_shell_kernel(
  const GatherView * __restrict__ gviews, const GatherColumn * __restrict__ gcols,
  const signed char * __restrict__ colAxis, int nviews, int nu, int nv, const float * __restrict__ gval,
  const uint2 * __restrict__ gend, const float2 * __restrict__ gw, const int * __restrict__ shell, int nshell,
  float * vol, int nx, int ny, int nz)
{
  // Workgroups are dealt to the 8 XCDs in turn; XCD x takes a contiguous run of the
  // shell list instead of every 8th workgroup, so the voxels one XCD's L2 serves lie
  // together and read overlapping rays
  const unsigned nb = gridDim.x, full = nb / 8, rem = nb % 8, x = blockIdx.x % 8, q = blockIdx.x / 8;
  const unsigned lb = x * full + min(x, rem) + q;
  const int      n = int(lb * blockDim.x + threadIdx.x);
  if (n >= nshell)
    return;
  const int    idx = shell[n];
  const int    i = idx % nx, j = (idx / nx) % ny, k = idx / (nx * ny);
  const float  cx = 0.5f * float(nx - 1), cy = 0.5f * float(ny - 1), cz = 0.5f * float(nz - 1);
  const float  xc = float(i) - cx, zc = float(k) - cz, yv = float(j) - cy;
  const double xd = double(xc), zd = double(zc);
  const size_t viewRays = size_t(nu) * nv;
  float        acc = 0.f;
  for (int p = 0; p < nviews; p++)
  {
    const GatherView & V = gviews[p];
    for (int ax = 0; ax < 2; ax++)
    {
      const int axis = ax == 0 ? 0 : 2;
      if ((axis == 0 && !V.hasX) || (axis == 2 && !V.hasZ))
        continue;
      const float  kc = axis == 0 ? xc : zc, hc = axis == 0 ? zc : xc;
      const double kd = axis == 0 ? xd : zd;
      const int    kidx = axis == 0 ? i : k;
      const float  hmax = axis == 0 ? cz : cx; // in-slice horizontal box [-hmax, hmax]
      const float  xa = axis == 0 ? kc : hc - 1.f, za = axis == 0 ? hc - 1.f : kc;
      const float  xb = axis == 0 ? kc : hc + 1.f, zb = axis == 0 ? hc + 1.f : kc;
      const float  ua = (V.pu[3] + V.pu[0] * xa + V.pu[2] * za) * __builtin_amdgcn_rcpf(V.pw[3] + V.pw[0] * xa + V.pw[2] * za);
      const float  ub = (V.pu[3] + V.pu[0] * xb + V.pu[2] * zb) * __builtin_amdgcn_rcpf(V.pw[3] + V.pw[0] * xb + V.pw[2] * zb);
      const int    iuLo = max(int(ceilf(fminf(ua, ub) - 1e-3f)), 0);
      const int    iuHi = min(int(floorf(fmaxf(ua, ub) + 1e-3f)), nu - 1);
      for (int iu = iuLo; iu <= iuHi; iu++)
      {
        // both loads issued before the test, so they overlap
        const int          cax = int(colAxis[size_t(p) * nu + iu]);
        const GatherColumn C = gcols[size_t(p) * nu + iu];
        if (cax != axis)
          continue;
        const double       X = C.h0 + kd * C.h1;
        const double       tt = C.t0 + kd * C.t1;
        const double       a = V.y0 + tt * (V.sy - V.y0);
        const double       bd = V.y1 * (1. - tt);
        const float        bf = float(bd), rb = __builtin_amdgcn_rcpf(bf), sg = bf > 0.f ? 1.f : -1.f;
        // rows relative to jref = the row nearest this voxel's height
        const int          jref = int(rintf(float(double(yv) - a) * rb));
        const float        base = float(a + bd * jref - double(yv)); // y' - j of row jref
        const int          d0 = int(ceilf((-sg - base) * rb));
        const float        wh = fmaxf(0.f, 1.f - fabsf(float(X) - hc));
        const float        whc = fmaxf(0.f, 1.f - fabsf(fminf(fmaxf(float(X), -hmax), hmax) - hc));
        // the rows' slices and values, all loads issued together (rows outside the
        // detector read a clamped row and take weight zero; a ray the gather does not
        // take has value zero and weights zero)
        const size_t colBase = size_t(p) * viewRays + iu;
        uint2        ve[kGRows];
#pragma unroll
        for (int r = 0; r < kGRows; r++)
          ve[r] = gend[colBase + size_t(min(max(jref + d0 + r, 0), nv - 1)) * nu];
#pragma unroll
        for (int r = 0; r < kGRows; r++)
        {
          const int jv = jref + d0 + r;
          // the middle-step and end-step weights of this row
          const float yabs = fmaf(bf, float(d0 + r), base) + yv;
          const float wm = wh * fmaxf(0.f, 1.f - fabsf(yabs - yv));
          const float we = whc * fmaxf(0.f, 1.f - fabsf(fminf(fmaxf(yabs, -cy), cy) - yv));
          const float val = __uint_as_float(ve[r].x);
          const int   ns = int(ve[r].y & 0xffffu), fs = int(ve[r].y >> 16);
          if (jv < 0 || jv >= nv || kidx < min(ns, fs) || kidx > max(ns, fs))
            continue;
          float w = wm;
          if (kidx == ns || kidx == fs)
          {
            const float2 w2 = gw[colBase + size_t(jv) * nu];
            w = (kidx == ns ? w2.x : w2.y) * we;
          }
          acc = fmaf(w, val, acc);
        }
      }
    }
  }
  vol[idx] += acc;
}

// Forward projection as a gather, for the geometries the gather backprojection takes
// (detector columns vertical in the volume).  A column's rays lie in one vertical
// plane: at slice k of their dominant axis they share the horizontal position X(k),
// and a middle sample's bilinear interpolation (.hxx:182-543) is the linear
// interpolation, at the ray's height y'(k), of the profile g(y) = (1 - fx) v(hx, y) +
// fx v(hx + 1, y) with hx = floor(X), fx = X - hx, the same for every ray of the
// column.  A workgroup takes 16 columns x kFRows rows of one view; per slice its
// threads build the 16 profiles over the rows' height range in LDS (double-buffered,
// one barrier per slice), then each thread adds the middle samples of its kFRays rays
// (rows 16 apart) from two LDS reads each.  First and last samples, at the ray's own (clamped)
// positions, read their four taps from global memory as the per-ray projector does.
// Rays of columns that mix dominant axes are walked per ray.
constexpr int kFRows = 64; // detector rows per workgroup
constexpr int kFRays = kFRows / 16; // rays per thread (rows 16 apart)
constexpr int kFProf = 48; // profile entries per column: kFRows |b| + 3 with |b| < 2/3
constexpr int kFPad = 49;  // LDS stride of a column's profile (its columns on distinct banks)

// The first and last samples, and every sample of a ray walked alone (a column that
// mixes dominant axes), as the per-ray forward projector computes them.
__device__ __forceinline__ float
RayBorder(const Ray & R, double step, const float * bb, double x, double y)
{
  int   tt[4];
  float w[4];
  BorderTaps(R, x, y, tt, w);
  float r = 0.f;
  r += w[0] * bb[tt[0]];
  r += w[1] * bb[tt[1]];
  r += w[2] * bb[tt[2]];
  r += w[3] * bb[tt[3]];
  return float(double(r) * step);
}

__device__ __forceinline__ float
RayEnds(const Ray & R, const float * vol, const float * volT)
{
  const float * b = (R.transposed ? volT : vol) + R.start;
  if (R.fs == R.ns)
    return RayBorder(R, R.wOne, b, R.cx, R.cy);
  const float first = RayBorder(R, R.wB, b, R.cx, R.cy);
  const double n = double(R.nmid + 1);
  return first + RayBorder(R, R.wE, b + (R.nmid + 1) * R.oz, R.cx + n * R.stepx, R.cy + n * R.stepy);
}

__device__ __forceinline__ float
RayWalk(const Ray & R, const float * vol, const float * volT)
{
  const float * b = (R.transposed ? volT : vol) + R.start;
  if (R.fs == R.ns)
    return RayBorder(R, R.wOne, b, R.cx, R.cy);
  float sum = RayBorder(R, R.wB, b, R.cx, R.cy);
  b += R.oz;
  Walker wk;
  wk.Init(R.cx + R.stepx, R.cy + R.stepy);
  for (int s2 = 0; s2 < R.nmid; ++s2)
  {
    float       lx, ly;
    const int   idx = wk.Stage(R, lx, ly);
    const float lxc = 1.f - lx, lyc = 1.f - ly;
    sum += (lxc * lyc) * b[idx] + (lx * lyc) * b[idx + 1] + (lxc * ly) * b[idx + R.oy] + (lx * ly) * b[idx + 1 + R.oy];
    b += R.oz;
  }
  return sum + RayBorder(R, R.wE, b, wk.PosX(), wk.PosY());
}

// Middle samples of the columns of dominant axis A (0: slices along x in the z-fastest
// copy, horizontal z; 2: slices along z, horizontal x), slices klo..khi (uniform over
// the workgroup).  Ray q of this thread takes the slices lo[q]..hi[q].  The profile
// values of slice k + 1 are loaded while the rays of slice k are summed.
template <int A>
__device__ __forceinline__ void
ForwardGatherSweep(const GatherView & V, const GatherColumn & C, bool colOn, const int (&lo)[kFRays], const int (&hi)[kFRays],
                   int jv0, int rb, int rlast, const float * __restrict__ buf, int nsl, int nh, int ny, int klo,
                   int khi, float (*prof)[16][kFPad], float (&acc)[kFRays])
{
  const int    t = threadIdx.x, c = t & 15, rq = t >> 4;
  const size_t sk = size_t(ny) * nh; // slice stride; rows nh apart, the horizontal axis unit-stride
  const double ck = 0.5 * (nsl - 1), ch = 0.5 * (nh - 1), cy = 0.5 * (ny - 1);
  constexpr int NM = kFProf / 16; // profile entries per thread
  // the column at slice k: the rows' heights a + b jv relative to the profile's first
  // height ylo (Af + Bf jv), and this thread's profile entries as tap pairs
  struct Slice
  {
    float  Af, Bf, fx;
    int    ylo;
    float2 v[NM];
  };
  // X and t advance by h1 and t1 per slice; a and b are affine in t
  const double kd0 = double(klo) - ck;
  double       X = C.h0 + kd0 * C.h1 + ch, tt = C.t0 + kd0 * C.t1;
  const double dsy = V.sy - V.y0, a0 = V.y0 + cy;
  auto         fetch = [&](int k, Slice & S) {
    if (!colOn || k > khi)
      return;
    const double a = fma(tt, dsy, a0), b = fma(-V.y1, tt, V.y1);
    const double fl = floor(X);
    const int    hx = min(max(int(fl), 0), nh - 2);
    S.fx = float(X - fl);
    S.ylo = int(floor(fmin(a + b * rb, a + b * rlast)));
    S.Af = float(a - double(S.ylo));
    S.Bf = float(b);
    // entries rq + 16 m: heights ylo + rq + 16 m (zero outside the volume)
    const int     y0 = S.ylo + rq;
    const float * q0 = buf + (size_t(k) * sk + hx) + ptrdiff_t(y0) * nh;
#pragma unroll
    for (int m = 0; m < NM; m++)
    {
      const float * q = q0 + 16 * m * nh;
      S.v[m] = unsigned(y0 + 16 * m) < unsigned(ny) ? make_float2(q[0], q[1]) : make_float2(0.f, 0.f);
    }
    X += C.h1;
    tt += C.t1;
  };
  Slice cur{}, nxt{};
  fetch(klo, cur);
  for (int k = klo; k <= khi; k++)
  {
    float(*g)[kFPad] = prof[k & 1];
    if (colOn)
    {
#pragma unroll
      for (int m = 0; m < NM; m++)
        g[c][rq + 16 * m] = (1.f - cur.fx) * cur.v[m].x + cur.fx * cur.v[m].y;
    }
    __syncthreads();
    fetch(k + 1, nxt);
    if (colOn)
    {
#pragma unroll
      for (int q = 0; q < kFRays; q++)
      {
        if (k < lo[q] || k > hi[q])
          continue;
        // the row pair of the ray's height, its index clamped as the per-ray walk
        // clamps it, the fraction unclamped
        // (entries outside the volume are zero; the index bound only guards rounding)
        const float u = fmaxf(fmaf(cur.Bf, float(jv0 + 16 * q), cur.Af), 0.f); // y' - ylo (>= 0 but for rounding)
        const float fu = floorf(u);
        const float fy = u - fu;
        const int   iy = min(int(fu), kFProf - 2);
        const float g0 = g[c][iy];
        acc[q] += fmaf(fy, g[c][iy + 1] - g0, g0);
      }
    }
    cur = nxt;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void joseph_fp_gather_kernel(
  const JosephView * __restrict__ views, const GatherView * __restrict__ gviews,
  const GatherColumn * __restrict__ gcols, const signed char * __restrict__ colAxis, int nu, int nv,
  const float * pin, float * pout, const float * __restrict__ vol, const float * __restrict__ volT, int nx, int ny,
  int nz, double spx, double spy, double spz, double infClip, double supClip, int xcdOrder)
{
  __shared__ float prof[2][16][kFPad];
  __shared__ int   krange[2][2]; // [x / z columns][first, last middle slice]
  const int        t = threadIdx.x, c = t & 15, rq = t >> 4;
  // grid (column blocks, views, row bands), as the per-ray projector: workgroups in
  // flight share a band of detector rows, whose rays stay in one slab of the volume,
  // so the slab stays in L2 while every view reads it
  // Workgroups are dealt to the 8 XCDs in turn in linear order.  xcdOrder 1 renumbers
  // them so that XCD x takes a contiguous run of (column block, view, row band)
  // triples: its L2 then serves one band's rays, and the 8 XCDs read 8 bands' slabs of
  // the volume at once (for volumes whose two copies fit the last-level cache).
  // xcdOrder 2 keeps the bands in order and, within a band, gives XCD x a contiguous
  // run of its (column block, view) pairs: all XCDs read one slab, each for a range of
  // neighbouring views.
  const unsigned   nblk = gridDim.x * gridDim.y * gridDim.z, full = nblk / 8, rem = nblk % 8;
  const unsigned   lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  unsigned         blk = lin;
  if (xcdOrder == 1)
    blk = (lin % 8) * full + min(lin % 8, rem) + lin / 8;
  else if (xcdOrder == 2)
  {
    // within each row band: XCD x takes a contiguous run of its (column block, view) pairs
    const unsigned P = gridDim.x * gridDim.y, w = lin % P, f = P / 8, r = P % 8, x = w % 8;
    blk = (lin / P) * P + x * f + min(x, r) + w / 8;
  }
  const int        p = int((blk / gridDim.x) % gridDim.y), iu = int(blk % gridDim.x) * 16 + c,
                   rb = int(blk / (gridDim.x * gridDim.y)) * kFRows;
  const int        rlast = min(rb + kFRows, nv) - 1;
  const int        jv0 = rb + rq; // this thread's rays: rows jv0 + 16 q
  if (t < 2)
  {
    krange[t][0] = INT_MAX;
    krange[t][1] = INT_MIN;
  }
  __syncthreads();
  const int cax = iu < nu ? int(colAxis[size_t(p) * nu + iu]) : -1;
  float     acc[kFRays];
  double    norm[kFRays];
  int       lo[kFRays], hi[kFRays];
  bool      hit[kFRays];
  int       klo = INT_MAX, khi = INT_MIN;
#pragma unroll
  for (int q = 0; q < kFRays; q++)
  {
    const int jv = jv0 + 16 * q;
    Ray       R;
    acc[q] = 0.f;
    norm[q] = 0.;
    hit[q] = iu < nu && jv < nv && SetupRay<false>(views[p], iu, jv, p, nx, ny, nz, spx, spy, spz, infClip, supClip, R);
    lo[q] = 1;
    hi[q] = 0;
    if (!hit[q])
      continue;
    norm[q] = R.stepNorm;
    if (cax < 0 || R.axis != cax)
      acc[q] = RayWalk(R, vol, volT); // a column that mixes dominant axes
    else
    {
      acc[q] = RayEnds(R, vol, volT);
      lo[q] = min(R.ns, R.fs) + 1; // middle slices
      hi[q] = max(R.ns, R.fs) - 1;
      klo = min(klo, lo[q]);
      khi = max(khi, hi[q]);
    }
  }
  if (cax >= 0 && klo <= khi)
  {
    atomicMin(&krange[cax >> 1][0], klo);
    atomicMax(&krange[cax >> 1][1], khi);
  }
  __syncthreads();
  const GatherView V = gviews[p];
  GatherColumn     C{};
  if (cax >= 0)
    C = gcols[size_t(p) * nu + iu];
  if (krange[0][0] <= krange[0][1])
    ForwardGatherSweep<0>(V, C, cax == 0, lo, hi, jv0, rb, rlast, volT, nx, nz, ny, krange[0][0], krange[0][1], prof,
                          acc);
  if (krange[1][0] <= krange[1][1])
    ForwardGatherSweep<2>(V, C, cax == 2, lo, hi, jv0, rb, rlast, vol, nz, nx, ny, krange[1][0], krange[1][1], prof,
                          acc);
#pragma unroll
  for (int q = 0; q < kFRays; q++)
  {
    const int jv = jv0 + 16 * q;
    if (iu >= nu || jv >= nv)
      continue;
    const long pix = (long(p) * nv + jv) * nu + iu;
    if (hit[q])
      pout[pix] = float(double(pin ? pin[pix] : 0.f) + double(acc[q]) * norm[q]);
    else
      pout[pix] = pin ? pin[pix] : 0.f;
  }
}

struct JosephCache
{
  std::map<hipStream_t, DeviceBuffer<JosephView>>          views;
  std::map<hipStream_t, DeviceBuffer<float>>               volT;
  std::map<hipStream_t, DeviceBuffer<unsigned long long>> count;
  // gather backprojection: per-view and per-column tables, block lists, per-ray data
  std::map<hipStream_t, DeviceBuffer<GatherView>>   gviews;
  std::map<hipStream_t, DeviceBuffer<GatherColumn>> gcols;
  std::map<hipStream_t, DeviceBuffer<signed char>>  gaxis;
  std::map<hipStream_t, DeviceBuffer<int>>          gblocks;
  std::map<hipStream_t, DeviceBuffer<float>>        gval;
  std::map<hipStream_t, DeviceBuffer<uint2>>        gend;
  std::map<hipStream_t, DeviceBuffer<float2>>       gw;
  std::map<hipStream_t, DeviceBuffer<int>>          gshell; // voxels of the outer layers along x and z
  std::map<hipStream_t, std::array<int, 3>>         gshellDims;
  std::map<hipStream_t, int>                        gshellCount;
  // the arguments the uploaded tables were made from: a call with the same ones (every
  // iteration of an iterative reconstruction) reuses them without host work or copies
  std::map<hipStream_t, std::vector<double>> viewsKey;
  std::map<hipStream_t, std::vector<double>> gplanKey;
  std::map<hipStream_t, long long>           gplanBlocks; // -1: the gather does not apply
  std::map<hipStream_t, int>                 gplanFree;
  std::map<hipStream_t, bool>                gplanForward;
};
thread_local JosephCache g_joseph;

// Host side of the gather backprojection: eligibility, per-view and per-column tables.
struct GatherPlan
{
  std::vector<GatherView>   views;
  std::vector<GatherColumn> cols;
  std::vector<signed char>  axis;
  std::vector<int>          blocks; // packed block coordinates (bx | bz << 10 | by << 20)
  int                       freeRows = 1; // leading candidate rows whose weight needs no clamp
  bool                      forward = false; // the forward gather's profiles hold kFRows rows
};

// True when the geometry lets the gather take the backprojection: default clips,
// detector columns vertical in the volume (no in-plane or out-of-plane tilt), every
// voxel between the detector and the source, and rows dense enough that kGRows rows
// hold every ray within one voxel of a voxel's height.  Columns whose rays are not all
// x- or all z-dominant are marked -1 (their rays take the per-ray path).
bool
PlanGather(const int proj_dim[3], const int vol_dim[3], const double * proj_to_vol, const double * source,
           double inferior_clip, double superior_clip, GatherPlan & G)
{
  if (inferior_clip != 0. || superior_clip != 1.)
    return false;
  const int nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
  const int nx = vol_dim[0], ny = vol_dim[1], nz = vol_dim[2];
  if (nx < 5 || ny < 5 || nz < 5 || nx > 16 * 1023 || nz > 16 * 1023 || ny > kGNJ * 2047 || nu > 32767)
    return false;
  const double cx = 0.5 * (nx - 1), cy = 0.5 * (ny - 1), cz = 0.5 * (nz - 1);
  const double ctr[3] = { cx, cy, cz };
  G.views.assign(np, GatherView{});
  G.cols.assign(size_t(np) * nu, GatherColumn{});
  G.axis.assign(size_t(np) * nu, -1);
  double minRowStep = DBL_MAX, maxRowStep = 0.;
  for (int p = 0; p < np; p++)
  {
    const double * m = proj_to_vol + 12 * p;
    const double * S = source + 3 * p;
    const double   eu[3] = { m[0], m[4], m[8] }, ev[3] = { m[1], m[5], m[9] };
    const double   o[3] = { m[3] + m[2] * p, m[7] + m[6] * p, m[11] + m[10] * p };
    const double   scale = std::max(std::abs(eu[0]) + std::abs(eu[2]), std::abs(ev[1]));
    if (!(scale > 0.) || std::abs(eu[1]) > 1e-9 * scale || std::abs(ev[0]) > 1e-9 * scale ||
        std::abs(ev[2]) > 1e-9 * scale)
      return false;
    // volume index -> detector (u w, v w, w): iu = (c x ev).d / n.d, jv = (eu x c).d / n.d
    // with c = S - o, n = eu x ev, d = P - S (the ray from the source through P meets the
    // detector plane o + iu eu + jv ev)
    auto cross = [](const double * a, const double * b, double * r) {
      r[0] = a[1] * b[2] - a[2] * b[1];
      r[1] = a[2] * b[0] - a[0] * b[2];
      r[2] = a[0] * b[1] - a[1] * b[0];
    };
    const double c[3] = { S[0] - o[0], S[1] - o[1], S[2] - o[2] };
    double       n[3], au[3], av[3];
    cross(eu, ev, n);
    cross(c, ev, au);
    cross(eu, c, av);
    GatherView & V = G.views[p];
    auto         row = [&](const double * a, float * out) {
      double off = 0.;
      for (int r = 0; r < 3; r++)
      {
        out[r] = float(a[r]);
        off += a[r] * (ctr[r] - S[r]);
      }
      out[3] = float(off);
    };
    row(au, V.pu);
    row(av, V.pv);
    row(n, V.pw);
    V.y0 = o[1] - cy;
    V.y1 = ev[1];
    V.sy = S[1] - cy;
    const double dyTop = std::abs(S[1] - o[1]), dyBot = std::abs(S[1] - (o[1] + (nv - 1) * ev[1]));
    const double mh = std::max(dyTop, dyBot);
    for (int iu = 0; iu < nu; iu++)
    {
      const double px = o[0] + iu * eu[0], pz = o[2] + iu * eu[2];
      const double dx = S[0] - px, dz = S[2] - pz;
      int          ax = -1;
      if (std::abs(dx) >= std::abs(dz) && std::abs(dx) >= mh)
        ax = 0;
      else if (std::abs(dz) > std::abs(dx) && std::abs(dz) > mh)
        ax = 2;
      if (ax < 0)
        continue;
      GatherColumn & C = G.cols[size_t(p) * nu + iu];
      const double   dk = ax == 0 ? dx : dz, dh = ax == 0 ? dz : dx;
      const double   pk = ax == 0 ? px : pz, ph = ax == 0 ? pz : px;
      const double   ck = ax == 0 ? cx : cz, ch = ax == 0 ? cz : cx;
      C.t1 = 1. / dk;
      C.t0 = (ck - pk) * C.t1;
      C.h0 = ph - ch + C.t0 * dh;
      C.h1 = C.t1 * dh;
      // the column's rays cross the volume's horizontal square, widened by the one voxel
      // by which a candidate's position may lie outside it, where t is in [tIn, tOut];
      // there every voxel must lie between the pixel (t = 0) and the source (t = 1), and
      // the rows' vertical spacing, |ev_y| (1 - t), is smallest at tOut, largest at tIn
      double tIn = -DBL_MAX, tOut = DBL_MAX;
      const double hp[2] = { px - cx, pz - cz }, hd[2] = { dx, dz }, hc[2] = { cx + 1., cz + 1. };
      for (int r = 0; r < 2; r++)
      {
        if (hd[r] == 0.)
        {
          if (std::abs(hp[r]) > hc[r])
            tIn = DBL_MAX;
          continue;
        }
        double ta = (-hc[r] - hp[r]) / hd[r], tb = (hc[r] - hp[r]) / hd[r];
        if (ta > tb)
          std::swap(ta, tb);
        tIn = std::max(tIn, ta);
        tOut = std::min(tOut, tb);
      }
      if (tIn <= tOut)
      {
        if (!(tIn > 0. && tOut < 1.))
          return false;
        minRowStep = std::min(minRowStep, std::abs(ev[1] * (1. - tOut)));
        maxRowStep = std::max(maxRowStep, std::abs(ev[1] * (1. - tIn)));
      }
      G.axis[size_t(p) * nu + iu] = static_cast<signed char>(ax);
      (ax == 0 ? V.hasX : V.hasZ) = 1;
    }
  }
  // the rows within one voxel of a height form an open interval of 2 / step rows
  if (!(minRowStep > 0.) || 2. / minRowStep > double(kGRows))
    return false;
  // the first candidate row has |y' - j| < 1 and rows advance by the step b, so rows
  // 0..2 stay inside the window when 3 |b| < 2 (a margin covers the FP32 arithmetic)
  G.freeRows = maxRowStep < (2. / 3.) * (1. - 1e-4) ? 3 : 1;
  // kFRows rows span (kFRows - 1) |b| + 1 heights above the profile's first one, and
  // a middle sample reads two entries
  G.forward = maxRowStep * (kFRows - 1) * (1. + 1e-4) + 3. <= double(kFProf);
  const int nbx = (nx + 15) / 16, nbz = (nz + 15) / 16, nby = (ny + kGNJ - 1) / kGNJ;
  G.blocks.clear();
  for (int by = 0; by < nby; by++)
    for (int bz = 0; bz < nbz; bz++)
      for (int bx = 0; bx < nbx; bx++)
        G.blocks.push_back(bx | (bz << 10) | (by << 20));
  return true;
}

// Runs the gather backprojection of every view (in chunks that keep the per-ray data,
// 20 bytes a ray, under 1 GiB): per chunk, the per-ray pass (which also splats the rays
// the gather does not take into acc / accT), then the gather over every block (columns
// off the two outer layers along x and z) and the shell kernel (those layers).
// Plans the gather for these arguments and uploads its tables, or reuses the tables of
// the previous call on stream s when its arguments were the same.  Returns the number
// of gather blocks, or -1 when the gather does not apply.
long long
PrepareGather(const int proj_dim[3], const int vol_dim[3], const double * proj_to_vol, const double * source,
              double inferior_clip, double superior_clip, hipStream_t s, int & freeRows, bool * forward = nullptr)
{
  const int           np = proj_dim[2], nu = proj_dim[0];
  std::vector<double> key = { double(proj_dim[0]), double(proj_dim[1]), double(np),           double(vol_dim[0]),
                              double(vol_dim[1]),  double(vol_dim[2]),  inferior_clip,        superior_clip };
  key.insert(key.end(), proj_to_vol, proj_to_vol + 12 * size_t(np));
  key.insert(key.end(), source, source + 3 * size_t(np));
  int dev = 0;
  RTKHIP_CHECK(hipGetDevice(&dev));
  key.push_back(double(dev));
  auto & cached = g_joseph.gplanKey[s];
  if (cached == key)
  {
    freeRows = g_joseph.gplanFree[s];
    if (forward)
      *forward = g_joseph.gplanForward[s];
    return g_joseph.gplanBlocks[s];
  }
  cached.clear(); // a failure below leaves no stale entry
  GatherPlan G;
  long long  nblocks = -1;
  if (PlanGather(proj_dim, vol_dim, proj_to_vol, source, inferior_clip, superior_clip, G))
  {
    auto & gv = g_joseph.gviews[s];
    auto & gc = g_joseph.gcols[s];
    auto & ga = g_joseph.gaxis[s];
    auto & gb = g_joseph.gblocks[s];
    gv.Resize(np);
    gc.Resize(size_t(np) * nu);
    ga.Resize(size_t(np) * nu);
    gb.Resize(G.blocks.size());
    RTKHIP_CHECK(hipMemcpyAsync(gv.get(), G.views.data(), sizeof(GatherView) * np, hipMemcpyHostToDevice, s));
    RTKHIP_CHECK(hipMemcpyAsync(gc.get(), G.cols.data(), sizeof(GatherColumn) * G.cols.size(),
                                hipMemcpyHostToDevice, s));
    RTKHIP_CHECK(hipMemcpyAsync(ga.get(), G.axis.data(), G.axis.size(), hipMemcpyHostToDevice, s));
    RTKHIP_CHECK(
      hipMemcpyAsync(gb.get(), G.blocks.data(), sizeof(int) * G.blocks.size(), hipMemcpyHostToDevice, s));
    nblocks = (long long)G.blocks.size();
  }
  g_joseph.gplanBlocks[s] = nblocks;
  g_joseph.gplanFree[s] = freeRows = G.freeRows;
  g_joseph.gplanForward[s] = nblocks >= 0 && G.forward;
  if (forward)
    *forward = g_joseph.gplanForward[s];
  cached = std::move(key);
  return nblocks;
}

void
GatherBackproject(long long nblocks, int freeRows, const JosephView * dv, const int proj_dim[3], const int vol_dim[3],
                  const double vol_spacing[3], double infClip, double supClip, const void * dev_proj, int proj_dtype,
                  float * vol, float * volT, hipStream_t s)
{
  const int nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
  const int nx = vol_dim[0], ny = vol_dim[1], nz = vol_dim[2];
  auto &    gv = g_joseph.gviews[s];
  auto &    gc = g_joseph.gcols[s];
  auto &    ga = g_joseph.gaxis[s];
  auto &    gb = g_joseph.gblocks[s];
  // the voxels of the two outer layers along x and z, listed once per volume size
  auto &                   gs = g_joseph.gshell[s];
  const std::array<int, 3> dims = { nx, ny, nz };
  if (!gs.get() || g_joseph.gshellDims[s] != dims)
  {
    // consecutive threads take neighbouring voxels of one row across the face (x on the
    // faces z = 0, 1, nz - 2, nz - 1; z on the faces x = 0, 1, nx - 2, nx - 1), whose
    // candidate rays lie in the same detector rows
    std::vector<int> shell;
    for (int z : { 0, 1, nz - 2, nz - 1 })
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++)
          shell.push_back((z * ny + y) * nx + x);
    for (int x : { 0, 1, nx - 2, nx - 1 })
      for (int y = 0; y < ny; y++)
        for (int z = 2; z < nz - 2; z++)
          shell.push_back((z * ny + y) * nx + x);
    gs.Resize(shell.size());
    RTKHIP_CHECK(hipMemcpy(gs.get(), shell.data(), sizeof(int) * shell.size(), hipMemcpyHostToDevice));
    g_joseph.gshellDims[s] = dims;
    g_joseph.gshellCount[s] = int(shell.size());
  }
  const int nshell = g_joseph.gshellCount[s];
  const size_t rays = size_t(nu) * nv;
  const int    chunk = int(std::max<size_t>(1, std::min<size_t>(np, (size_t(1) << 30) / (20 * rays))));
  auto &       dval = g_joseph.gval[s];
  auto &       dend = g_joseph.gend[s];
  auto &       dw = g_joseph.gw[s];
  dval.Resize(rays * chunk);
  dend.Resize(rays * chunk);
  dw.Resize(rays * chunk);
  for (int p0 = 0; p0 < np; p0 += chunk)
  {
    const int c = std::min(chunk, np - p0);
    dim3      pg((nu + 15) / 16, c, (nv + 15) / 16);
#define RTKHIP_GPREP(TP)                                                                                             \
  hipLaunchKernelGGL(joseph_gather_prep_kernel<TP>, pg, dim3(256), 0, s, dv, p0, ga.get() + size_t(p0) * nu, nu, nv, \
                     static_cast<const TP *>(dev_proj), vol, volT, nx, ny, nz, vol_spacing[0], vol_spacing[1],     \
                     vol_spacing[2], infClip, supClip, dval.get(), dend.get(), dw.get())
    if (proj_dtype == RTKHIP_F16)
      RTKHIP_GPREP(__half);
    else if (proj_dtype == RTKHIP_BF16)
      RTKHIP_GPREP(__hip_bfloat16);
    else
      RTKHIP_GPREP(float);
#undef RTKHIP_GPREP
    RTKHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(freeRows == 3 ? joseph_gather_kernel<3> : joseph_gather_kernel<1>, dim3(unsigned(nblocks)),
                       dim3(256), 0, s, gv.get() + p0, gc.get() + size_t(p0) * nu, ga.get() + size_t(p0) * nu, c, nu,
                       nv, dval.get(), dend.get(), dw.get(), gb.get(), vol, nx, ny, nz);
    RTKHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(joseph_gather_shell_kernel, dim3(unsigned((nshell + 255) / 256)), dim3(256), 0, s,
                       gv.get() + p0, gc.get() + size_t(p0) * nu, ga.get() + size_t(p0) * nu, c, nu, nv, dval.get(),
                       dend.get(), dw.get(), gs.get(), nshell, vol, nx, ny, nz);
    RTKHIP_CHECK_LAUNCH();
  }
}

void
CheckArgs(const int proj_dim[3], const int vol_dim[3], const double * proj_to_vol, const double * source,
          const double vol_spacing[3])
{
  RTKHIP_REQUIRE(proj_dim && vol_dim && proj_to_vol && source && vol_spacing, RTKHIP_ERR_INVALID_ARGUMENT,
                 "joseph: null argument");
  RTKHIP_REQUIRE(proj_dim[0] > 0 && proj_dim[1] > 0 && proj_dim[2] >= 0 && vol_dim[0] > 0 && vol_dim[1] > 0 &&
                   vol_dim[2] > 0,
                 RTKHIP_ERR_INVALID_ARGUMENT, "joseph: sizes must be positive");
  RTKHIP_REQUIRE(vol_dim[0] >= 2 && vol_dim[1] >= 2 && vol_dim[2] >= 2, RTKHIP_ERR_UNSUPPORTED,
                 "joseph: every volume dimension must be at least 2 (bilinear taps)");
  RTKHIP_REQUIRE(proj_dim[2] <= 65535, RTKHIP_ERR_UNSUPPORTED, "joseph: at most 65535 views per call");
  RTKHIP_REQUIRE(vol_dim[1] <= 65535, RTKHIP_ERR_UNSUPPORTED, "joseph: at most 65535 volume rows");
  const long long nvox = (long long)vol_dim[0] * vol_dim[1] * vol_dim[2];
  RTKHIP_REQUIRE(nvox < (1LL << 31), RTKHIP_ERR_UNSUPPORTED, "joseph: volumes are limited to 2^31 voxels");
}

JosephView *
UploadViews(int np, const double * proj_to_vol, const double * source, hipStream_t s)
{
  auto &              dv = g_joseph.views[s];
  auto &              cached = g_joseph.viewsKey[s];
  std::vector<double> key(proj_to_vol, proj_to_vol + 12 * size_t(np));
  key.insert(key.end(), source, source + 3 * size_t(np));
  int dev = 0;
  RTKHIP_CHECK(hipGetDevice(&dev));
  key.push_back(double(dev));
  if (dv.get() && cached == key)
    return dv.get();
  cached.clear();
  std::vector<JosephView> hv(np);
  for (int p = 0; p < np; p++)
  {
    for (int j = 0; j < 12; j++)
      hv[p].m[j] = proj_to_vol[12 * p + j];
    for (int j = 0; j < 3; j++)
      hv[p].src[j] = source[3 * p + j];
    hv[p].src[3] = 1.;
  }
  dv.Resize(np);
  RTKHIP_CHECK(hipMemcpyAsync(dv.get(), hv.data(), sizeof(JosephView) * np, hipMemcpyHostToDevice, s));
  cached = std::move(key);
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
    dim3         grid((nu + 15) / 16, np, (nv + 15) / 16);
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
      // detector columns vertical in the volume: the gather (same tables as the gather
      // backprojection, kept across calls with the same geometry)
      int       freeRows = 1;
      bool      forward = false;
      const int nblocks =
        int(PrepareGather(proj_dim, vol_dim, proj_to_vol, source, inferior_clip, superior_clip, s, freeRows, &forward));
      if (nblocks >= 0 && forward)
      {
        dim3 fg((nu + 15) / 16, np, (nv + kFRows - 1) / kFRows);
        // XCD-contiguous workgroups over the whole grid when the volume and its
        // transposed copy fit well in the 256 MB last-level cache (256^3 <- 90 x 512^2:
        // 5.50 -> 4.73 ms; 4.82 ms within bands), within each row band otherwise
        // (512^3 <- 720 x 1024^2: 318 -> 283 ms; 340 ms over the whole grid)
        const int xcdOrder = 2 * size_t(vol_dim[0]) * vol_dim[1] * vol_dim[2] * sizeof(float) <= (size_t(192) << 20) ? 1 : 2;
        hipLaunchKernelGGL(joseph_fp_gather_kernel, fg, dim3(256), 0, s, dv, g_joseph.gviews[s].get(),
                           g_joseph.gcols[s].get(), g_joseph.gaxis[s].get(), nu, nv, dev_proj_in, dev_proj_out,
                           dev_vol, dT.get(), vol_dim[0], vol_dim[1], vol_dim[2], vol_spacing[0], vol_spacing[1],
                           vol_spacing[2], infClip, supClip, xcdOrder);
      }
      else
        hipLaunchKernelGGL(joseph_kernel<false>, grid, dim3(256), 0, s, dv, nu, nv, dev_proj_in, dev_proj_out,
                           dev_vol, dT.get(), vol_dim[0], vol_dim[1], vol_dim[2], vol_spacing[0], vol_spacing[1],
                           vol_spacing[2], infClip, supClip, nullptr);
      RTKHIP_CHECK_LAUNCH();
    }
  });
}

extern "C" int
rtkhip_joseph_backproject(const int proj_dim[3], const int vol_dim[3], const double * proj_to_vol,
                          const double * source, const double vol_spacing[3], double inferior_clip,
                          double superior_clip, const float * dev_vol_in, float * dev_vol_out, const void * dev_proj,
                          int proj_dtype, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    CheckArgs(proj_dim, vol_dim, proj_to_vol, source, vol_spacing);
    RTKHIP_REQUIRE(dev_vol_out && dev_proj, RTKHIP_ERR_INVALID_ARGUMENT, "joseph_backproject: null buffer");
    RTKHIP_REQUIRE(proj_dtype == RTKHIP_F32 || proj_dtype == RTKHIP_F16 || proj_dtype == RTKHIP_BF16,
                   RTKHIP_ERR_INVALID_ARGUMENT, "joseph_backproject: unknown projection dtype");
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
    int             freeRows = 1;
    const long long nblocks =
      PrepareGather(proj_dim, vol_dim, proj_to_vol, source, inferior_clip, superior_clip, s, freeRows);
    if (nblocks >= 0)
    {
      GatherBackproject(nblocks, freeRows, dv, proj_dim, vol_dim, vol_spacing, infClip, supClip, dev_proj, proj_dtype,
                        dev_vol_out, dT.get(), s);
      dim3 tg((vol_dim[0] + 31) / 32, (vol_dim[2] + 31) / 32, vol_dim[1]);
      hipLaunchKernelGGL(add_transposed_kernel, tg, dim3(256), 0, s, dT.get(), dev_vol_out, vol_dim[0], vol_dim[1],
                         vol_dim[2]);
      RTKHIP_CHECK_LAUNCH();
      return;
    }
    dim3 grid((nu + 15) / 16, np, (nv + 15) / 16);
#define RTKHIP_JBP_LAUNCH(TP)                                                                                      \
  hipLaunchKernelGGL(joseph_bp_tile_kernel<TP>, grid, dim3(256), 0, s, dv, nu, nv, static_cast<const TP *>(dev_proj), \
                     dev_vol_out, dT.get(), vol_dim[0], vol_dim[1], vol_dim[2], vol_spacing[0], vol_spacing[1],       \
                     vol_spacing[2], infClip, supClip)
    if (proj_dtype == RTKHIP_F16)
      RTKHIP_JBP_LAUNCH(__half);
    else if (proj_dtype == RTKHIP_BF16)
      RTKHIP_JBP_LAUNCH(__hip_bfloat16);
    else
      RTKHIP_JBP_LAUNCH(float);
#undef RTKHIP_JBP_LAUNCH
    RTKHIP_CHECK_LAUNCH();
    dim3 tg((vol_dim[0] + 31) / 32, (vol_dim[2] + 31) / 32, vol_dim[1]);
    hipLaunchKernelGGL(add_transposed_kernel, tg, dim3(256), 0, s, dT.get(), dev_vol_out, vol_dim[0], vol_dim[1],
                       vol_dim[2]);
    RTKHIP_CHECK_LAUNCH();
  });
}
