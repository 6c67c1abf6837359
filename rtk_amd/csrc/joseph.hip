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

#include <cfloat>
#include <climits>
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
    dim3         grid((nu + 15) / 16, np, (nv + 15) / 16);
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
