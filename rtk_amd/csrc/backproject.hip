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
//  * A workgroup is 16 (x) x 16 (z) such columns.  Per view it stages the detector
//    rectangle its voxel block projects onto (bounded by the block's corners) in
//    LDS with 16-byte row loads, double-buffered, and reads its bilinear taps from
//    there: the gathers no longer go through the texture-address path, which
//    bounds a global-memory gather kernel (4 lane-loads per update).
//  * Many views per launch: the volume is read and written once per launch,
//    the per-view matrices are wave-uniform (scalar loads).
//  * Everything else (X-path views, ITK generic interpolation, non-zero
//    y-dependence of u or w) runs in a per-voxel general kernel.
#include "common.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cfloat>
#include <cmath>
#include <type_traits>
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

// Column kernel, for views whose u and w do not depend on y (c[2] == c[10] == 0)
// and that use the OPTIMIZED border rule (RTK's OptimizedBackprojectionY case).
// The row stride, 76 floats (12 banks past a multiple of 64), spreads the taps of a
// wave's 8 x 8 voxel patch, whose detector columns span about 16 and whose rows differ
// by one or two, over distinct LDS banks; it is a multiple of 4, so a lane stages its
// four pixels with one 16-byte store.
constexpr int kTileU = 64, kTileV = 48, kTileStride = 76; // staged detector footprint, pixels
constexpr int kStagePasses = (kTileV + 15) / 16;           // 16 rows per staging pass
constexpr int kPlaceViews = 64;                            // footprints placed per chunk, one per lane
// A footprint whose every tap lies at least kEdge pixels inside the detector's
// interpolation range (0 <= u, v and u < nu - 1, v < nv - 1) is flagged in the bit
// kInside of its height: then every column of the block passes the border rule and
// reads the staged rectangle, and the column loop runs without its per-voxel tests.
// kEdge is far above the rounding of the per-voxel u and v (a few ulps of ~1e3).
constexpr float kEdge = 1.f / 16.f;
constexpr int   kInside = 1 << 16;
// 48 rows hold a 16-voxel column's taps up to 2.6 detector rows per voxel (the
// 512^3 <- 1024^2 FDK reaches 2.34); each extra staging pass costs ~1 VALU per update

// Detector rectangle [u0, u0 + W) x [v0, v0 + H) holding every tap of the voxel block
// [xlo, xhi] x [ylo, yhi] x [zlo, zhi] (centred coordinates) for view c.  u, v are
// linear-fractional in the voxel position; where the denominator w keeps one sign
// over the block (positive for the FDK's normalised matrices, negative for the plain
// backprojector's) the block's corners bound them.  One pixel of margin covers
// rounding (including the 1-ulp reciprocal).  W = 0: no tap can pass the border rule
// (skip the view); W < 0: too large, or w changes sign (read global memory for this
// view).  One lane evaluates the four (x, z) corners in turn, so a wavefront places
// the footprints of 64 views at once.
__device__ __forceinline__ int4
FootprintLane(const ViewCoef & c, float xlo, float xhi, float ylo, float yhi, float zlo, float zhi, int nu, int nv)
{
  float umin = FLT_MAX, umax = -FLT_MAX, vmin = FLT_MAX, vmax = -FLT_MAX, smin = 1.f, smax = -1.f;
#pragma unroll
  for (int q = 0; q < 4; q++)
  {
    const float x = (q & 1) ? xhi : xlo, z = (q & 2) ? zhi : zlo;
    const float w = c.c[8] + c.c[9] * x + c.c[11] * z;
    const float rw = __builtin_amdgcn_rcpf(w);
    const float u = (c.c[0] + c.c[1] * x + c.c[3] * z) * rw;
    const float vl = (c.c[4] + c.c[5] * x + c.c[7] * z + c.c[6] * ylo) * rw;
    const float vh = (c.c[4] + c.c[5] * x + c.c[7] * z + c.c[6] * yhi) * rw;
    const float sgn = (w > 0.f) ? 1.f : (w < 0.f ? -1.f : 0.f);
    umin = fminf(umin, u);
    umax = fmaxf(umax, u);
    vmin = fminf(vmin, fminf(vl, vh));
    vmax = fmaxf(vmax, fmaxf(vl, vh));
    smin = fminf(smin, sgn);
    smax = fmaxf(smax, sgn);
  }
  if (!(smin == smax && smin != 0.f) || !(umin <= umax) || !(vmin <= vmax) || !(umax - umin < 1e6f) ||
      !(vmax - vmin < 1e6f))
    return make_int4(0, 0, -1, 0);
  const float fuLo = floorf(fmaxf(umin, -2.f)), fuHi = floorf(fminf(umax, float(nu)));
  const float fvLo = floorf(fmaxf(vmin, -2.f)), fvHi = floorf(fminf(vmax, float(nv)));
  if (fuHi < 0.f || fuLo > float(nu - 2) || fvHi < 0.f || fvLo > float(nv - 2))
    return make_int4(0, 0, 0, 0);
  const int  u0 = max(int(fuLo) - 1, 0) & ~3;
  const int  v0 = max(int(fvLo) - 1, 0);
  const int  u1 = min(int(fuHi) + 2, nu - 1), v1 = min(int(fvHi) + 2, nv - 1);
  const int  W = u1 - u0 + 1, H = v1 - v0 + 1;
  const bool fits = W <= kTileU && H <= kTileV;
  const bool inside = umin >= kEdge && umax <= float(nu - 1) - kEdge && vmin >= kEdge && vmax <= float(nv - 1) - kEdge;
  return make_int4(u0, v0, fits ? W : -1, H | (fits && inside ? kInside : 0));
}

// Each thread accumulates a column of NJ voxels along y; a workgroup is 16 (x) x 16
// (z) columns, a wavefront an 8 x 8 (x, z) patch.  Each view's detector footprint of
// the workgroup's voxel block is staged in LDS: the footprint of view p + 1 is loaded
// into registers before view p is backprojected from LDS and written to the other
// buffer after it, so there is one barrier per view and the loads overlap the
// arithmetic.  Taps outside the staged rectangle (rounding) and views whose
// footprint does not fit read global memory; results are identical either way.
// Four consecutive pixels from an aligned address, as FP32.
__device__ __forceinline__ void
Load4(const float * p, float * d)
{
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 f = *reinterpret_cast<const f4 *>(p);
  d[0] = f.x;
  d[1] = f.y;
  d[2] = f.z;
  d[3] = f.w;
}
template <typename TP>
__device__ __forceinline__ void
Load4(const TP * p, float * d)
{
  typedef unsigned short u4 __attribute__((ext_vector_type(4)));
  const u4 f = *reinterpret_cast<const u4 *>(p);
  const unsigned short e[4] = { f.x, f.y, f.z, f.w };
#pragma unroll
  for (int q = 0; q < 4; q++)
  {
    TP v;
    __builtin_memcpy(&v, &e[q], 2);
    d[q] = LoadPix(&v, 0);
  }
}

template <typename TP, bool FDK, int NJ, bool VEC>
__global__ __launch_bounds__(256) void bp_lds_kernel(const ViewCoef * __restrict__ views, int nviews,
                                                     const TP * __restrict__ proj, int nu, int nv,
                                                     const float * vol_in, float * vol_out, int nx, int ny, int nz,
                                                     float ci, float cj, float ck)
{
  __shared__ __align__(16) float stage[2][kTileV * kTileStride];
  __shared__ int4  rects[2][kPlaceViews]; // footprints (u0, v0, W, H) of two chunks of views
  const int        t = threadIdx.x;
  const int        bx0 = blockIdx.x * 16, bz0 = blockIdx.y * 16, j0 = blockIdx.z * NJ;
  const int        i = bx0 + (t & 7) + 8 * ((t >> 6) & 1);
  const int        k = bz0 + ((t >> 3) & 7) + 8 * (t >> 7);
  const bool       live = i < nx && k < nz;
  const long       sliceStride = long(nx) * ny;
  const long       base = long(k) * sliceStride + long(j0) * nx + i;
  const int        jn = min(NJ, ny - j0);
  // the views are summed from zero and vol_in is added at the end: no global load is
  // outstanding when the view loop starts, so the compiler's wait for the staging
  // loads stays where their data is stored
  float            acc[NJ];
#pragma unroll
  for (int jj = 0; jj < NJ; jj++)
    acc[jj] = 0.f;

  const float x = float(i) - ci, z = float(k) - ck, y0 = float(j0) - cj;
  const float xlo = float(bx0) - ci, xhi = float(min(bx0 + 15, nx - 1)) - ci;
  const float zlo = float(bz0) - ck, zhi = float(min(bz0 + 15, nz - 1)) - ck;
  const float ylo = y0, yhi = float(j0 + jn - 1) - cj;
  const long  projStride = long(nu) * nv;
  const float numax = float(nu - 1);
  // staging: lane = 4 consecutive pixels (one 16-byte load for FP32 when rows are
  // 4-aligned), 16 lanes per 64-pixel row, 16 rows per pass
  const int   sc = 4 * (t & 15), sr = t >> 4;

  float sreg[4 * kStagePasses];
  int   cu0, cv0, cW, cH; // footprint of the view being backprojected
  bool  cIn;              // ... and whether it lies inside the detector (kInside)
  // VEC (nu a multiple of 4, 16-byte aligned stack): every lane loads its four 4-pixel
  // chunks, 16 bytes (8 for 16-bit stacks) each, with no branch, so the loaded
  // registers are never copied before the arithmetic that hides their latency.  With
  // nu, u0 and sc multiples of 4, a chunk that starts inside the rectangle (sc < W)
  // lies inside its detector row; lanes outside read the view's first 4 pixels
  // instead, which store() discards with the same test.
  // the view's base address and row offsets are wave-uniform (scalar registers);
  // each lane adds one 32-bit offset (the host keeps a view under 4 GiB)
  const unsigned laneOff = unsigned(sr * nu + sc); // pixel offset of this lane's chunk in pass 0
  auto           load = [&](int p, int u0, int v0, int W, int H) {
    const long vbase = long(p) * projStride;
    if constexpr (VEC)
    {
      const TP * vp = proj + vbase;
      const bool colIn = sc < W;
#pragma unroll
      for (int q = 0; q < kStagePasses; q++)
      {
        const unsigned rowOff = unsigned((v0 + 16 * q) * nu + u0); // wave-uniform
        const unsigned o = (colIn && sr + 16 * q < H) ? rowOff + laneOff : 0u;
        Load4(vp + o, sreg + 4 * q);
      }
    }
    else
    {
#pragma unroll
      for (int q = 0; q < kStagePasses; q++)
#pragma unroll
        for (int e = 0; e < 4; e++)
        {
          const int r = sr + 16 * q;
          sreg[4 * q + e] = (r < H && sc + e < W) ? LoadPix(proj, vbase + long(v0 + r) * nu + (u0 + sc + e)) : 0.f;
        }
    }
  };
  // whole 4-pixel chunks are stored: cells at or beyond W hold pixels (or zeros) that
  // no tap reads, and sc + 3 <= 63 keeps them inside the tile
  auto store = [&](int b, int W, int H) {
#pragma unroll
    for (int q = 0; q < kStagePasses; q++)
    {
      const int r = sr + 16 * q;
      if (r < H && sc < W)
      {
#pragma unroll
        for (int e = 0; e < 4; e++)
          stage[b][r * kTileStride + sc + e] = sreg[4 * q + e];
      }
    }
  };
  // Wavefront 0 places the footprints of views [64 c, 64 c + 64) in rects[c & 1], one
  // view per lane: chunk 0 before the loop, chunk c + 1 at view 64 c.  The buffer it
  // overwrites held chunk c - 1, whose last footprint was read before the barrier
  // that ended view 64 c - 1, and chunk c + 1 is first read at view 64 c + 63.
  const int wave = t >> 6;
  auto      place = [&](int chunk) {
    const int q = chunk * kPlaceViews + (t & 63);
    if (q < nviews)
      rects[chunk & 1][t & 63] = FootprintLane(views[q], xlo, xhi, ylo, yhi, zlo, zhi, nu, nv);
  };
  auto rectOf = [&](int q, int & u0, int & v0, int & W, int & H, bool & in) {
    const int4 r = rects[(q / kPlaceViews) & 1][q % kPlaceViews];
    u0 = __builtin_amdgcn_readfirstlane(r.x);
    v0 = __builtin_amdgcn_readfirstlane(r.y);
    W = __builtin_amdgcn_readfirstlane(r.z);
    const int h = __builtin_amdgcn_readfirstlane(r.w);
    H = h & (kInside - 1);
    in = (h & kInside) != 0;
  };
  if (wave == 0)
    place(0);
  __syncthreads();
  rectOf(0, cu0, cv0, cW, cH, cIn);
  load(0, cu0, cv0, cW, cH);
  store(0, cW, cH);
  __syncthreads();
  for (int p = 0; p < nviews; p++)
  {
    const int b = p & 1;
    int       nu0 = 0, nv0 = 0, nW = 0, nH = 0;
    bool      nIn = false;
    if (p + 1 < nviews)
    {
      rectOf(p + 1, nu0, nv0, nW, nH, nIn);
      load(p + 1, nu0, nv0, nW, nH);
    }
    if (wave == 0 && p % kPlaceViews == 0 && p + kPlaceViews < nviews)
      place(p / kPlaceViews + 1);
    // per-view limits of the fast path (wave-uniform)
    const bool  viewFast = jn == NJ && cW > 0;
    const float fvLo = float(cv0), fvHi = float(min(nv - 1, cv0 + cH - 1));
    if (live && cW != 0)
    {
      const ViewCoef & c = views[p];
      const float      w = c.c[8] + c.c[9] * x + c.c[11] * z;
      const float      rw = __builtin_amdgcn_rcpf(w); // 1 ulp
      const float      u = (c.c[0] + c.c[1] * x + c.c[3] * z) * rw;
      const float      vb = (c.c[4] + c.c[5] * x + c.c[7] * z + c.c[6] * y0) * rw;
      const float      dv = c.c[6] * rw;
      const float      wgt = FDK ? rw * rw : 1.f;
      // The column's detector column pair and whether it runs from the staged
      // rectangle with no per-voxel test.  When the block's footprint lies inside
      // the detector (wave-uniform) every column does, with no test at all: u >= kEdge,
      // so truncation is the floor and fract(u) is the exact u - floor(u).
      int   lc = 0;
      float fu = 0.f;
      bool  fast = true, inU = true;
      if (!(viewFast && cIn))
      {
        const float uf = floorf(u);
        inU = uf >= 0.f && uf < numax;
        lc = inU ? int(uf) - cu0 : 0;
        fu = u - uf;
        // v is monotonic along the column (one FMA per voxel), so its two ends bound
        // every tap: when both pass the border rule and lie in the staged rectangle,
        // the whole column runs without per-voxel tests
        const float vlast = vb + float(NJ - 1) * dv;
        const float vmn = vb < vlast ? vb : vlast, vmx = vb < vlast ? vlast : vb;
        // one combined test (no short-circuit branches): the column lies inside the
        // staged rows [cv0, cv0 + cH - 1) and inside [0, nv - 1), and its detector
        // column pair inside the staged columns; cv0 >= 0 by construction
        fast = inU & viewFast & (unsigned(lc) < unsigned(cW - 1)) & (floorf(vmn) >= fvLo) & (floorf(vmx) < fvHi);
      }
      else
      {
        lc = int(u) - cu0;
        fu = __builtin_amdgcn_fractf(u);
      }
      if (fast)
      {
        int off = lc - cv0 * kTileStride;
        // opaque to the optimiser: otherwise it folds cv0 into every voxel's row index
        // and replaces the one 24-bit multiply-add per voxel by four instructions
        __asm__("" : "+v"(off));
        const float * S = stage[b];
        const float   w0 = wgt * (1.f - fu), w1 = wgt * fu;
#pragma unroll
        for (int jj = 0; jj < NJ; jj++)
        {
          const float   vv = vb + float(jj) * dv;
          // vv >= 0 on this path: truncation is the floor, and fract(vv) = vv - floor(vv)
          const float   fv = __builtin_amdgcn_fractf(vv);
          // 24-bit multiply: the LDS offset is small and the 32 x 32 form is slower
          const float * s0 = S + (int(__umul24(unsigned(int(vv)), unsigned(kTileStride))) + off);
          const float   t0 = s0[0] + fv * (s0[kTileStride] - s0[0]);
          const float   t1 = s0[1] + fv * (s0[kTileStride + 1] - s0[1]);
          acc[jj] = fmaf(w1, t1, fmaf(w0, t0, acc[jj]));
        }
      }
      else if (inU)
      {
        const float * S = stage[b];
        const TP *    pp = proj + long(p) * projStride + (lc + cu0);
#pragma unroll
        for (int jj = 0; jj < NJ; jj++)
        {
          const float vv = vb + float(jj) * dv;
          const float vf = floorf(vv);
          const int   vi = int(vf);
          if (vf >= 0.f && vi < nv - 1)
          {
            const float fv = vv - vf;
            const int   lr = vi - cv0;
            float       a, bb, cc, d;
            if (cW > 0 && unsigned(lc) < unsigned(cW - 1) && unsigned(lr) < unsigned(cH - 1))
            {
              const float * s0 = S + lr * kTileStride + lc;
              a = s0[0];
              cc = s0[1];
              bb = s0[kTileStride];
              d = s0[kTileStride + 1];
            }
            else
            {
              const long o = long(vi) * nu;
              a = LoadPix(pp, o);
              bb = LoadPix(pp, o + nu);
              cc = LoadPix(pp, o + 1);
              d = LoadPix(pp, o + nu + 1);
            }
            const float t0 = (1.f - fv) * a + fv * bb;
            const float t1 = (1.f - fv) * cc + fv * d;
            acc[jj] += wgt * ((1.f - fu) * t0 + fu * t1);
          }
        }
      }
    }
    if (p + 1 < nviews)
      store(b ^ 1, nW, nH);
    __syncthreads();
    cu0 = nu0;
    cv0 = nv0;
    cW = nW;
    cH = nH;
    cIn = nIn;
  }
  if (live)
  {
#pragma unroll
    for (int jj = 0; jj < NJ; jj++)
      if (jj < jn)
        vol_out[base + long(jj) * nx] = (vol_in ? vol_in[base + long(jj) * nx] : 0.f) + acc[jj];
  }
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
    dim3       grid((nx + 15) / 16, (nz + 15) / 16, (ny + kNJ - 1) / kNJ);
    // the vector staging path addresses a view with 32-bit byte offsets
    const bool vec = (nu & 3) == 0 && (reinterpret_cast<size_t>(p) & 15) == 0 &&
                     size_t(nu) * size_t(nv) * sizeof(TP) < (size_t(1) << 32);
    if (vec)
      hipLaunchKernelGGL((bp_lds_kernel<TP, FDK, kNJ, true>), grid, dim3(256), 0, s, dcoef, nviews, p, nu, nv, vin,
                         vout, nx, ny, nz, ci, cj, ck);
    else
      hipLaunchKernelGGL((bp_lds_kernel<TP, FDK, kNJ, false>), grid, dim3(256), 0, s, dcoef, nviews, p, nu, nv, vin,
                         vout, nx, ny, nz, ci, cj, ck);
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
