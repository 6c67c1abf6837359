// FDK weighting and FFT ramp filtering for gfx950.
//
// Reference behaviour:
//   weighting  include/rtkFDKWeightProjectionFilter.hxx:67-128
//   ramp       include/rtkFFTProjectionsConvolutionImageFilter.hxx:121-276 (pad,
//              R2C, multiply by the kernel row, C2R with 1/N, crop) and
//              include/rtkFFTRampImageFilter.hxx:47-178 (kernel, built on the host)
//
// Design (CDNA4): one workgroup filters one detector row entirely in LDS.  The
// zero-padded real row of length N is packed into N/2 complex points
// (even + i odd), transformed by a radix-4/2 FFT in LDS, split into the N/2+1
// half spectrum, multiplied by the kernel, re-packed and inverse transformed;
// only the cropped centre is written.  The weighting is fused into the load
// and the output may be written as FP16/BF16 for the backprojector, so the
// stack crosses HBM once in and once out.
#include "common.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace rtkhip
{
namespace
{

// Per-view weighting constants (derived on the host in FP64 from the 7 values
// of include/rtkFDKWeightProjectionFilter.hxx:85-116).
struct WeightView
{
  float x0;       // physical x of pixel 0 + projOffsetX + tan(a) (sdd - sid)
  float y0;       // physical y of row 0 + projOffsetY - sourceOffsetY
  float numpart1; // sdd (cos a + tan a sin a)
  float sina;
  float sddtana;
  float sdd2;
  float factor; // angular-gap constant factor
  float parallel;
};

__device__ __forceinline__ float
CosGammaWeight(const WeightView & w, float dx, float dy, int i, int j)
{
  if (w.parallel != 0.f)
    return w.factor;
  const float px = w.x0 + dx * float(i);
  const float py = w.y0 + dy * float(j);
  const float t = px - w.sddtana;
  const float denom = sqrtf(w.sdd2 + py * py + t * t);
  return w.factor * (w.numpart1 - px * w.sina) / denom;
}

__global__ __launch_bounds__(256) void weight_kernel(const WeightView * __restrict__ views, int nu, int nv,
                                                      float dx, float dy, const float * in, float * out)
{
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int p = blockIdx.z;
  if (i >= nu || j >= nv)
    return;
  const long o = (long(p) * nv + j) * nu + i;
  out[o] = in[o] * CosGammaWeight(views[p], dx, dy, i, j);
}

__device__ __forceinline__ float2
cmul(float2 a, float2 b)
{
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2
conj2(float2 a)
{
  return make_float2(a.x, -a.y);
}

__device__ __forceinline__ void
StoreOut(float * p, long i, float v)
{
  p[i] = v;
}
__device__ __forceinline__ void
StoreOut(__half * p, long i, float v)
{
  p[i] = __float2half(v);
}
__device__ __forceinline__ void
StoreOut(__hip_bfloat16 * p, long i, float v)
{
  p[i] = __float2bfloat16(v);
}
// Two adjacent outputs (i even) in one store.
__device__ __forceinline__ void
StoreOut2(float * p, long i, float a, float b)
{
  *reinterpret_cast<float2 *>(p + i) = make_float2(a, b);
}
__device__ __forceinline__ void
StoreOut2(__half * p, long i, float a, float b)
{
  *reinterpret_cast<__half2 *>(p + i) = __floats2half2_rn(a, b);
}
__device__ __forceinline__ void
StoreOut2(__hip_bfloat16 * p, long i, float a, float b)
{
  __hip_bfloat16 v[2] = { __float2bfloat16(a), __float2bfloat16(b) };
  unsigned       u;
  __builtin_memcpy(&u, v, 4);
  *reinterpret_cast<unsigned *>(p + i) = u;
}

// In-place DIT FFT of M points in LDS whose input is stored in bit-reversed order;
// tw[t] = exp(-2 pi i t / M) for t < 3M/4.  INV uses conj(tw).  Pairs of radix-2
// stages are fused into radix-4 stages (one barrier and one LDS pass per two
// stages); an odd log2(M) starts with one radix-2 stage.  With L = 2^s and
// w = W_{4L}^p, the four points x0..x3 at p, p+L, p+2L, p+3L of a group become
//   a0 = x0, a1 = w^2 x1, a2 = w x2, a3 = w^3 x3,
//   y0 = (a0 + a1) + (a2 + a3), y2 = (a0 + a1) - (a2 + a3),
//   y1 = (a0 - a1) -/+ i (a2 - a3), y3 = (a0 - a1) +/- i (a2 - a3)  (forward/inverse),
// which is two radix-2 stages (twiddles W_{2L}^p = w^2, then w and w W_{4L}^L).
template <bool INV>
__device__ __forceinline__ void
LdsFFT(float2 * buf, const float2 * tw, int M, int logM)
{
  int s = 0;
  if (logM & 1)
  {
    __syncthreads();
    for (int b = threadIdx.x; b < (M >> 1); b += blockDim.x)
    {
      const float2 a = buf[2 * b], c = buf[2 * b + 1];
      buf[2 * b] = make_float2(a.x + c.x, a.y + c.y);
      buf[2 * b + 1] = make_float2(a.x - c.x, a.y - c.y);
    }
    s = 1;
  }
  for (; s < logM; s += 2)
  {
    __syncthreads();
    const int L = 1 << s;
    const int shift = logM - 2 - s; // twiddle index step M / (4L) = 2^shift
    for (int b = threadIdx.x; b < (M >> 2); b += blockDim.x)
    {
      const int pos = b & (L - 1);
      const int i0 = ((b >> s) << (s + 2)) + pos;
      float2    w1 = tw[pos << shift], w2 = tw[(2 * pos) << shift], w3 = tw[(3 * pos) << shift];
      if (INV)
      {
        w1.y = -w1.y;
        w2.y = -w2.y;
        w3.y = -w3.y;
      }
      const float2 a0 = buf[i0];
      const float2 a1 = cmul(buf[i0 + L], w2);
      const float2 a2 = cmul(buf[i0 + 2 * L], w1);
      const float2 a3 = cmul(buf[i0 + 3 * L], w3);
      const float2 s01 = make_float2(a0.x + a1.x, a0.y + a1.y), d01 = make_float2(a0.x - a1.x, a0.y - a1.y);
      const float2 s23 = make_float2(a2.x + a3.x, a2.y + a3.y), d23 = make_float2(a2.x - a3.x, a2.y - a3.y);
      // -i d = (d.y, -d.x); +i d = (-d.y, d.x)
      const float2 r = INV ? make_float2(-d23.y, d23.x) : make_float2(d23.y, -d23.x);
      buf[i0] = make_float2(s01.x + s23.x, s01.y + s23.y);
      buf[i0 + L] = make_float2(d01.x + r.x, d01.y + r.y);
      buf[i0 + 2 * L] = make_float2(s01.x - s23.x, s01.y - s23.y);
      buf[i0 + 3 * L] = make_float2(d01.x - r.x, d01.y - r.y);
    }
  }
  __syncthreads();
}

__device__ __forceinline__ int
BitRev(int m, int logM)
{
  return int(__brev(unsigned(m)) >> (32 - logM));
}

// One workgroup per (row, view).  tables: [0, 3M/4) FFT twiddles exp(-2 pi i t/M),
// [3M/4, 3M/4+M+1) split twiddles exp(-2 pi i k/N), then kernel K[k] for k in [0, M].
template <typename TO>
__global__ __launch_bounds__(256) void ramp_kernel(const float * __restrict__ in, TO * __restrict__ out, int nu,
                                                    int nv, int N, int logM, int zeroext,
                                                    const float2 * __restrict__ tables, int next,
                                                    const float * __restrict__ mirror,
                                                    const WeightView * __restrict__ wviews, float dx, float dy)
{
  extern __shared__ float2 lds[];
  const int M = N >> 1;
  float2 *  buf = lds;          // M: forward transform
  float2 *  buf2 = lds + M;     // M: inverse transform
  float2 *  tw = lds + 2 * M;   // 3M/4
  const int j = blockIdx.x;
  const int p = blockIdx.y;
  const float * row = in + (long(p) * nv + j) * nu;
  TO *          orow = out + (long(p) * nv + j) * nu;
  const int ntw = (3 * M) >> 2;
  for (int t = threadIdx.x; t < ntw; t += blockDim.x)
    tw[t] = tables[t];
  const WeightView * wv = wviews ? wviews + p : nullptr;
  auto               W = [&](int i) -> float { return wv ? row[i] * CosGammaWeight(*wv, dx, dy, i, j) : row[i]; };
  auto               X = [&](int n) -> float {
    const int i = n - zeroext;
    if (i >= 0 && i < nu)
      return W(i);
    if (next > 0)
    {
      if (i < 0 && -i <= next)
        return mirror[-i] * (2.f * W(1) - W(-i));
      if (i >= nu && i - (nu - 1) <= next)
      {
        const int b = i - (nu - 1);
        return mirror[b] * (2.f * W(nu - 1) - W(nu - 1 - b));
      }
    }
    return 0.f;
  };
  for (int m = threadIdx.x; m < M; m += blockDim.x)
    buf[BitRev(m, logM)] = make_float2(X(2 * m), X(2 * m + 1));
  LdsFFT<false>(buf, tw, M, logM);

  // Split to the half spectrum, multiply by the kernel, re-pack for the inverse
  // into the second buffer (bit-reversed for the inverse DIT FFT).
  const float2 * split = tables + ntw;
  const float2 * K = split + (M + 1);
  const int      npairs = (M >> 1) + 1; // k = 0 .. M/2, partner M-k
  for (int k = threadIdx.x; k < npairs; k += blockDim.x)
  {
    const int    km = (M - k) & (M - 1);
    const float2 Zk = buf[k & (M - 1)], Zm = buf[km];
    // X[k] = (Zk + conj Zm)/2 + W^k (Zk - conj Zm)/(2i), and the same for M-k.
    auto Xof = [&](float2 A, float2 B, float2 w) {
      const float2 e = make_float2(0.5f * (A.x + B.x), 0.5f * (A.y - B.y));
      const float2 o = make_float2(0.5f * (A.y + B.y), -0.5f * (A.x - B.x));
      const float2 wo = cmul(w, o);
      return make_float2(e.x + wo.x, e.y + wo.y);
    };
    float2 Xk, Xm;
    if (k == 0)
    {
      Xk = make_float2(Zk.x + Zk.y, 0.f); // X[0]
      Xm = make_float2(Zk.x - Zk.y, 0.f); // X[M] (Nyquist)
    }
    else
    {
      Xk = Xof(Zk, Zm, split[k]);
      Xm = Xof(Zm, Zk, split[M - k]);
    }
    const float2 Yk = cmul(Xk, K[k]);
    const float2 Ym = cmul(Xm, K[M - k]);
    // Z'[k] = Ye + i Yo, Ye = (Y[k] + conj Y[M-k])/2, Yo = (Y[k] - conj Y[M-k]) conj(W^k)/2.
    auto Zof = [&](float2 A, float2 B, float2 w) {
      const float2 e = make_float2(0.5f * (A.x + B.x), 0.5f * (A.y - B.y));
      const float2 d = make_float2(0.5f * (A.x - B.x), 0.5f * (A.y + B.y));
      const float2 o = cmul(d, conj2(w));
      return make_float2(e.x - o.y, e.y + o.x);
    };
    if (k < M)
      buf2[BitRev(k, logM)] = Zof(Yk, Ym, split[k]);
    const int mk = M - k;
    if (mk != k && mk < M)
      buf2[BitRev(mk, logM)] = Zof(Ym, Yk, split[mk]);
  }
  LdsFFT<true>(buf2, tw, M, logM);
  const float scale = 1.f / float(M);
  for (int i = threadIdx.x; i < nu; i += blockDim.x)
  {
    const int    n = zeroext + i;
    const float2 z = buf2[n >> 1];
    StoreOut(orow, i, ((n & 1) ? z.y : z.x) * scale);
  }
}

// ------------------------------------------------------------ register FFTs --
// Wavefront-per-row variant for M = N/2 in {256, 512, 1024, 2048}: each of the 64
// lanes holds M/64 complex points in registers; the M-point FFT runs as a mixed-radix
// Stockham autosort FFT (Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC 2008: FftIteration with expand()), whose
// radix-R butterflies are register DFTs and whose passes exchange data through LDS.
// One wavefront per workgroup: the barriers between passes cost nothing and a row
// crosses LDS four times per transform instead of once per radix-4 stage.

constexpr int
BitRevC(int i, int bits)
{
  int r = 0;
  for (int b = 0; b < bits; b++)
    r |= ((i >> b) & 1) << (bits - 1 - b);
  return r;
}

// R-point DFT (R = 4, 8, 16) of v in registers: radix-2 DIT after a compile-time bit
// reversal; W_16^m = exp(-2 pi i m / 16), conjugated for the inverse.
template <int R, bool INV>
__device__ __forceinline__ void
DftReg(float2 (&v)[R])
{
  constexpr int   bits = R == 4 ? 2 : (R == 8 ? 3 : 4);
  constexpr float c1 = 0.923879532511286756f, s1 = 0.382683432365089772f, h = 0.707106781186547524f;
  constexpr float wr[8] = { 1.f, c1, h, s1, 0.f, -s1, -h, -c1 };
  constexpr float wi[8] = { 0.f, -s1, -h, -c1, -1.f, -c1, -h, -s1 };
  float2          t[R];
#pragma unroll
  for (int i = 0; i < R; i++)
    t[i] = v[BitRevC(i, bits)];
#pragma unroll
  for (int len = 2; len <= R; len <<= 1)
  {
#pragma unroll
    for (int i = 0; i < R; i += len)
    {
#pragma unroll
      for (int k = 0; k < len / 2; k++)
      {
        const int m = k * (16 / len);
        float2    b = t[i + k + len / 2];
        if (m == 4) // -i (forward), +i (inverse)
          b = INV ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
        else if (m != 0)
          b = cmul(b, make_float2(wr[m], INV ? -wi[m] : wi[m]));
        const float2 a = t[i + k];
        t[i + k] = make_float2(a.x + b.x, a.y + b.y);
        t[i + k + len / 2] = make_float2(a.x - b.x, a.y - b.y);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < R; i++)
    v[i] = t[i];
}

// One Stockham pass of radix R after passes whose radices multiply to NS: butterfly j
// reads x[j + r M/R], multiplies by W_{NS R}^{r (j mod NS)} (tw[t] = exp(-2 pi i t/M)),
// transforms, and writes y[(j - j mod NS) R + j mod NS + r NS].  Each lane does the
// butterflies j = lane + 64 b; all of them are read before any is written, so the
// pass can work in place in one LDS buffer.
template <int M, int R, int NS, bool INV, typename LoadF, typename StoreF>
__device__ __forceinline__ void
StockhamPass(LoadF && load, StoreF && store, const float2 * __restrict__ tw, int lane)
{
  constexpr int B = M / (R * 64);
  float2        v[B][R];
#pragma unroll
  for (int b = 0; b < B; b++)
#pragma unroll
    for (int r = 0; r < R; r++)
      v[b][r] = load(lane + 64 * b + r * (M / R));
  __syncthreads();
#pragma unroll
  for (int b = 0; b < B; b++)
  {
    const int j = lane + 64 * b, k = j & (NS - 1);
    if constexpr (NS > 1)
    {
#pragma unroll
      for (int r = 1; r < R; r++)
      {
        float2 w = tw[r * k * (M / (NS * R))];
        if (INV)
          w.y = -w.y;
        v[b][r] = cmul(v[b][r], w);
      }
    }
    DftReg<R, INV>(v[b]);
    const int d = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; r++)
      store(d + r * NS, v[b][r]);
  }
}

template <int M, bool INV, typename LoadF, typename StoreF, typename LdsLoadF, typename LdsStoreF>
__device__ __forceinline__ void
StockhamFFT(LoadF && first, StoreF && last, LdsLoadF && lload, LdsStoreF && lstore, const float2 * tw, int lane)
{
  if constexpr (M == 256)
  {
    StockhamPass<M, 4, 1, INV>(first, lstore, tw, lane);
    __syncthreads();
    StockhamPass<M, 4, 4, INV>(lload, lstore, tw, lane);
    __syncthreads();
    StockhamPass<M, 4, 16, INV>(lload, lstore, tw, lane);
    __syncthreads();
    StockhamPass<M, 4, 64, INV>(lload, last, tw, lane);
  }
  else if constexpr (M == 512)
  {
    StockhamPass<M, 8, 1, INV>(first, lstore, tw, lane);
    __syncthreads();
    StockhamPass<M, 8, 8, INV>(lload, lstore, tw, lane);
    __syncthreads();
    StockhamPass<M, 8, 64, INV>(lload, last, tw, lane);
  }
  else if constexpr (M == 1024)
  {
    StockhamPass<M, 16, 1, INV>(first, lstore, tw, lane);
    __syncthreads();
    StockhamPass<M, 16, 16, INV>(lload, lstore, tw, lane);
    __syncthreads();
    StockhamPass<M, 4, 256, INV>(lload, last, tw, lane);
  }
  else
  {
    static_assert(M == 2048, "register FFT sizes: 256, 512, 1024, 2048");
    StockhamPass<M, 16, 1, INV>(first, lstore, tw, lane);
    __syncthreads();
    StockhamPass<M, 16, 16, INV>(lload, lstore, tw, lane);
    __syncthreads();
    StockhamPass<M, 8, 256, INV>(lload, last, tw, lane);
  }
}

// Same filter as ramp_kernel, one wavefront per (row, view).  tables: [0, M) FFT
// twiddles exp(-2 pi i t/M), [M, 2M+1) split twiddles exp(-2 pi i k/N), then the
// kernel K[k] for k in [0, M].
// TRUNC: the Ohnesorge mirror extends the row into the padding (next > 0); without it
// a padded sample is in-range data or zero, loaded without branches.
template <typename TO, int M, bool TRUNC>
__global__ __launch_bounds__(64) void ramp_wave_kernel(const float * __restrict__ in, TO * __restrict__ out, int nu,
                                                        int nv, int zeroext, const float2 * __restrict__ tables,
                                                        int next, const float * __restrict__ mirror,
                                                        const WeightView * __restrict__ wviews, float dx, float dy)
{
  __shared__ float4 buf4[M / 2];
  float2 *          buf = reinterpret_cast<float2 *>(buf4);
  const int         lane = threadIdx.x;
  const int         j = blockIdx.x;
  const int         p = blockIdx.y;
  const float *     row = in + (long(p) * nv + j) * nu;
  TO *              orow = out + (long(p) * nv + j) * nu;
  const float2 *    tw = tables;
  const float2 *    split = tables + M;
  const float2 *    K = split + (M + 1);
  const WeightView * wv = wviews ? wviews + p : nullptr;
  // cos(gamma) weighting with the row's constant part hoisted and one reciprocal
  // square root (1 ulp) per pixel
  float wx0 = 0.f, wsddtana = 0.f, wnum = 0.f, wsina = 0.f, wfac = 1.f, wC = 0.f;
  bool  wcos = false;
  if (wv)
  {
    wfac = wv->factor;
    if (wv->parallel == 0.f)
    {
      wcos = true;
      const float py = wv->y0 + dy * float(j);
      wx0 = wv->x0;
      wsddtana = wv->sddtana;
      wnum = wv->numpart1;
      wsina = wv->sina;
      wC = wv->sdd2 + py * py;
    }
  }
  auto W = [&](int i) -> float {
    if (!wcos)
      return row[i] * wfac;
    const float px = wx0 + dx * float(i), t = px - wsddtana;
    return row[i] * (wfac * (wnum - px * wsina) * __builtin_amdgcn_rsqf(wC + t * t));
  };
  auto X = [&](int n) -> float {
    const int i = n - zeroext;
    if (!TRUNC)
      return unsigned(i) < unsigned(nu) ? W(min(max(i, 0), nu - 1)) : 0.f;
    if (i >= 0 && i < nu)
      return W(i);
    {
      if (i < 0 && -i <= next)
        return mirror[-i] * (2.f * W(1) - W(-i));
      if (i >= nu && i - (nu - 1) <= next)
      {
        const int b = i - (nu - 1);
        return mirror[b] * (2.f * W(nu - 1) - W(nu - 1 - b));
      }
    }
    return 0.f;
  };
  // the zero-padded, weighted row goes to LDS first (four reals per lane and step),
  // so the first FFT pass does not hold its register block across global loads
  for (int q = lane; q < M / 2; q += 64)
    buf4[q] = make_float4(X(4 * q), X(4 * q + 1), X(4 * q + 2), X(4 * q + 3));
  __syncthreads();
  auto lload = [&](int m) { return buf[m]; };
  auto lstore = [&](int m, float2 v) { buf[m] = v; };
  StockhamFFT<M, false>(lload, lstore, lload, lstore, tw, lane);
  __syncthreads();

  // Split to the half spectrum, multiply by the kernel and re-pack, in place: lane
  // k owns the pair (k, M - k).
  for (int k = lane; k <= (M >> 1); k += 64)
  {
    const int    km = (M - k) & (M - 1);
    const float2 Zk = buf[k & (M - 1)], Zm = buf[km];
    auto         Xof = [&](float2 A, float2 Bv, float2 w) {
      const float2 e = make_float2(0.5f * (A.x + Bv.x), 0.5f * (A.y - Bv.y));
      const float2 o = make_float2(0.5f * (A.y + Bv.y), -0.5f * (A.x - Bv.x));
      const float2 wo = cmul(w, o);
      return make_float2(e.x + wo.x, e.y + wo.y);
    };
    float2 Xk, Xm;
    if (k == 0)
    {
      Xk = make_float2(Zk.x + Zk.y, 0.f);
      Xm = make_float2(Zk.x - Zk.y, 0.f);
    }
    else
    {
      Xk = Xof(Zk, Zm, split[k]);
      Xm = Xof(Zm, Zk, split[M - k]);
    }
    const float2 Yk = cmul(Xk, K[k]);
    const float2 Ym = cmul(Xm, K[M - k]);
    auto         Zof = [&](float2 A, float2 Bv, float2 w) {
      const float2 e = make_float2(0.5f * (A.x + Bv.x), 0.5f * (A.y - Bv.y));
      const float2 d = make_float2(0.5f * (A.x - Bv.x), 0.5f * (A.y + Bv.y));
      const float2 o = cmul(d, conj2(w));
      return make_float2(e.x - o.y, e.y + o.x);
    };
    if (k < M)
      buf[k] = Zof(Yk, Ym, split[k]);
    const int mk = M - k;
    if (mk != k && mk < M)
      buf[mk] = Zof(Ym, Yk, split[mk]);
  }
  __syncthreads();
  const float scale = 1.f / float(M);
  auto        gstore = [&](int m, float2 z) {
    const int i0 = 2 * m - zeroext;
    if (i0 >= 0 && i0 < nu)
      StoreOut(orow, i0, z.x * scale);
    if (i0 + 1 >= 0 && i0 + 1 < nu)
      StoreOut(orow, i0 + 1, z.y * scale);
  };
  StockhamFFT<M, true>(lload, gstore, lload, lstore, tw, lane);
}

// ------------------------------------------------- persistent wavefront FFT --
// The twiddles of one Stockham pass for the butterflies of one lane, held in
// registers: w[b][r - 1] = W_M^{r k M / (NS R)} for butterfly j = lane + 64 b and
// k = j mod NS (the factors StockhamPass reads from its table).
// With several butterflies per lane (B > 1, the last pass: NS R = M, k = lane + 64 b)
// only W^{r lane} is kept per lane and W^{r k} = W^{r lane} W^{64 r b}, whose second
// factor is the same on every lane (scalar registers).
template <int M, int R, int NS>
struct PassTwiddles
{
  static constexpr int B = M / (R * 64);
  static_assert(B == 1 || NS * R == M, "several butterflies per lane only in the last pass");
  float2 base[R - 1];
  float2 fac[B][R - 1];
  __device__ __forceinline__ void load(const float2 * __restrict__ tw, int lane)
  {
    const int k = lane & (NS - 1);
#pragma unroll
    for (int r = 1; r < R; r++)
    {
      base[r - 1] = tw[r * k * (M / (NS * R))];
#pragma unroll
      for (int b = 1; b < B; b++)
        fac[b][r - 1] = tw[(r * 64 * b) & (M - 1)];
    }
  }
  __device__ __forceinline__ float2 w(int b, int r) const
  {
    return b == 0 ? base[r - 1] : cmul(base[r - 1], fac[b][r - 1]);
  }
};

// StockhamPass with the twiddles from registers (NS > 1).
template <int M, int R, int NS, bool INV, typename LoadF, typename StoreF>
__device__ __forceinline__ void
StockhamPassReg(LoadF && load, StoreF && store, const PassTwiddles<M, R, NS> & T, int lane)
{
  constexpr int B = M / (R * 64);
  float2        v[B][R];
#pragma unroll
  for (int b = 0; b < B; b++)
#pragma unroll
    for (int r = 0; r < R; r++)
      v[b][r] = load(lane + 64 * b + r * (M / R));
  __syncthreads();
#pragma unroll
  for (int b = 0; b < B; b++)
  {
    const int j = lane + 64 * b, k = j & (NS - 1);
#pragma unroll
    for (int r = 1; r < R; r++)
    {
      float2 w = T.w(b, r);
      if (INV)
        w.y = -w.y;
      v[b][r] = cmul(v[b][r], w);
    }
    DftReg<R, INV>(v[b]);
    const int d = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; r++)
      store(d + r * NS, v[b][r]);
  }
}

// The passes of StockhamFFT for M = 256, 512, 1024 with every twiddle of the lane in
// registers (9, 14 and 27 complex values).
template <int M>
struct WaveTwiddles;
template <>
struct WaveTwiddles<256>
{
  PassTwiddles<256, 4, 4>  a;
  PassTwiddles<256, 4, 16> b;
  PassTwiddles<256, 4, 64> c;
  __device__ __forceinline__ void load(const float2 * tw, int lane)
  {
    a.load(tw, lane);
    b.load(tw, lane);
    c.load(tw, lane);
  }
  template <bool INV, typename F, typename L, typename LL, typename LS>
  __device__ __forceinline__ void fft(F && first, L && last, LL && lload, LS && lstore, const float2 * tw,
                                      int lane) const
  {
    StockhamPass<256, 4, 1, INV>(first, lstore, tw, lane);
    __syncthreads();
    StockhamPassReg<256, 4, 4, INV>(lload, lstore, a, lane);
    __syncthreads();
    StockhamPassReg<256, 4, 16, INV>(lload, lstore, b, lane);
    __syncthreads();
    StockhamPassReg<256, 4, 64, INV>(lload, last, c, lane);
  }
};
template <>
struct WaveTwiddles<512>
{
  PassTwiddles<512, 8, 8>  a;
  PassTwiddles<512, 8, 64> b;
  __device__ __forceinline__ void load(const float2 * tw, int lane)
  {
    a.load(tw, lane);
    b.load(tw, lane);
  }
  template <bool INV, typename F, typename L, typename LL, typename LS>
  __device__ __forceinline__ void fft(F && first, L && last, LL && lload, LS && lstore, const float2 * tw,
                                      int lane) const
  {
    StockhamPass<512, 8, 1, INV>(first, lstore, tw, lane);
    __syncthreads();
    StockhamPassReg<512, 8, 8, INV>(lload, lstore, a, lane);
    __syncthreads();
    StockhamPassReg<512, 8, 64, INV>(lload, last, b, lane);
  }
};
template <>
struct WaveTwiddles<1024>
{
  PassTwiddles<1024, 16, 16> a;
  PassTwiddles<1024, 4, 256> b;
  __device__ __forceinline__ void load(const float2 * tw, int lane)
  {
    a.load(tw, lane);
    b.load(tw, lane);
  }
  template <bool INV, typename F, typename L, typename LL, typename LS>
  __device__ __forceinline__ void fft(F && first, L && last, LL && lload, LS && lstore, const float2 * tw,
                                      int lane) const
  {
    StockhamPass<1024, 16, 1, INV>(first, lstore, tw, lane);
    __syncthreads();
    StockhamPassReg<1024, 16, 16, INV>(lload, lstore, a, lane);
    __syncthreads();
    StockhamPassReg<1024, 4, 256, INV>(lload, last, b, lane);
  }
};

// ramp_wave_kernel for a real ramp kernel (K[k] real, as for every window of
// include/rtkFFTRampImageFilter.hxx: the spatial kernel is even) and M <= 1024, as
// persistent wavefronts: each wavefront filters rows blockIdx.x, blockIdx.x +
// gridDim.x, ... with every table value it needs loaded once into registers (the
// FFT twiddles, and for the split step exp(-2 pi i k / N) and K[k], K[M - k] of its
// pairs k = 1 + lane + 64 i), so the rows read only the stack.  Same tables and
// arithmetic as ramp_wave_kernel, except that the imaginary part of K (zero) is not
// multiplied and split[M - k] = -conj(split[k]) is not read.
// VEC (no truncation correction, nu and the zero extension multiples of 4, 16-byte
// aligned rows): four pixels per 16-byte load, two outputs per store.
template <typename TO, int M, bool TRUNC, bool VEC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void ramp_rows_kernel(const float * __restrict__ in, TO * __restrict__ out, int nu,
                                                        int nv, long nrows, int zeroext,
                                                        const float2 * __restrict__ tables, int next,
                                                        const float * __restrict__ mirror,
                                                        const WeightView * __restrict__ wviews, float dx, float dy)
{
  __shared__ float4 buf4[M / 2];
  __shared__ float  kre[M + 1]; // the kernel, real
  float2 *          buf = reinterpret_cast<float2 *>(buf4);
  const int         lane = threadIdx.x;
  const float2 *    tw = tables;
  const float2 *    split = tables + M;
  const float2 *    K = split + (M + 1);
  constexpr int     SPL = (M / 2) / 64; // split pairs per lane
  WaveTwiddles<M>   T;
  T.load(tw, lane);
  for (int k = lane; k <= M; k += 64)
    kre[k] = K[k].x;
  // split[1 + lane + 64 it] = split[1 + lane] exp(-2 pi i 64 it / N), N = 2M: the
  // second factor is split[64 it], the same on every lane
  const float2 sw0 = split[1 + lane];
  float2       swf[SPL];
#pragma unroll
  for (int it = 0; it < SPL; it++)
    swf[it] = split[64 * it];
  const float scale = 1.f / float(M);

  for (long rowIdx = blockIdx.x; rowIdx < nrows; rowIdx += gridDim.x)
  {
    const int     p = int(rowIdx / nv), j = int(rowIdx - long(p) * nv);
    const float * row = in + rowIdx * nu;
    TO *          orow = out + rowIdx * nu;
    const WeightView * wv = wviews ? wviews + p : nullptr;
    float wx0 = 0.f, wsddtana = 0.f, wnum = 0.f, wsina = 0.f, wfac = 1.f, wC = 0.f;
    bool  wcos = false;
    if (wv)
    {
      wfac = wv->factor;
      if (wv->parallel == 0.f)
      {
        wcos = true;
        const float py = wv->y0 + dy * float(j);
        wx0 = wv->x0;
        wsddtana = wv->sddtana;
        wnum = wv->numpart1;
        wsina = wv->sina;
        wC = wv->sdd2 + py * py;
      }
    }
    auto W = [&](int i) -> float {
      if (!wcos)
        return row[i] * wfac;
      const float px = wx0 + dx * float(i), t = px - wsddtana;
      return row[i] * (wfac * (wnum - px * wsina) * __builtin_amdgcn_rsqf(wC + t * t));
    };
    auto X = [&](int n) -> float {
      const int i = n - zeroext;
      if (!TRUNC)
        return unsigned(i) < unsigned(nu) ? W(min(max(i, 0), nu - 1)) : 0.f;
      if (i >= 0 && i < nu)
        return W(i);
      {
        if (i < 0 && -i <= next)
          return mirror[-i] * (2.f * W(1) - W(-i));
        if (i >= nu && i - (nu - 1) <= next)
        {
          const int b = i - (nu - 1);
          return mirror[b] * (2.f * W(nu - 1) - W(nu - 1 - b));
        }
      }
      return 0.f;
    };
    // the previous row's last LDS reads are done before its buffer is overwritten
    __syncthreads();
    if constexpr (VEC)
    {
      for (int q = lane; q < M / 2; q += 64)
      {
        const int i = 4 * q - zeroext; // a chunk lies wholly inside the row or outside it
        float4    v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (unsigned(i) < unsigned(nu))
        {
          v = *reinterpret_cast<const float4 *>(row + i);
          if (wcos)
          {
            auto wt = [&](int e) {
              const float px = wx0 + dx * float(i + e), t = px - wsddtana;
              return wfac * (wnum - px * wsina) * __builtin_amdgcn_rsqf(wC + t * t);
            };
            v = make_float4(v.x * wt(0), v.y * wt(1), v.z * wt(2), v.w * wt(3));
          }
          else
            v = make_float4(v.x * wfac, v.y * wfac, v.z * wfac, v.w * wfac);
        }
        buf4[q] = v;
      }
    }
    else
      for (int q = lane; q < M / 2; q += 64)
        buf4[q] = make_float4(X(4 * q), X(4 * q + 1), X(4 * q + 2), X(4 * q + 3));
    __syncthreads();
    auto lload = [&](int m) { return buf[m]; };
    auto lstore = [&](int m, float2 v) { buf[m] = v; };
    T.template fft<false>(lload, lstore, lload, lstore, tw, lane);
    __syncthreads();

    // split to the half spectrum, multiply by the kernel and re-pack, in place: lane
    // pairs (k, M - k) for k = 1 + lane + 64 it; k = 0 by lane 0 (split[0] = 1)
    auto Xof = [](float2 A, float2 Bv, float2 w) {
      const float2 e = make_float2(0.5f * (A.x + Bv.x), 0.5f * (A.y - Bv.y));
      const float2 o = make_float2(0.5f * (A.y + Bv.y), -0.5f * (A.x - Bv.x));
      const float2 wo = cmul(w, o);
      return make_float2(e.x + wo.x, e.y + wo.y);
    };
    auto Zof = [](float2 A, float2 Bv, float2 w) {
      const float2 e = make_float2(0.5f * (A.x + Bv.x), 0.5f * (A.y - Bv.y));
      const float2 d = make_float2(0.5f * (A.x - Bv.x), 0.5f * (A.y + Bv.y));
      const float2 o = cmul(d, conj2(w));
      return make_float2(e.x - o.y, e.y + o.x);
    };
#pragma unroll
    for (int it = 0; it < SPL; it++)
    {
      const int    k = 1 + lane + 64 * it, mk = M - k;
      const float2 Zk = buf[k], Zm = buf[mk];
      const float2 s = it == 0 ? sw0 : cmul(sw0, swf[it]);
      const float2 sm = make_float2(-s.x, s.y); // split[M - k] = -conj(split[k])
      const float2 Xk = Xof(Zk, Zm, s), Xm = Xof(Zm, Zk, sm);
      const float  kk = kre[k], km = kre[mk];
      const float2 Yk = make_float2(Xk.x * kk, Xk.y * kk);
      const float2 Ym = make_float2(Xm.x * km, Xm.y * km);
      buf[k] = Zof(Yk, Ym, s);
      if (mk != k)
        buf[mk] = Zof(Ym, Yk, sm);
    }
    if (lane == 0)
    {
      const float2 Z0 = buf[0];
      const float2 Y0 = make_float2((Z0.x + Z0.y) * kre[0], 0.f), YM = make_float2((Z0.x - Z0.y) * kre[M], 0.f);
      buf[0] = Zof(Y0, YM, make_float2(1.f, 0.f));
    }
    __syncthreads();
    auto gstore = [&](int m, float2 z) {
      const int i0 = 2 * m - zeroext;
      if constexpr (VEC)
      {
        if (unsigned(i0) < unsigned(nu)) // i0 even, nu even: both outputs or neither
          StoreOut2(orow, i0, z.x * scale, z.y * scale);
      }
      else
      {
        if (i0 >= 0 && i0 < nu)
          StoreOut(orow, i0, z.x * scale);
        if (i0 + 1 >= 0 && i0 + 1 < nu)
          StoreOut(orow, i0 + 1, z.y * scale);
      }
    };
    T.template fft<true>(lload, gstore, lload, lstore, tw, lane);
  }
}

// ------------------------------------------------------- mixed-radix FFTs --
// Padded lengths whose prime factors are all at most 13: RTK pads to such lengths
// when its FFT is FFTW, and its CUDA convolution filter always does (greatest prime
// factor 13, include/rtkCudaFFTProjectionsConvolutionImageFilter.hxx:42).  These run
// as complex Stockham autosort FFTs in LDS (the same pass as StockhamPass, with the
// butterfly's radix chosen at run time from 2, 3, 4, 5, 7, 11, 13 and the R-point DFT
// done directly from the R roots of unity held in registers), ping-ponging between two
// LDS buffers.  A filtered row or column is real, so two of them travel as the real and
// imaginary parts of one complex sequence: for a real kernel h the spectrum K is
// Hermitian, and IFFT(K FFT(a + i b)) = (h * a) + i (h * b) exactly.
constexpr int kMaxRadices = 24;

// One pass of radix R over `batch` transforms of N points stored back to back:
// butterfly j reads src[j + r N/R], multiplies by W_{NS R}^{r (j mod NS)} (tw[t] =
// exp(-2 pi i t / N)), applies the R-point DFT and writes dst[(j - j mod NS) R + j mod NS
// + r NS].
template <int R, bool INV>
__device__ __forceinline__ void
MixedPass(const float2 * __restrict__ src, float2 * __restrict__ dst, const float2 * __restrict__ tw, int N, int NS,
          int batch)
{
  const int NR = N / R, step = N / (NS * R);
  float2    root[R];
#pragma unroll
  for (int m = 0; m < R; m++)
    root[m] = INV ? conj2(tw[m * NR]) : tw[m * NR];
  for (int jj = threadIdx.x; jj < batch * NR; jj += blockDim.x)
  {
    const int      f = jj / NR, j = jj - f * NR, k = j % NS;
    const float2 * sp = src + f * N;
    float2 *       dp = dst + f * N;
    float2         v[R];
#pragma unroll
    for (int r = 0; r < R; r++)
      v[r] = sp[j + r * NR];
#pragma unroll
    for (int r = 1; r < R; r++)
      v[r] = cmul(v[r], INV ? conj2(tw[r * k * step]) : tw[r * k * step]);
    const int o = (j - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; q++)
    {
      float2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; r++)
      {
        const float2 t = cmul(v[r], root[(r * q) % R]);
        acc.x += t.x;
        acc.y += t.y;
      }
      dp[o + q * NS] = acc;
    }
  }
}

// `batch` N-point FFTs of a (b is the other buffer); returns the buffer holding the
// result.  The caller synchronises before (input written) and after (result read).
template <bool INV>
__device__ float2 *
MixedFFT(float2 * a, float2 * b, const float2 * __restrict__ tw, int N, const int * __restrict__ rad, int nrad,
         int batch)
{
  int NS = 1;
  for (int s = 0; s < nrad; s++)
  {
    const int R = rad[s];
    switch (R)
    {
      case 2:
        MixedPass<2, INV>(a, b, tw, N, NS, batch);
        break;
      case 3:
        MixedPass<3, INV>(a, b, tw, N, NS, batch);
        break;
      case 4:
        MixedPass<4, INV>(a, b, tw, N, NS, batch);
        break;
      case 5:
        MixedPass<5, INV>(a, b, tw, N, NS, batch);
        break;
      case 7:
        MixedPass<7, INV>(a, b, tw, N, NS, batch);
        break;
      case 11:
        MixedPass<11, INV>(a, b, tw, N, NS, batch);
        break;
      default:
        MixedPass<13, INV>(a, b, tw, N, NS, batch);
        break;
    }
    __syncthreads();
    float2 * t = a;
    a = b;
    b = t;
    NS *= R;
  }
  return a;
}

// Sample n of a weighted, padded row (zero padding, optional Ohnesorge mirror), as the
// LDS and wavefront kernels build it.
__device__ __forceinline__ float
PaddedSample(const float * row, const WeightView * wv, float dx, float dy, int j, int nu, int zeroext, int next,
             const float * mirror, int n)
{
  auto      W = [&](int i) -> float { return wv ? row[i] * CosGammaWeight(*wv, dx, dy, i, j) : row[i]; };
  const int i = n - zeroext;
  if (i >= 0 && i < nu)
    return W(i);
  if (next > 0)
  {
    if (i < 0 && -i <= next)
      return mirror[-i] * (2.f * W(1) - W(-i));
    if (i >= nu && i - (nu - 1) <= next)
    {
      const int b = i - (nu - 1);
      return mirror[b] * (2.f * W(nu - 1) - W(nu - 1 - b));
    }
  }
  return 0.f;
}

// Ramp filter of two detector rows per workgroup (rows 2 blockIdx.x and 2 blockIdx.x + 1
// of the slab, in (view, row) order) for any padded length N with prime factors <= 13.
// K: the full N-point kernel spectrum.  LDS: 2N complex values.
template <typename TO>
__global__ __launch_bounds__(256) void ramp_mixed_kernel(const float * __restrict__ in, TO * __restrict__ out, int nu,
                                                          int nv, long nrows, int N, int zeroext,
                                                          const float2 * __restrict__ tw,
                                                          const float2 * __restrict__ K, const int * __restrict__ rad,
                                                          int nrad, int next, const float * __restrict__ mirror,
                                                          const WeightView * __restrict__ wviews, float dx, float dy)
{
  extern __shared__ float2 lds[];
  const long r0 = 2 * long(blockIdx.x);
  const bool two = r0 + 1 < nrows;
  const int  p0 = int(r0 / nv), j0 = int(r0 % nv);
  const int  p1 = two ? int((r0 + 1) / nv) : p0, j1 = two ? int((r0 + 1) % nv) : j0;
  const float * row0 = in + r0 * nu;
  const float * row1 = in + (two ? r0 + 1 : r0) * nu;
  const WeightView * wv0 = wviews ? wviews + p0 : nullptr;
  const WeightView * wv1 = wviews ? wviews + p1 : nullptr;
  for (int n = threadIdx.x; n < N; n += blockDim.x)
    lds[n] = make_float2(PaddedSample(row0, wv0, dx, dy, j0, nu, zeroext, next, mirror, n),
                         two ? PaddedSample(row1, wv1, dx, dy, j1, nu, zeroext, next, mirror, n) : 0.f);
  __syncthreads();
  float2 * z = MixedFFT<false>(lds, lds + N, tw, N, rad, nrad, 1);
  float2 * other = z == lds ? lds + N : lds;
  for (int k = threadIdx.x; k < N; k += blockDim.x)
    z[k] = cmul(z[k], K[k]);
  __syncthreads();
  z = MixedFFT<true>(z, other, tw, N, rad, nrad, 1);
  const float scale = 1.f / float(N);
  TO *        o0 = out + r0 * nu;
  TO *        o1 = out + (r0 + 1) * nu;
  for (int i = threadIdx.x; i < nu; i += blockDim.x)
  {
    const float2 v = z[zeroext + i];
    StoreOut(o0, i, v.x * scale);
    if (two)
      StoreOut(o1, i, v.y * scale);
  }
}

// Filters every detector column (along v) of view blockIdx.y, columns 2F blockIdx.x ..
// 2F blockIdx.x + 2F - 1: each column is zero padded at its end to N samples, circularly
// convolved through its N-point spectrum with K (full spectrum) and cropped back to nv.
// Columns travel in pairs as complex sequences; F pairs per workgroup, 2 F N complex
// values of LDS.  Every workgroup reads all of its columns before writing any, and
// workgroups own disjoint columns, so in == out is allowed (FP32).
template <typename TO>
__global__ __launch_bounds__(256) void column_filter_kernel(const float * in, TO * out, int nu, int nv, int N, int F,
                                                             const float2 * __restrict__ tw,
                                                             const float2 * __restrict__ K,
                                                             const int * __restrict__ rad, int nrad)
{
  extern __shared__ float2 lds[];
  const int     p = blockIdx.y, u0 = 2 * F * blockIdx.x, C = 2 * F;
  const float * vin = in + long(p) * nv * nu;
  TO *          vout = out + long(p) * nv * nu;
  float *       ldsf = reinterpret_cast<float *>(lds);
  for (int e = threadIdx.x; e < N * C; e += blockDim.x)
  {
    const int v = e / C, c = e - v * C, u = u0 + c;
    // pair f = c / 2 of sample v: real part for even c, imaginary for odd c
    ldsf[2 * ((c >> 1) * N + v) + (c & 1)] = (v < nv && u < nu) ? vin[long(v) * nu + u] : 0.f;
  }
  __syncthreads();
  float2 * z = MixedFFT<false>(lds, lds + F * N, tw, N, rad, nrad, F);
  float2 * other = z == lds ? lds + F * N : lds;
  for (int e = threadIdx.x; e < F * N; e += blockDim.x)
    z[e] = cmul(z[e], K[e % N]);
  __syncthreads();
  z = MixedFFT<true>(z, other, tw, N, rad, nrad, F);
  const float   scale = 1.f / float(N);
  const float * zf = reinterpret_cast<const float *>(z);
  for (int e = threadIdx.x; e < nv * C; e += blockDim.x)
  {
    const int v = e / C, c = e - v * C, u = u0 + c;
    if (u < nu)
      StoreOut(vout, long(v) * nu + u, zf[2 * ((c >> 1) * N + v) + (c & 1)] * scale);
  }
}

// Radices of an N-point transform (4s first, then 2, 3, 5, 7, 11, 13); empty if N has
// a prime factor above 13.
std::vector<int>
MixedRadices(int N)
{
  std::vector<int> r;
  while (N % 4 == 0)
  {
    r.push_back(4);
    N /= 4;
  }
  for (int q : { 2, 3, 5, 7, 11, 13 })
    while (N % q == 0)
    {
      r.push_back(q);
      N /= q;
    }
  if (N != 1 || r.size() > size_t(kMaxRadices))
    r.clear();
  return r;
}

// Device tables of an N-point mixed-radix filter: tw[0, N) = exp(-2 pi i t / N), then the
// N-point spectrum K, then the radices (as ints, after the float2 values).
struct MixedTables
{
  DeviceBuffer<float2> buf;
  int                  nrad = 0;
  const float2 *       tw() const { return buf.get(); }
  const float2 *       K(int N) const { return buf.get() + N; }
  const int *          rad(int N) const { return reinterpret_cast<const int *>(buf.get() + 2 * N); }
};

// Uploads the tables; spectrum(k) gives K[k] for k in [0, N) as doubles (re, im).
template <class F>
void
UploadMixedTables(MixedTables & t, int N, const std::vector<int> & rad, F && spectrum, hipStream_t s)
{
  const double        pi = 3.141592653589793238462643383279502884;
  const size_t        nr2 = (rad.size() + 1) / 2; // radices packed two per float2 slot
  std::vector<float2> h(2 * size_t(N) + nr2);
  for (int q = 0; q < N; q++)
  {
    h[q] = make_float2(float(std::cos(-2 * pi * q / N)), float(std::sin(-2 * pi * q / N)));
    double re = 0., im = 0.;
    spectrum(q, re, im);
    h[N + q] = make_float2(float(re), float(im));
  }
  std::memcpy(reinterpret_cast<int *>(h.data() + 2 * N), rad.data(), rad.size() * sizeof(int));
  t.buf.Resize(h.size());
  t.nrad = int(rad.size());
  RTKHIP_CHECK(hipMemcpyAsync(t.buf.get(), h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice, s));
}

size_t
MaxLds()
{
  int dev = 0, maxLds = 0;
  RTKHIP_CHECK(hipGetDevice(&dev));
  RTKHIP_CHECK(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  return size_t(maxLds);
}

std::vector<WeightView>
MakeWeightViews(int np, const double * g7, const double origin[2])
{
  std::vector<WeightView> v(np);
  for (int p = 0; p < np; p++)
  {
    const double * g = g7 + 7 * p;
    const double   sdd = g[0], sid = g[1], pox = g[2], poy = g[3], soy = g[4], fac = g[5], tilt = g[6];
    WeightView &   w = v[p];
    w.factor = float(fac);
    if (sdd == 0.)
    {
      w.parallel = 1.f;
      continue;
    }
    const double cosa = std::cos(tilt), sina = std::sin(tilt), tana = std::tan(tilt);
    w.parallel = 0.f;
    w.x0 = float(origin[0] + pox + tana * (sdd - sid));
    w.y0 = float(origin[1] + poy - soy);
    w.numpart1 = float(sdd * (cosa + tana * sina));
    w.sina = float(sina);
    w.sddtana = float(sdd * tana);
    w.sdd2 = float(sdd * sdd);
  }
  return v;
}

struct RampCache
{
  std::map<hipStream_t, DeviceBuffer<WeightView>> views;
  std::map<hipStream_t, DeviceBuffer<float2>>     tables;
  std::map<hipStream_t, DeviceBuffer<float2>>     waveTables;
  std::map<hipStream_t, DeviceBuffer<float>>      mirror;
  std::map<hipStream_t, MixedTables>              mixedRows, mixedColumns;
};
thread_local RampCache g_ramp;

} // namespace
} // namespace rtkhip

extern "C" int
rtkhip_fdk_weight(const int proj_dim[3], const double * geom7, const double proj_origin[2],
                  const double proj_spacing[2], const float * dev_in, float * dev_out, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(proj_dim && geom7 && proj_origin && proj_spacing && dev_in && dev_out,
                   RTKHIP_ERR_INVALID_ARGUMENT, "fdk_weight: null argument");
    const int nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
    RTKHIP_REQUIRE(nu > 0 && nv > 0 && np >= 0 && np <= 65535, RTKHIP_ERR_INVALID_ARGUMENT,
                   "fdk_weight: bad sizes");
    RequireDevice();
    if (np == 0)
      return;
    hipStream_t s = Stream(stream);
    auto        hv = MakeWeightViews(np, geom7, proj_origin);
    auto &      dv = g_ramp.views[s];
    dv.Resize(np);
    RTKHIP_CHECK(hipMemcpyAsync(dv.get(), hv.data(), sizeof(WeightView) * np, hipMemcpyHostToDevice, s));
    dim3 grid((nu + 63) / 64, (nv + 3) / 4, np);
    hipLaunchKernelGGL(weight_kernel, grid, dim3(256), 0, s, dv.get(), nu, nv, float(proj_spacing[0]),
                       float(proj_spacing[1]), dev_in, dev_out);
    RTKHIP_CHECK_LAUNCH();
  });
}

extern "C" int
rtkhip_ramp_filter(const int proj_dim[3], int xpad, const double * kernel, int trunc_next,
                   const double * mirror_weights, const double * weight_geom7, const double proj_origin[2],
                   const double proj_spacing[2], const float * dev_in, void * dev_out, int out_dtype,
                   void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(proj_dim && kernel && dev_in && dev_out, RTKHIP_ERR_INVALID_ARGUMENT, "ramp_filter: null argument");
    const int nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
    RTKHIP_REQUIRE(nu > 1 && nv > 0 && np >= 0 && np <= 65535 && nv <= 2147483647, RTKHIP_ERR_INVALID_ARGUMENT,
                   "ramp_filter: bad sizes");
    const bool pow2 = xpad >= 4 && (xpad & (xpad - 1)) == 0;
    const std::vector<int> radices = pow2 ? std::vector<int>() : MixedRadices(xpad);
    RTKHIP_REQUIRE(xpad >= 2 && xpad <= 8192 && xpad >= nu && (pow2 || !radices.empty()), RTKHIP_ERR_UNSUPPORTED,
                   "ramp_filter: padded size must lie in [max(2, Nu), 8192] and have no prime factor above 13");
    RTKHIP_REQUIRE(trunc_next >= 0 && (trunc_next == 0 || (mirror_weights && trunc_next < nu - 1)),
                   RTKHIP_ERR_INVALID_ARGUMENT, "ramp_filter: bad truncation arguments");
    RTKHIP_REQUIRE(out_dtype == RTKHIP_F32 || out_dtype == RTKHIP_F16 || out_dtype == RTKHIP_BF16,
                   RTKHIP_ERR_INVALID_ARGUMENT, "ramp_filter: unknown output dtype");
    RTKHIP_REQUIRE(!weight_geom7 || (proj_origin && proj_spacing), RTKHIP_ERR_INVALID_ARGUMENT,
                   "ramp_filter: fused weighting needs origin and spacing");
    RTKHIP_REQUIRE((dev_in != dev_out) || out_dtype == RTKHIP_F32, RTKHIP_ERR_INVALID_ARGUMENT,
                   "ramp_filter: in-place filtering needs an FP32 output");
    RequireDevice();
    if (np == 0)
      return;
    hipStream_t s = Stream(stream);
    const int   M = xpad / 2;
    int         logM = 0;
    while ((1 << logM) < M)
      logM++;
    const long zeroext = (long(xpad) - nu) / 2;
    RTKHIP_REQUIRE(zeroext >= trunc_next, RTKHIP_ERR_INVALID_ARGUMENT, "ramp_filter: mirror wider than padding");
    const double pi = 3.141592653589793238462643383279502884;
    const float * dmirror = nullptr;
    if (trunc_next > 0)
    {
      std::vector<float> mw(trunc_next + 1);
      for (int b = 0; b <= trunc_next; b++)
        mw[b] = float(mirror_weights[b]);
      auto & dm = g_ramp.mirror[s];
      dm.Resize(mw.size());
      RTKHIP_CHECK(hipMemcpyAsync(dm.get(), mw.data(), sizeof(float) * mw.size(), hipMemcpyHostToDevice, s));
      dmirror = dm.get();
    }
    const WeightView * dw = nullptr;
    float              dx = 0.f, dy = 0.f;
    if (weight_geom7)
    {
      auto   hv = MakeWeightViews(np, weight_geom7, proj_origin);
      auto & dv = g_ramp.views[s];
      dv.Resize(np);
      RTKHIP_CHECK(hipMemcpyAsync(dv.get(), hv.data(), sizeof(WeightView) * np, hipMemcpyHostToDevice, s));
      dw = dv.get();
      dx = float(proj_spacing[0]);
      dy = float(proj_spacing[1]);
    }
    if (!pow2)
    {
      // any 13-smooth length: the full spectrum from the half one (K[N - k] = conj K[k])
      const int N = xpad;
      auto &    t = g_ramp.mixedRows[s];
      UploadMixedTables(
        t, N, radices,
        [&](int k, double & re, double & im) {
          const int  h = k <= N / 2 ? k : N - k;
          const bool c = k > N / 2;
          re = kernel[2 * h];
          im = c ? -kernel[2 * h + 1] : kernel[2 * h + 1];
        },
        s);
      const size_t lds = sizeof(float2) * 2 * size_t(N);
      RTKHIP_REQUIRE(lds <= MaxLds(), RTKHIP_ERR_UNSUPPORTED, "ramp_filter: padded row does not fit in LDS");
      const long nrows = long(nv) * np;
      RTKHIP_REQUIRE((nrows + 1) / 2 <= 2147483647L, RTKHIP_ERR_UNSUPPORTED, "ramp_filter: too many rows");
      dim3 grid(unsigned((nrows + 1) / 2));
#define RTKHIP_RAMPM(TO)                                                                                             \
  hipLaunchKernelGGL(ramp_mixed_kernel<TO>, grid, dim3(256), lds, s, dev_in, static_cast<TO *>(dev_out), nu, nv,    \
                     nrows, N, int(zeroext), t.tw(), t.K(N), t.rad(N), t.nrad, trunc_next, dmirror, dw, dx, dy)
      if (out_dtype == RTKHIP_F32)
        RTKHIP_RAMPM(float);
      else if (out_dtype == RTKHIP_F16)
        RTKHIP_RAMPM(__half);
      else
        RTKHIP_RAMPM(__hip_bfloat16);
#undef RTKHIP_RAMPM
      RTKHIP_CHECK_LAUNCH();
      return;
    }
    const int           ntw = (3 * M) / 4;
    std::vector<float2> tab(ntw + (M + 1) + (M + 1));
    for (int t = 0; t < ntw; t++)
      tab[t] = make_float2(float(std::cos(-2 * pi * t / M)), float(std::sin(-2 * pi * t / M)));
    for (int k = 0; k <= M; k++)
      tab[ntw + k] = make_float2(float(std::cos(-2 * pi * k / xpad)), float(std::sin(-2 * pi * k / xpad)));
    for (int k = 0; k <= M; k++)
      tab[ntw + M + 1 + k] = make_float2(float(kernel[2 * k]), float(kernel[2 * k + 1]));
    auto & dt = g_ramp.tables[s];
    dt.Resize(tab.size());
    RTKHIP_CHECK(hipMemcpyAsync(dt.get(), tab.data(), sizeof(float2) * tab.size(), hipMemcpyHostToDevice, s));
    if (M == 256 || M == 512 || M == 1024 || M == 2048)
    {
      // wavefront-per-row register FFT; its own table layout
      std::vector<float2> wt(M + 2 * (M + 1));
      for (int t = 0; t < M; t++)
        wt[t] = make_float2(float(std::cos(-2 * pi * t / M)), float(std::sin(-2 * pi * t / M)));
      for (int k = 0; k <= M; k++)
      {
        wt[M + k] = make_float2(float(std::cos(-2 * pi * k / xpad)), float(std::sin(-2 * pi * k / xpad)));
        wt[2 * M + 1 + k] = make_float2(float(kernel[2 * k]), float(kernel[2 * k + 1]));
      }
      auto & dwt = g_ramp.waveTables[s];
      dwt.Resize(wt.size());
      RTKHIP_CHECK(hipMemcpyAsync(dwt.get(), wt.data(), sizeof(float2) * wt.size(), hipMemcpyHostToDevice, s));
      // a real kernel spectrum (the spatial kernel is even: every window of the ramp
      // filter) and M <= 1024: persistent wavefronts with their tables in registers
      double kre = 0., kim = 0.;
      for (int k = 0; k <= M; k++)
      {
        kre = std::max(kre, std::abs(kernel[2 * k]));
        kim = std::max(kim, std::abs(kernel[2 * k + 1]));
      }
      if (M <= 1024 && kim <= 1e-9 * kre)
      {
        const long nrows = long(nv) * np;
        int        dev = 0, cus = 0, perCu = 0;
        RTKHIP_CHECK(hipGetDevice(&dev));
        RTKHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        // rows of 16-byte aligned 4-pixel chunks, wholly inside the row or in the padding
        const bool vec = trunc_next == 0 && nu % 4 == 0 && zeroext % 4 == 0 &&
                         reinterpret_cast<uintptr_t>(dev_in) % 16 == 0 &&
                         reinterpret_cast<uintptr_t>(dev_out) % 16 == 0;
#define RTKHIP_RAMPR(TO, MM, TR, VE)                                                                                 \
  do                                                                                                                 \
  {                                                                                                                  \
    RTKHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, ramp_rows_kernel<TO, MM, TR, VE>, 64, 0));     \
    const long g = std::min<long>(nrows, long(std::max(cus, 1)) * std::max(perCu, 1));                               \
    hipLaunchKernelGGL((ramp_rows_kernel<TO, MM, TR, VE>), dim3(unsigned(g)), dim3(64), 0, s, dev_in,                 \
                       static_cast<TO *>(dev_out), nu, nv, nrows, int(zeroext), dwt.get(), trunc_next, dmirror, dw,   \
                       dx, dy);                                                                                      \
  } while (0)
#define RTKHIP_RAMPR_T(TO, MM)                                                                                       \
  do                                                                                                                 \
  {                                                                                                                  \
    if (trunc_next > 0)                                                                                              \
      RTKHIP_RAMPR(TO, MM, true, false);                                                                             \
    else if (vec)                                                                                                    \
      RTKHIP_RAMPR(TO, MM, false, true);                                                                             \
    else                                                                                                             \
      RTKHIP_RAMPR(TO, MM, false, false);                                                                            \
  } while (0)
#define RTKHIP_RAMPR_M(TO)                                                                                           \
  do                                                                                                                 \
  {                                                                                                                  \
    if (M == 256)                                                                                                    \
      RTKHIP_RAMPR_T(TO, 256);                                                                                       \
    else if (M == 512)                                                                                               \
      RTKHIP_RAMPR_T(TO, 512);                                                                                       \
    else                                                                                                             \
      RTKHIP_RAMPR_T(TO, 1024);                                                                                      \
  } while (0)
        if (out_dtype == RTKHIP_F32)
          RTKHIP_RAMPR_M(float);
        else if (out_dtype == RTKHIP_F16)
          RTKHIP_RAMPR_M(__half);
        else
          RTKHIP_RAMPR_M(__hip_bfloat16);
#undef RTKHIP_RAMPR_M
#undef RTKHIP_RAMPR_T
#undef RTKHIP_RAMPR
        RTKHIP_CHECK_LAUNCH();
        return;
      }
      dim3 wgrid(nv, np);
#define RTKHIP_RAMPW(TO, MM)                                                                                         \
  do                                                                                                                 \
  {                                                                                                                  \
    if (trunc_next > 0)                                                                                              \
      hipLaunchKernelGGL((ramp_wave_kernel<TO, MM, true>), wgrid, dim3(64), 0, s, dev_in,                            \
                         static_cast<TO *>(dev_out), nu, nv, int(zeroext), dwt.get(), trunc_next, dmirror, dw, dx,   \
                         dy);                                                                                        \
    else                                                                                                             \
      hipLaunchKernelGGL((ramp_wave_kernel<TO, MM, false>), wgrid, dim3(64), 0, s, dev_in,                           \
                         static_cast<TO *>(dev_out), nu, nv, int(zeroext), dwt.get(), trunc_next, dmirror, dw, dx,   \
                         dy);                                                                                        \
  } while (0)
#define RTKHIP_RAMPW_M(TO)                                                                                           \
  if (M == 256)                                                                                                      \
    RTKHIP_RAMPW(TO, 256);                                                                                           \
  else if (M == 512)                                                                                                 \
    RTKHIP_RAMPW(TO, 512);                                                                                           \
  else if (M == 1024)                                                                                                \
    RTKHIP_RAMPW(TO, 1024);                                                                                          \
  else                                                                                                               \
    RTKHIP_RAMPW(TO, 2048);
      if (out_dtype == RTKHIP_F32)
      {
        RTKHIP_RAMPW_M(float)
      }
      else if (out_dtype == RTKHIP_F16)
      {
        RTKHIP_RAMPW_M(__half)
      }
      else
      {
        RTKHIP_RAMPW_M(__hip_bfloat16)
      }
#undef RTKHIP_RAMPW_M
#undef RTKHIP_RAMPW
      RTKHIP_CHECK_LAUNCH();
      return;
    }
    const size_t lds = sizeof(float2) * (2 * M + (3 * M) / 4);
    const size_t maxLds = MaxLds();
    RTKHIP_REQUIRE(lds <= maxLds, RTKHIP_ERR_UNSUPPORTED,
                   "ramp_filter: padded row does not fit in LDS (" + std::to_string(lds) + " > " +
                     std::to_string(maxLds) + " bytes)");
    dim3         grid(nv, np);
#define RTKHIP_RAMP_LAUNCH(TO)                                                                                       \
  hipLaunchKernelGGL(ramp_kernel<TO>, grid, dim3(256), lds, s, dev_in, static_cast<TO *>(dev_out), nu, nv, xpad,    \
                     logM, int(zeroext), dt.get(), trunc_next, dmirror, dw, dx, dy)
    if (out_dtype == RTKHIP_F32)
      RTKHIP_RAMP_LAUNCH(float);
    else if (out_dtype == RTKHIP_F16)
      RTKHIP_RAMP_LAUNCH(__half);
    else
      RTKHIP_RAMP_LAUNCH(__hip_bfloat16);
#undef RTKHIP_RAMP_LAUNCH
    RTKHIP_CHECK_LAUNCH();
  });
}

extern "C" int
rtkhip_column_filter(const int proj_dim[3], int ypad, const double * kernel, const float * dev_in, void * dev_out,
                     int out_dtype, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(proj_dim && kernel && dev_in && dev_out, RTKHIP_ERR_INVALID_ARGUMENT,
                   "column_filter: null argument");
    const int nu = proj_dim[0], nv = proj_dim[1], np = proj_dim[2];
    RTKHIP_REQUIRE(nu > 0 && nv > 0 && np >= 0 && np <= 65535, RTKHIP_ERR_INVALID_ARGUMENT,
                   "column_filter: bad sizes");
    const std::vector<int> radices = MixedRadices(ypad);
    RTKHIP_REQUIRE(ypad >= 2 && ypad <= 8192 && ypad >= nv && !radices.empty(), RTKHIP_ERR_UNSUPPORTED,
                   "column_filter: padded size must lie in [max(2, Nv), 8192] and have no prime factor above 13");
    RTKHIP_REQUIRE(out_dtype == RTKHIP_F32 || out_dtype == RTKHIP_F16 || out_dtype == RTKHIP_BF16,
                   RTKHIP_ERR_INVALID_ARGUMENT, "column_filter: unknown output dtype");
    RTKHIP_REQUIRE((dev_in != dev_out) || out_dtype == RTKHIP_F32, RTKHIP_ERR_INVALID_ARGUMENT,
                   "column_filter: in-place filtering needs an FP32 output");
    RequireDevice();
    if (np == 0)
      return;
    hipStream_t s = Stream(stream);
    const int   N = ypad;
    auto &      t = g_ramp.mixedColumns[s];
    UploadMixedTables(
      t, N, radices,
      [&](int k, double & re, double & im) {
        re = kernel[2 * k];
        im = kernel[2 * k + 1];
      },
      s);
    // pairs of columns per workgroup: 64 KB of LDS or one pair
    const int    F = std::max(1, std::min(8, 4096 / N));
    const size_t lds = sizeof(float2) * 2 * size_t(F) * N;
    RTKHIP_REQUIRE(lds <= MaxLds(), RTKHIP_ERR_UNSUPPORTED, "column_filter: padded column does not fit in LDS");
    dim3 grid(unsigned((nu + 2 * F - 1) / (2 * F)), unsigned(np));
#define RTKHIP_COLF(TO)                                                                                              \
  hipLaunchKernelGGL(column_filter_kernel<TO>, grid, dim3(256), lds, s, dev_in, static_cast<TO *>(dev_out), nu, nv,  \
                     N, F, t.tw(), t.K(N), t.rad(N), t.nrad)
    if (out_dtype == RTKHIP_F32)
      RTKHIP_COLF(float);
    else if (out_dtype == RTKHIP_F16)
      RTKHIP_COLF(__half);
    else
      RTKHIP_COLF(__hip_bfloat16);
#undef RTKHIP_COLF
    RTKHIP_CHECK_LAUNCH();
  });
}
