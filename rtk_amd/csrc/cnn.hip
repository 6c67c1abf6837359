// 3D residual CNN for learned CG warm starts (no counterpart in the reference).
//
// Network (fixed): six 3x3x3 convolutions, 32 channels, ReLU after the first five,
// residual output:  y = x + conv6(relu(conv5(...relu(conv1(s x))...))) / s
// where s is a fixed input scale.  Convolutions are cross-correlations with zero
// padding ("same" size): out[v][co] = b[co] + sum_t sum_ci W[t][co][ci] in[v + d(t)][ci],
// tap t = (dz*3 + dy)*3 + dx, d(t) = (dx-1, dy-1, dz-1).
//
// Storage: activations BF16 channels-last [b][z][y][x][32]; master parameters FP32,
// packed: W1[27][32], b1[32], then for layers 2..5 W[27][32 co][32 ci], b[32], then
// W6[27][32 ci], b6[1].
//
// Kernels (gfx950):
//  * 32->32 layers (forward and data gradient) are implicit GEMMs on
//    v_mfma_f32_16x16x32_bf16: M = 16 voxels along x, N = 16 output channels,
//    K = 32 input channels per tap, 27 taps accumulated in FP32.  A workgroup is
//    16 (x) x 4 (y) x 4 (z) voxels; wave w owns z-plane w and four M-tiles (y rows).
//    The block's input with its one-voxel halo is staged in LDS once; A fragments
//    are 16-byte channel chunks of neighbouring voxels read from there (1 KB per
//    instruction, conflict-free), B fragments 16-byte chunks of the weights.  Bias, ReLU
//    and the ReLU-derivative mask of the data gradient are fused into the epilogue.
//  * The 1->32 and 32->1 layers and all weight gradients run on the VALU in FP32
//    (they are 64x smaller or reduction-shaped); weight gradients reduce per block in
//    registers and once per block with FP32 atomics.
#include "common.h"

#include <cfloat>
#include <cmath>
#include <vector>

typedef __bf16 bf16_t;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

namespace rtkhip
{
namespace cnn
{

constexpr int C = 32;
constexpr int T = 27;
constexpr int kL1W = 0, kL1B = kL1W + T * C;
constexpr int kMidBase = kL1B + C;
constexpr int kMidSize = T * C * C + C;
constexpr int kL6W = kMidBase + 4 * kMidSize, kL6B = kL6W + T * C;
constexpr int kNumParams = kL6B + 1;

__device__ __forceinline__ int
TapOffset(int t, int X, int Y, int & dx, int & dy, int & dz)
{
  dx = t % 3 - 1;
  dy = (t / 3) % 3 - 1;
  dz = t / 9 - 1;
  return (dz * Y + dy) * X + dx;
}

// ------------------------------------------------------------------ layer 1
// A1[v][co] = relu(b1[co] + sum_t W1[t][co] * s * x[v + d(t)])
__global__ __launch_bounds__(256) void conv_in_kernel(const float * __restrict__ x, const float * __restrict__ p,
                                                       bf16_t * __restrict__ a1, int B, int X, int Y, int Z,
                                                       float scale)
{
  const long n = long(B) * X * Y * Z;
  const long v = long(blockIdx.x) * 256 + threadIdx.x;
  if (v >= n)
    return;
  const int i = int(v % X), j = int((v / X) % Y), k = int((v / (long(X) * Y)) % Z);
  float     xin[T];
#pragma unroll
  for (int t = 0; t < T; t++)
  {
    int       dx, dy, dz;
    const int off = TapOffset(t, X, Y, dx, dy, dz);
    const bool in = i + dx >= 0 && i + dx < X && j + dy >= 0 && j + dy < Y && k + dz >= 0 && k + dz < Z;
    xin[t] = in ? scale * x[v + off] : 0.f;
  }
  bf16x8 out[4];
#pragma unroll
  for (int co = 0; co < C; co++)
  {
    float acc = p[kL1B + co];
#pragma unroll
    for (int t = 0; t < T; t++)
      acc += p[kL1W + t * C + co] * xin[t];
    out[co / 8][co % 8] = bf16_t(fmaxf(acc, 0.f));
  }
  bf16x8 * dst = reinterpret_cast<bf16x8 *>(a1 + v * C);
#pragma unroll
  for (int q = 0; q < 4; q++)
    dst[q] = out[q];
}

// --------------------------------------------------------- 32->32 on MFMA
// out[v][co] = epilogue(b[co] + sum_t sum_ci W[t][co][ci] in[v + d(t)][ci])
// epilogue: optional ReLU; optional mask (out *= mask[v][co] > 0).
// w layout [27][32 co][32 ci] BF16.
constexpr int kRowsY = 4; // y rows (M-tiles) per wave: a block is 16 x kRowsY x 4 voxels (8: 2.09 ms vs 1.85 ms at 256^3, occupancy)

template <bool RELU, bool MASK>
__global__ __launch_bounds__(256) void conv32_mfma_kernel(const bf16_t * __restrict__ in,
                                                           const bf16_t * __restrict__ w,
                                                           const float * __restrict__ bias,
                                                           const bf16_t * __restrict__ mask, bf16_t * __restrict__ out,
                                                           int X, int Y, int Z, int tilesX, int tilesY, int tilesZ)
{
  // The block's input with a one-voxel halo, 18 (x) x 6 (y) x 6 (z) voxels of 32
  // channels, is staged in LDS once (zero outside the volume); every tap's A
  // fragment is then a conflict-free 1 KB LDS read instead of a global load.
  constexpr int HX = 18, HY = kRowsY + 2, HZ = 6, NCHUNK = HX * HY * HZ * 4;
  __shared__ bf16x8 halo[NCHUNK];
  const int  lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  long       blk = blockIdx.x;
  const int  tx = int(blk % tilesX);
  blk /= tilesX;
  const int  ty = int(blk % tilesY);
  blk /= tilesY;
  const int  tz = int(blk % tilesZ);
  const int  b = int(blk / tilesZ);
  const int  z = tz * 4 + wave;
  const int  y0 = ty * kRowsY;
  const int  xl = lane & 15;  // voxel of this lane's A row (x offset in the block)
  const int  kg = lane >> 4;  // channel group: ci = 8 kg .. 8 kg + 7
  const long volBase = long(b) * Z * Y * X;

  const bf16x8 zero8 = bf16x8{ bf16_t(0.f), bf16_t(0.f), bf16_t(0.f), bf16_t(0.f),
                               bf16_t(0.f), bf16_t(0.f), bf16_t(0.f), bf16_t(0.f) };
  for (int c = threadIdx.x; c < NCHUNK; c += 256)
  {
    const int part = c & 3, vi = c >> 2;
    const int xs = tx * 16 - 1 + vi % HX, ys = y0 - 1 + (vi / HX) % HY, zs = tz * 4 - 1 + vi / (HX * HY);
    bf16x8    v = zero8;
    if (xs >= 0 && xs < X && ys >= 0 && ys < Y && zs >= 0 && zs < Z)
      v = *reinterpret_cast<const bf16x8 *>(in + ((volBase + (long(zs) * Y + ys) * X + xs) * C + 8 * part));
    halo[c] = v;
  }
  __syncthreads();

  floatx4 acc[kRowsY][2];
#pragma unroll
  for (int r = 0; r < kRowsY; r++)
    for (int nn = 0; nn < 2; nn++)
      acc[r][nn] = floatx4{ 0.f, 0.f, 0.f, 0.f };

  for (int t = 0; t < T; t++)
  {
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const bf16x8 b0 = *reinterpret_cast<const bf16x8 *>(w + (long(t) * C + (lane & 15)) * C + 8 * kg);
    const bf16x8 b1 = *reinterpret_cast<const bf16x8 *>(w + (long(t) * C + 16 + (lane & 15)) * C + 8 * kg);
    const int    hbase = ((wave + dz + 1) * HY + dy + 1) * HX + xl + dx + 1;
#pragma unroll
    for (int r = 0; r < kRowsY; r++)
    {
      const bf16x8 a = halo[(hbase + r * HX) * 4 + kg];
      acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc[r][0], 0, 0, 0);
      acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc[r][1], 0, 0, 0);
    }
  }
  if (z >= Z)
    return;
  // D layout: lane holds rows (voxels) 4*(lane/16)+q, column (channel) lane%16.
#pragma unroll
  for (int r = 0; r < kRowsY; r++)
  {
    const int y = y0 + r;
    if (y >= Y)
      continue;
#pragma unroll
    for (int nn = 0; nn < 2; nn++)
    {
      const int   co = 16 * nn + (lane & 15);
      const float bb = bias ? bias[co] : 0.f;
#pragma unroll
      for (int q = 0; q < 4; q++)
      {
        const int x = tx * 16 + 4 * (lane >> 4) + q;
        if (x >= X)
          continue;
        const long o = (volBase + (long(z) * Y + y) * X + x) * C + co;
        float      v = acc[r][nn][q] + bb;
        if (RELU)
          v = fmaxf(v, 0.f);
        if (MASK)
          v = (float(mask[o]) > 0.f) ? v : 0.f;
        out[o] = bf16_t(v);
      }
    }
  }
}

// ------------------------------------------------------------------ layer 6
// y[v] = x[v] + (b6 + sum_t sum_ci W6[t][ci] a5[v + d(t)][ci]) / s
__global__ __launch_bounds__(256) void conv_out_kernel(const bf16_t * __restrict__ a5, const float * __restrict__ x,
                                                        const float * __restrict__ p, float * __restrict__ y, int B,
                                                        int X, int Y, int Z, float inv_scale)
{
  const long n = long(B) * X * Y * Z;
  const long v = long(blockIdx.x) * 256 + threadIdx.x;
  if (v >= n)
    return;
  const int i = int(v % X), j = int((v / X) % Y), k = int((v / (long(X) * Y)) % Z);
  float     acc = p[kL6B];
  for (int t = 0; t < T; t++)
  {
    int       dx, dy, dz;
    const int off = TapOffset(t, X, Y, dx, dy, dz);
    if (!(i + dx >= 0 && i + dx < X && j + dy >= 0 && j + dy < Y && k + dz >= 0 && k + dz < Z))
      continue;
    const bf16x8 * src = reinterpret_cast<const bf16x8 *>(a5 + (v + off) * C);
#pragma unroll
    for (int q = 0; q < 4; q++)
    {
      const bf16x8 a = src[q];
#pragma unroll
      for (int e = 0; e < 8; e++)
        acc += p[kL6W + t * C + 8 * q + e] * float(a[e]);
    }
  }
  y[v] = x[v] + acc * inv_scale;
}

// dZ5[u][ci] = (sum_t W6[t][ci] g[u - d(t)]) * (a5[u][ci] > 0), g = dy * (1/s)
__global__ __launch_bounds__(256) void conv_out_bwd_data_kernel(const float * __restrict__ dy,
                                                                 const float * __restrict__ p,
                                                                 const bf16_t * __restrict__ a5,
                                                                 bf16_t * __restrict__ dz5, int B, int X, int Y,
                                                                 int Z, float inv_scale)
{
  const long n = long(B) * X * Y * Z;
  const long u = long(blockIdx.x) * 256 + threadIdx.x;
  if (u >= n)
    return;
  const int i = int(u % X), j = int((u / X) % Y), k = int((u / (long(X) * Y)) % Z);
  float     g[T];
#pragma unroll
  for (int t = 0; t < T; t++)
  {
    int       dx, dy_, dz;
    const int off = TapOffset(t, X, Y, dx, dy_, dz);
    // v = u - d(t) must be inside
    const bool in = i - dx >= 0 && i - dx < X && j - dy_ >= 0 && j - dy_ < Y && k - dz >= 0 && k - dz < Z;
    g[t] = in ? dy[u - off] * inv_scale : 0.f;
  }
  const bf16x8 * msk = reinterpret_cast<const bf16x8 *>(a5 + u * C);
  bf16x8 *       dst = reinterpret_cast<bf16x8 *>(dz5 + u * C);
#pragma unroll
  for (int q = 0; q < 4; q++)
  {
    const bf16x8 m = msk[q];
    bf16x8       o;
#pragma unroll
    for (int e = 0; e < 8; e++)
    {
      const int ci = 8 * q + e;
      float     acc = 0.f;
#pragma unroll
      for (int t = 0; t < T; t++)
        acc += p[kL6W + t * C + ci] * g[t];
      o[e] = bf16_t(float(m[e]) > 0.f ? acc : 0.f);
    }
    dst[q] = o;
  }
}

// --------------------------------------------------------- weight gradients
// Grid-stride over voxels; every block accumulates its share in registers and
// adds it to the FP32 gradient once with atomics.

// dW[t][co][ci] += sum_v dz[v][co] a[v + d(t)][ci];  db[co] += sum_v dz[v][co]
// thread: co = tid / 8, ci = 4 (tid % 8) .. +3, all 27 taps.
__global__ __launch_bounds__(256) void wgrad32_kernel(const bf16_t * __restrict__ a, const bf16_t * __restrict__ dz,
                                                       float * __restrict__ dW, float * __restrict__ db, int B, int X,
                                                       int Y, int Z)
{
  const int  co = threadIdx.x >> 3, ci0 = 4 * (threadIdx.x & 7);
  const long n = long(B) * X * Y * Z;
  float      acc[T][4];
  float      accb = 0.f;
#pragma unroll
  for (int t = 0; t < T; t++)
    for (int e = 0; e < 4; e++)
      acc[t][e] = 0.f;
  for (long v = blockIdx.x; v < n; v += gridDim.x)
  {
    const int   i = int(v % X), j = int((v / X) % Y), k = int((v / (long(X) * Y)) % Z);
    const float g = float(dz[v * C + co]);
    accb += g;
#pragma unroll
    for (int t = 0; t < T; t++)
    {
      int       dx, dy, dzz;
      const int off = TapOffset(t, X, Y, dx, dy, dzz);
      if (i + dx >= 0 && i + dx < X && j + dy >= 0 && j + dy < Y && k + dzz >= 0 && k + dzz < Z)
      {
        const bf16_t * s = a + (v + off) * C + ci0;
#pragma unroll
        for (int e = 0; e < 4; e++)
          acc[t][e] += g * float(s[e]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < T; t++)
    for (int e = 0; e < 4; e++)
      atomicAdd(dW + (t * C + co) * C + ci0 + e, acc[t][e]);
  if ((threadIdx.x & 7) == 0)
    atomicAdd(db + co, accb);
}

// Layer 1: dW1[t][co] += sum_v dz1[v][co] s x[v + d(t)]; db1[co] += sum_v dz1[v][co].
// thread: co = tid % 32, tap group tg = tid / 32 (taps tg, tg+8, tg+16, tg+24).
__global__ __launch_bounds__(256) void wgrad_in_kernel(const float * __restrict__ x, const bf16_t * __restrict__ dz,
                                                        float * __restrict__ dW, float * __restrict__ db, int B, int X,
                                                        int Y, int Z, float scale)
{
  const int  co = threadIdx.x & 31, tg = threadIdx.x >> 5;
  const long n = long(B) * X * Y * Z;
  float      acc[4] = { 0.f, 0.f, 0.f, 0.f }, accb = 0.f;
  for (long v = blockIdx.x; v < n; v += gridDim.x)
  {
    const int   i = int(v % X), j = int((v / X) % Y), k = int((v / (long(X) * Y)) % Z);
    const float g = float(dz[v * C + co]);
    if (tg == 0)
      accb += g;
#pragma unroll
    for (int q = 0; q < 4; q++)
    {
      const int t = tg + 8 * q;
      if (t >= T)
        break;
      int       dx, dy, dzz;
      const int off = TapOffset(t, X, Y, dx, dy, dzz);
      if (i + dx >= 0 && i + dx < X && j + dy >= 0 && j + dy < Y && k + dzz >= 0 && k + dzz < Z)
        acc[q] += g * scale * x[v + off];
    }
  }
  for (int q = 0; q < 4; q++)
  {
    const int t = tg + 8 * q;
    if (t < T)
      atomicAdd(dW + t * C + co, acc[q]);
  }
  if (tg == 0)
    atomicAdd(db + co, accb);
}

// Layer 6: dW6[t][ci] += sum_v g[v] a5[v + d(t)][ci]; db6 += sum_v g[v], g = dy / s.
__global__ __launch_bounds__(256) void wgrad_out_kernel(const bf16_t * __restrict__ a5, const float * __restrict__ dy,
                                                         float * __restrict__ dW, float * __restrict__ db, int B,
                                                         int X, int Y, int Z, float inv_scale)
{
  const int  ci = threadIdx.x & 31, tg = threadIdx.x >> 5;
  const long n = long(B) * X * Y * Z;
  float      acc[4] = { 0.f, 0.f, 0.f, 0.f }, accb = 0.f;
  for (long v = blockIdx.x; v < n; v += gridDim.x)
  {
    const int   i = int(v % X), j = int((v / X) % Y), k = int((v / (long(X) * Y)) % Z);
    const float g = dy[v] * inv_scale;
    if (threadIdx.x == 0)
      accb += g;
#pragma unroll
    for (int q = 0; q < 4; q++)
    {
      const int t = tg + 8 * q;
      if (t >= T)
        break;
      int       dx, dyy, dzz;
      const int off = TapOffset(t, X, Y, dx, dyy, dzz);
      if (i + dx >= 0 && i + dx < X && j + dyy >= 0 && j + dyy < Y && k + dzz >= 0 && k + dzz < Z)
        acc[q] += g * float(a5[(v + off) * C + ci]);
    }
  }
  for (int q = 0; q < 4; q++)
  {
    const int t = tg + 8 * q;
    if (t < T)
      atomicAdd(dW + t * C + ci, acc[q]);
  }
  if (threadIdx.x == 0)
    atomicAdd(db, accb);
}

// ------------------------------------------------------------ loss, Adam, packing
// dy = 2 s^2 (y - t) / N; partial sums of s^2 (y - t)^2 (FP64, fixed order).
__global__ __launch_bounds__(256) void mse_kernel(const float * __restrict__ y, const float * __restrict__ t,
                                                   float * __restrict__ dy, long n, float scale2, double * partial)
{
  __shared__ double sh[4];
  double            acc = 0.;
  for (long i = long(blockIdx.x) * 256 + threadIdx.x; i < n; i += long(gridDim.x) * 256)
  {
    const float d = y[i] - t[i];
    dy[i] = 2.f * scale2 * d / float(n);
    acc += double(scale2) * double(d) * double(d);
  }
  for (int o = 32; o > 0; o >>= 1)
    acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0)
    sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void adam_kernel(float * __restrict__ p, const float * __restrict__ g, float * __restrict__ m,
                            float * __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n)
    return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of the four 32->32 layers: forward [t][co][ci] and data-gradient
// [t'][ci][co] = W[26 - t'][co][ci] (flipped taps, swapped channel roles).
__global__ void pack_kernel(const float * __restrict__ p, bf16_t * __restrict__ wf, bf16_t * __restrict__ wb)
{
  const int idx = blockIdx.x * 256 + threadIdx.x; // over 4 * 27 * 32 * 32
  if (idx >= 4 * T * C * C)
    return;
  const int l = idx / (T * C * C), r = idx % (T * C * C);
  const int t = r / (C * C), co = (r / C) % C, ci = r % C;
  const float wv = p[kMidBase + l * kMidSize + r];
  wf[idx] = bf16_t(wv);
  wb[l * T * C * C + ((T - 1 - t) * C + ci) * C + co] = bf16_t(wv);
}

} // namespace cnn
} // namespace rtkhip

// ------------------------------------------------------------------ handle
struct rtkhip_cnn
{
  rtkhip::DeviceBuffer<float>   params, grads, m, v, dy, out;
  rtkhip::DeviceBuffer<bf16_t>  wf, wb, act, dz0, dz1;
  rtkhip::DeviceBuffer<double>  partial;
  long                          nvox = 0; // voxels of the activations kept for backward
  int                           dims[4] = { 0, 0, 0, 0 };
  float                         scale = 1.f;
  const float *                 lastIn = nullptr;
};

namespace rtkhip
{
namespace cnn
{
namespace
{

void
Reserve(rtkhip_cnn & net, long nvox)
{
  net.act.Resize(size_t(5) * nvox * C); // A1..A5
  net.out.Resize(size_t(nvox));
}

void
ForwardImpl(rtkhip_cnn & net, int B, const int d[3], const float * in, float * out, float scale, hipStream_t s)
{
  const int  X = d[0], Y = d[1], Z = d[2];
  const long n = long(B) * X * Y * Z;
  Reserve(net, n);
  bf16_t *  A = net.act.get();
  const unsigned g = unsigned((n + 255) / 256);
  hipLaunchKernelGGL(conv_in_kernel, dim3(g), dim3(256), 0, s, in, net.params.get(), A, B, X, Y, Z, scale);
  RTKHIP_CHECK_LAUNCH();
  const int tilesX = (X + 15) / 16, tilesY = (Y + kRowsY - 1) / kRowsY, tilesZ = (Z + 3) / 4;
  const long nb = long(B) * tilesX * tilesY * tilesZ;
  RTKHIP_REQUIRE(nb < (1L << 31), RTKHIP_ERR_UNSUPPORTED, "cnn: volume too large");
  for (int l = 0; l < 4; l++)
  {
    const bf16_t * ain = A + size_t(l) * n * C;
    bf16_t *       aout = A + size_t(l + 1) * n * C;
    hipLaunchKernelGGL((conv32_mfma_kernel<true, false>), dim3(unsigned(nb)), dim3(256), 0, s, ain,
                       net.wf.get() + size_t(l) * T * C * C, net.params.get() + kMidBase + l * kMidSize + T * C * C,
                       nullptr, aout, X, Y, Z, tilesX, tilesY, tilesZ);
    RTKHIP_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(conv_out_kernel, dim3(g), dim3(256), 0, s, A + size_t(4) * n * C, in, net.params.get(), out, B,
                     X, Y, Z, 1.f / scale);
  RTKHIP_CHECK_LAUNCH();
  net.nvox = n;
  net.dims[0] = B;
  net.dims[1] = X;
  net.dims[2] = Y;
  net.dims[3] = Z;
  net.scale = scale;
  net.lastIn = in;
}

int
BlocksFor(long n)
{
  return int(std::min<long>(std::max<long>(n / 64, 1), 2048));
}

} // namespace
} // namespace cnn
} // namespace rtkhip

using namespace rtkhip;
using namespace rtkhip::cnn;

extern "C" int
rtkhip_cnn_num_params(void)
{
  return kNumParams;
}

extern "C" rtkhip_cnn *
rtkhip_cnn_create(void)
{
  rtkhip_cnn * net = nullptr;
  int          rc = Guard([&] {
    RequireDevice();
    net = new rtkhip_cnn();
    net->params.Resize(kNumParams);
    net->grads.Resize(kNumParams);
    net->m.Resize(kNumParams);
    net->v.Resize(kNumParams);
    net->wf.Resize(size_t(4) * T * C * C);
    net->wb.Resize(size_t(4) * T * C * C);
    net->partial.Resize(2048);
    RTKHIP_CHECK(hipMemset(net->params.get(), 0, sizeof(float) * kNumParams));
    RTKHIP_CHECK(hipMemset(net->m.get(), 0, sizeof(float) * kNumParams));
    RTKHIP_CHECK(hipMemset(net->v.get(), 0, sizeof(float) * kNumParams));
  });
  if (rc != RTKHIP_OK)
  {
    delete net;
    return nullptr;
  }
  return net;
}

extern "C" void
rtkhip_cnn_destroy(rtkhip_cnn * net)
{
  delete net;
}

extern "C" int
rtkhip_cnn_set_params(rtkhip_cnn * net, const float * host_params, void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(net && host_params, RTKHIP_ERR_INVALID_ARGUMENT, "cnn_set_params: null argument");
    hipStream_t s = Stream(stream);
    RTKHIP_CHECK(hipMemcpyAsync(net->params.get(), host_params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, s));
    RTKHIP_CHECK(hipMemsetAsync(net->m.get(), 0, sizeof(float) * kNumParams, s));
    RTKHIP_CHECK(hipMemsetAsync(net->v.get(), 0, sizeof(float) * kNumParams, s));
    hipLaunchKernelGGL(pack_kernel, dim3((4 * T * C * C + 255) / 256), dim3(256), 0, s, net->params.get(),
                       net->wf.get(), net->wb.get());
    RTKHIP_CHECK_LAUNCH();
    RTKHIP_CHECK(hipStreamSynchronize(s));
  });
}

extern "C" int
rtkhip_cnn_get_params(rtkhip_cnn * net, float * host_params, float * host_grads, void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(net, RTKHIP_ERR_INVALID_ARGUMENT, "cnn_get_params: null network");
    hipStream_t s = Stream(stream);
    if (host_params)
      RTKHIP_CHECK(
        hipMemcpyAsync(host_params, net->params.get(), sizeof(float) * kNumParams, hipMemcpyDeviceToHost, s));
    if (host_grads)
      RTKHIP_CHECK(
        hipMemcpyAsync(host_grads, net->grads.get(), sizeof(float) * kNumParams, hipMemcpyDeviceToHost, s));
    RTKHIP_CHECK(hipStreamSynchronize(s));
  });
}

extern "C" int
rtkhip_cnn_forward(rtkhip_cnn * net, int batch, const int dim[3], const float * dev_in, float * dev_out, float scale,
                   void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(net && dim && dev_in && dev_out && batch > 0 && scale > 0.f, RTKHIP_ERR_INVALID_ARGUMENT,
                   "cnn_forward: bad argument");
    RTKHIP_REQUIRE(dim[0] > 0 && dim[1] > 0 && dim[2] > 0, RTKHIP_ERR_INVALID_ARGUMENT, "cnn_forward: bad size");
    RTKHIP_REQUIRE(dev_in != dev_out, RTKHIP_ERR_INVALID_ARGUMENT, "cnn_forward: in-place is not supported");
    ForwardImpl(*net, batch, dim, dev_in, dev_out, scale, Stream(stream));
  });
}

extern "C" int
rtkhip_cnn_backward(rtkhip_cnn * net, const float * dev_out, const float * dev_target, double * loss, void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(net && dev_out && dev_target && net->nvox > 0 && net->lastIn, RTKHIP_ERR_INVALID_ARGUMENT,
                   "cnn_backward: call rtkhip_cnn_forward first");
    hipStream_t s = Stream(stream);
    const int   B = net->dims[0], X = net->dims[1], Y = net->dims[2], Z = net->dims[3];
    const long  n = net->nvox;
    const float sc = net->scale;
    net->dy.Resize(size_t(n));
    net->dz0.Resize(size_t(n) * C);
    net->dz1.Resize(size_t(n) * C);
    float *  P = net->params.get();
    float *  G = net->grads.get();
    bf16_t * A = net->act.get();
    RTKHIP_CHECK(hipMemsetAsync(G, 0, sizeof(float) * kNumParams, s));
    // loss and dL/dy
    const int gl = int(std::min<long>((n + 255) / 256, 2048));
    hipLaunchKernelGGL(mse_kernel, dim3(gl), dim3(256), 0, s, dev_out, dev_target, net->dy.get(), n, sc * sc,
                       net->partial.get());
    RTKHIP_CHECK_LAUNCH();
    const unsigned g = unsigned((n + 255) / 256);
    const int      wb = BlocksFor(n);
    // layer 6
    hipLaunchKernelGGL(wgrad_out_kernel, dim3(wb), dim3(256), 0, s, A + size_t(4) * n * C, net->dy.get(), G + kL6W,
                       G + kL6B, B, X, Y, Z, 1.f / sc);
    RTKHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(conv_out_bwd_data_kernel, dim3(g), dim3(256), 0, s, net->dy.get(), P, A + size_t(4) * n * C,
                       net->dz0.get(), B, X, Y, Z, 1.f / sc);
    RTKHIP_CHECK_LAUNCH();
    // layers 5..2: dZ_l in dz_cur; weight grads from A_{l-1}; data grad masked by A_{l-1}
    const int tilesX = (X + 15) / 16, tilesY = (Y + kRowsY - 1) / kRowsY, tilesZ = (Z + 3) / 4;
    const long nb = long(B) * tilesX * tilesY * tilesZ;
    bf16_t *   cur = net->dz0.get();
    bf16_t *   nxt = net->dz1.get();
    for (int l = 3; l >= 0; l--) // mid layer index l = layer (l + 2)
    {
      const bf16_t * aprev = A + size_t(l) * n * C; // A_{layer-1} = A[l]
      hipLaunchKernelGGL(wgrad32_kernel, dim3(wb), dim3(256), 0, s, aprev, cur, G + kMidBase + l * kMidSize,
                         G + kMidBase + l * kMidSize + T * C * C, B, X, Y, Z);
      RTKHIP_CHECK_LAUNCH();
      hipLaunchKernelGGL((conv32_mfma_kernel<false, true>), dim3(unsigned(nb)), dim3(256), 0, s, cur,
                         net->wb.get() + size_t(l) * T * C * C, nullptr, aprev, nxt, X, Y, Z, tilesX, tilesY,
                         tilesZ);
      RTKHIP_CHECK_LAUNCH();
      std::swap(cur, nxt);
    }
    // layer 1 (cur = dZ1)
    hipLaunchKernelGGL(wgrad_in_kernel, dim3(wb), dim3(256), 0, s, net->lastIn, cur, G + kL1W, G + kL1B, B, X, Y, Z,
                       sc);
    RTKHIP_CHECK_LAUNCH();
    if (loss)
    {
      std::vector<double> h(gl);
      RTKHIP_CHECK(hipMemcpyAsync(h.data(), net->partial.get(), sizeof(double) * gl, hipMemcpyDeviceToHost, s));
      RTKHIP_CHECK(hipStreamSynchronize(s));
      double sum = 0.;
      for (double x : h)
        sum += x;
      *loss = sum / double(n);
    }
  });
}

extern "C" int
rtkhip_cnn_adam_step(rtkhip_cnn * net, float lr, float beta1, float beta2, float eps, int step, void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(net && step >= 1, RTKHIP_ERR_INVALID_ARGUMENT, "cnn_adam_step: bad argument");
    hipStream_t s = Stream(stream);
    const float c1 = 1.f - std::pow(beta1, float(step)), c2 = 1.f - std::pow(beta2, float(step));
    hipLaunchKernelGGL(adam_kernel, dim3((kNumParams + 255) / 256), dim3(256), 0, s, net->params.get(),
                       net->grads.get(), net->m.get(), net->v.get(), kNumParams, lr, beta1, beta2, eps, c1, c2);
    RTKHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(pack_kernel, dim3((4 * T * C * C + 255) / 256), dim3(256), 0, s, net->params.get(),
                       net->wf.get(), net->wb.get());
    RTKHIP_CHECK_LAUNCH();
  });
}
