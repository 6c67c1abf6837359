// Host side of librtkhip: geometry, per-view matrices and the filter-level
// entry points (the work RTK's filter classes do around their kernels).
//
//   geometry           src/rtkThreeDCircularProjectionGeometry.cxx:55-153,407-637
//   BP matrices        include/rtkBackProjectionImageFilter.hxx:535-595, FDK
//                      normalisation include/rtkFDKBackProjectionImageFilter.hxx:79-104
//   Joseph rays        include/rtkProjectionsRegionConstIteratorRayBasedWithFlatPanel.hxx:40-71
//   FDK weights        include/rtkFDKWeightProjectionFilter.hxx:40-63
//   ramp kernel        include/rtkFFTRampImageFilter.hxx:47-178 and padding
//                      include/rtkFFTProjectionsConvolutionImageFilter.hxx:280-355
//   FDK pipeline       include/rtkFDKConeBeamReconstructionFilter.hxx:116-157
//   CG pipeline        include/rtkConjugateGradientConeBeamReconstructionFilter.hxx:195-305,
//                      include/rtkReconstructionConjugateGradientOperator.hxx:194-286,
//                      include/rtkConjugateGradientImageFilter.hxx:108-258
//
// All matrices are built in FP64 on the host; kernels receive buffer-relative
// indices (the ITK buffered-region start is folded into the matrices).
#include "common.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

namespace rtkhip
{

constexpr double kPi = 3.141592653589793238462643383279502884;

namespace
{
thread_local std::string g_lastError;
} // namespace

void
SetLastError(const std::string & msg)
{
  g_lastError = msg;
}

void
RequireDevice()
{
  int        n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
  {
    (void)hipGetLastError();
    throw Error(RTKHIP_ERR_NO_DEVICE, "no HIP device available");
  }
}

int ProjectQuadricsLaunch(int nu, int nv, int np, const double * rays16, int nshapes, const double * shapes,
                          const float * in, float * out, hipStream_t s);

// ------------------------------------------------------------- small algebra --
using M4 = std::array<double, 16>; // row-major 4x4
using M3 = std::array<double, 9>;  // row-major 3x3

M4
Mul(const M4 & a, const M4 & b)
{
  M4 c{};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++)
        c[4 * i + j] += a[4 * i + k] * b[4 * k + j];
  return c;
}

M4
Identity4()
{
  M4 m{};
  m[0] = m[5] = m[10] = m[15] = 1.;
  return m;
}

// Rigid/affine inverse of a homogeneous matrix whose 3x3 block is invertible.
M4
InverseAffine(const M4 & m)
{
  const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  M3           inv = { (e * i - f * h) / det, (c * h - b * i) / det, (b * f - c * e) / det,
                       (f * g - d * i) / det, (a * i - c * g) / det, (c * d - a * f) / det,
                       (d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det };
  M4           r{};
  for (int rr = 0; rr < 3; rr++)
  {
    for (int cc = 0; cc < 3; cc++)
      r[4 * rr + cc] = inv[3 * rr + cc];
    r[4 * rr + 3] = -(inv[3 * rr] * m[3] + inv[3 * rr + 1] * m[7] + inv[3 * rr + 2] * m[11]);
  }
  r[15] = 1.;
  return r;
}

double
Wrap2Pi(double a)
{
  return a - 2 * kPi * std::floor(a / (2 * kPi));
}

// Index (buffer-relative) -> physical point matrix of an image.
M4
BufferIndexToPhysical(const rtkhip_image & im)
{
  M4 m{};
  for (int r = 0; r < 3; r++)
  {
    m[5 * r] = im.spacing[r];
    m[4 * r + 3] = im.origin[r] + im.spacing[r] * im.index[r];
  }
  m[15] = 1.;
  return m;
}

} // namespace rtkhip

// ------------------------------------------------------------------ geometry --
struct rtkhip_geometry
{
  static constexpr double kPi = rtkhip::kPi;
  struct View
  {
    double sid, sdd, gantry, projX, projY, oop, ip, srcX, srcY; // angles in radians, wrapped
    rtkhip::M4 rot;                                             // homogeneous rotation R
    double     source[3];                                       // R^T (srcX, srcY, sid)
    double     sourceAngle;
  };
  std::vector<View> views;
  double            radius = 0.;

  // Rotation of src/rtkThreeDCircularProjectionGeometry.cxx:121 written in
  // closed form: R = Rz(-ip) Rx(-oop) Ry(-gantry) (documentation/docs/Geometry.md:60-90).
  static rtkhip::M4 Rotation(double oop, double ga, double ip)
  {
    const double cx = std::cos(-oop), sx = std::sin(-oop);
    const double cy = std::cos(-ga), sy = std::sin(-ga);
    const double cz = std::cos(-ip), sz = std::sin(-ip);
    // Rz * Rx * Ry expanded.
    rtkhip::M4 m{};
    m[0] = cz * cy - sz * sx * sy;
    m[1] = -sz * cx;
    m[2] = cz * sy + sz * sx * cy;
    m[4] = sz * cy + cz * sx * sy;
    m[5] = cz * cx;
    m[6] = sz * sy - cz * sx * cy;
    m[8] = -cx * sy;
    m[9] = sx;
    m[10] = cx * cy;
    m[15] = 1.;
    return m;
  }

  int Add(double sid, double sdd, double gDeg, double px, double py, double oDeg, double iDeg, double sx, double sy)
  {
    if (!views.empty() && ((sdd == 0.) != (views[0].sdd == 0.)))
      throw rtkhip::Error(RTKHIP_ERR_GEOMETRY, "cannot mix parallel and divergent projections in one geometry");
    const double d2r = std::atan(1.0) / 45.0;
    const double ga = d2r * gDeg, oa = d2r * oDeg, ia = d2r * iDeg;
    View         v;
    v.sid = sid;
    v.sdd = sdd;
    v.gantry = rtkhip::Wrap2Pi(ga);
    v.oop = rtkhip::Wrap2Pi(oa);
    v.ip = rtkhip::Wrap2Pi(ia);
    v.projX = px;
    v.projY = py;
    v.srcX = sx;
    v.srcY = sy;
    v.rot = Rotation(oa, ga, ia);
    // GetSourcePosition (.cxx:597-608): R^-1 = R^T for a rotation.
    const double s[3] = { sx, sy, sid };
    for (int r = 0; r < 3; r++)
      v.source[r] = v.rot[r] * s[0] + v.rot[4 + r] * s[1] + v.rot[8 + r] * s[2];
    // Source angle (.cxx:133-144): angle of the source projected on the x-z plane.
    const double n = std::sqrt(v.source[0] * v.source[0] + v.source[2] * v.source[2]);
    double       a = std::acos(v.source[2] / n);
    if (v.source[0] / n > 0.)
      a = 2. * kPi - a;
    v.sourceAngle = rtkhip::Wrap2Pi(a);
    views.push_back(v);
    return 0;
  }

  // Mag * SrcTrans * Rot (3x4, rows u, v, w in projection physical coordinates).
  std::array<double, 12> VolumePhysicalToProjection(size_t i) const
  {
    const View & v = views[i];
    // Magnification (.cxx:573-582) with sdd -> -sdd, sid -> -sid.
    const double           msdd = -v.sdd, msid = -v.sid;
    const double           m00 = (msdd == 0.) ? 1. : msdd, m22 = (msdd == 0.) ? 0. : 1., m23 = (msdd == 0.) ? 1. : msid;
    rtkhip::M4             st = rtkhip::Identity4();
    st[3] = -v.srcX;
    st[7] = -v.srcY;
    const rtkhip::M4       sr = rtkhip::Mul(st, v.rot);
    std::array<double, 12> out{};
    for (int c = 0; c < 4; c++)
    {
      out[c] = m00 * sr[c];
      out[4 + c] = m00 * sr[4 + c];
      out[8 + c] = m22 * sr[8 + c] + m23 * sr[12 + c];
    }
    return out;
  }

  // Full 3x4 projection matrix (.cxx:125-130).
  std::array<double, 12> Matrix(size_t i) const
  {
    const View & v = views[i];
    auto         p = VolumePhysicalToProjection(i);
    const double tx = v.srcX - v.projX, ty = v.srcY - v.projY;
    for (int c = 0; c < 4; c++)
    {
      p[c] += tx * p[8 + c];
      p[4 + c] += ty * p[8 + c];
    }
    return p;
  }

  // GetProjectionCoordinatesToFixedSystemMatrix (.cxx:611-637).
  rtkhip::M4 ProjectionToFixed(size_t i) const
  {
    const View & v = views[i];
    rtkhip::M4   d = rtkhip::Identity4();
    d[3] = v.projX;
    d[7] = v.projY;
    d[11] = (v.sdd == 0.) ? -v.sid : v.sid - v.sdd;
    d[10] = 0.;
    rtkhip::M4 rt{}; // R^T
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++)
        rt[4 * r + c] = v.rot[4 * c + r];
    rt[15] = 1.;
    return rtkhip::Mul(rt, d);
  }

  // ToUntiltedCoordinateAtIsocenter (.cxx:641-663).
  double Untilted(size_t i, double tilted) const
  {
    const View & v = views[i];
    const double sidu = std::sqrt(v.sid * v.sid + v.srcX * v.srcX);
    const double l = (tilted + v.projX - v.srcX) * v.sid / v.sdd + v.srcX;
    const double cosa = v.srcX / sidu;
    return l * std::abs(v.sid) / (sidu - l * cosa);
  }

  // GetAngularGapsWithNext over the gantry angles (.cxx:447-478).
  std::vector<double> GantryGapsWithNext() const
  {
    const size_t        n = views.size();
    std::vector<double> g(n, 0.);
    if (n == 1)
      g[0] = 2 * kPi;
    if (n < 2)
      return g;
    std::vector<std::pair<double, unsigned>> s(n);
    for (unsigned i = 0; i < n; i++)
      s[i] = { views[i].gantry, i };
    std::stable_sort(s.begin(), s.end(), [](const auto & a, const auto & b) { return a.first < b.first; });
    for (size_t k = 0; k + 1 < n; k++)
      g[s[k].second] = s[k + 1].first - s[k].first;
    g[s[n - 1].second] = s[0].first + 2 * kPi - s[n - 1].first;
    return g;
  }

  // GetAngularGaps over the source angles (.cxx:481-533).
  std::vector<double> AngularGaps() const
  {
    const size_t        n = views.size();
    std::vector<double> g(n, 0.);
    if (n == 1)
      g[0] = 2 * kPi;
    if (n < 2)
      return g;
    std::vector<std::pair<double, unsigned>> s(n);
    for (unsigned i = 0; i < n; i++)
      s[i] = { views[i].sourceAngle, i };
    std::stable_sort(s.begin(), s.end(), [](const auto & a, const auto & b) { return a.first < b.first; });
    g[s[0].second] = 0.5 * (s[1].first - s[n - 1].first + 2 * kPi);
    for (size_t k = 1; k + 1 < n; k++)
      g[s[k].second] = 0.5 * (s[k + 1].first - s[k - 1].first);
    g[s[n - 1].second] = 0.5 * (s[0].first + 2 * kPi - s[n - 2].first);
    if (views[0].sdd == 0.)
    {
      const double mx = *std::max_element(g.begin(), g.end());
      if (mx > kPi / 2)
        for (auto & x : g)
        {
          if (x > kPi / 2)
            x -= kPi / 2;
          x *= 2.;
        }
    }
    return g;
  }
};

// ------------------------------------------------------------------ context --
struct rtkhip_context
{
  rtkhip::DeviceBuffer<unsigned char> filtered;
  rtkhip::DeviceBuffer<float>         rowFiltered; // FP32 row-filtered slab before the 2-D kernel's column pass
  rtkhip::DeviceBuffer<float>         vB, vR, vP, vAP, pT;
  rtkhip::DeviceBuffer<unsigned short> pT16;   // 16-bit stack read by the CG backprojector
  rtkhip::DeviceBuffer<unsigned>       maxBits; // max |stack| for the FP16 scaling
  rtkhip::DeviceBuffer<double>        state;
  struct RampKey
  {
    int    xpad;
    double spacing;
    double opts[6];
    bool   operator<(const RampKey & o) const
    {
      if (xpad != o.xpad)
        return xpad < o.xpad;
      if (spacing != o.spacing)
        return spacing < o.spacing;
      return std::memcmp(opts, o.opts, sizeof(opts)) < 0;
    }
  };
  std::map<RampKey, std::vector<double>> rampKernels;
};

namespace rtkhip
{
namespace
{

void
CheckImage(const rtkhip_image * im, const char * what)
{
  RTKHIP_REQUIRE(im, RTKHIP_ERR_INVALID_ARGUMENT, std::string(what) + ": null image descriptor");
  for (int r = 0; r < 3; r++)
  {
    RTKHIP_REQUIRE(im->size[r] > 0, RTKHIP_ERR_INVALID_ARGUMENT, std::string(what) + ": sizes must be positive");
    RTKHIP_REQUIRE(im->spacing[r] != 0., RTKHIP_ERR_INVALID_ARGUMENT, std::string(what) + ": zero spacing");
  }
}

void
CheckGeometry(const rtkhip_geometry * g, const rtkhip_image & stack, bool flatOnly)
{
  RTKHIP_REQUIRE(g, RTKHIP_ERR_GEOMETRY, "Geometry has not been set.");
  const long last = long(stack.index[2]) + stack.size[2];
  RTKHIP_REQUIRE(stack.index[2] >= 0 && (long)g->views.size() >= last, RTKHIP_ERR_GEOMETRY,
                 "Mismatch between the number of projections and the geometry entries. Geometry has " +
                   std::to_string(g->views.size()) + " entries, which is less than the last index of the "
                   "projections stack, i.e., " + std::to_string(last) + ".");
  RTKHIP_REQUIRE(!flatOnly || g->radius == 0., RTKHIP_ERR_UNSUPPORTED,
                 "cylindrical detectors are not supported by this projector");
}

// Index-to-index BP matrices for the views of `stack` (buffer-relative indices),
// optionally FDK-normalised, plus the per-view border rule of the CPU filter.
void
BackprojectionMatrices(const rtkhip_geometry & g, const rtkhip_image & vol, const rtkhip_image & stack, bool fdk,
                       std::vector<double> & mats, std::vector<int> & border)
{
  const int np = stack.size[2];
  mats.assign(12 * size_t(np), 0.);
  border.assign(np, RTKHIP_BORDER_OPTIMIZED);
  const M4 volIdx = BufferIndexToPhysical(vol);
  // Rotation centre (physical origin) as a buffer-relative continuous index.
  double rc[3];
  for (int r = 0; r < 3; r++)
    rc[r] = (0. - vol.origin[r]) / vol.spacing[r] - vol.index[r];
  for (int p = 0; p < np; p++)
  {
    const size_t gi = size_t(stack.index[2]) + p;
    const auto   vp = g.VolumePhysicalToProjection(gi); // 3x4
    const auto & v = g.views[gi];
    // Projection physical -> buffer index: idx = (pp + (srcOff - projOff) - origin)/spacing - index.
    const double tx = v.srcX - v.projX, ty = v.srcY - v.projY;
    double       m[12];
    for (int c = 0; c < 4; c++)
    {
      double acc[3] = { 0., 0., 0. };
      for (int k = 0; k < 4; k++)
        for (int r = 0; r < 3; r++)
          acc[r] += vp[4 * r + k] * volIdx[4 * k + c];
      const double uw = acc[0] + tx * acc[2], vw = acc[1] + ty * acc[2];
      m[c] = (uw - (stack.origin[0] + stack.index[0] * stack.spacing[0]) * acc[2]) / stack.spacing[0];
      m[4 + c] = (vw - (stack.origin[1] + stack.index[1] * stack.spacing[1]) * acc[2]) / stack.spacing[1];
      m[8 + c] = acc[2];
    }
    if (fdk)
    {
      const double pf = m[11] + m[8] * rc[0] + m[9] * rc[1] + m[10] * rc[2];
      for (double & x : m)
        x /= pf;
    }
    // RTK chooses OptimizedBackprojectionX/Y or the generic ITK path on the
    // transposed matrix (include/rtkBackProjectionImageFilter.hxx:234-243).
    const bool xPath = std::abs(m[0]) < 1e-10 && std::abs(m[8]) < 1e-10;
    const bool yPath = std::abs(m[1]) < 1e-10 && std::abs(m[9]) < 1e-10;
    border[p] = (xPath || yPath) ? RTKHIP_BORDER_OPTIMIZED : RTKHIP_BORDER_ITK_LINEAR;
    std::copy(m, m + 12, mats.begin() + 12 * size_t(p));
  }
}

// Joseph rays: projection buffer index -> volume buffer index (rows 0..2, the
// view's stack index folded into the translation) and source in volume index.
void
JosephRays(const rtkhip_geometry & g, const rtkhip_image & stack, const rtkhip_image & vol,
           std::vector<double> & p2v, std::vector<double> & src)
{
  const int np = stack.size[2];
  p2v.assign(12 * size_t(np), 0.);
  src.assign(3 * size_t(np), 0.);
  const M4 volPPToIdx = InverseAffine(BufferIndexToPhysical(vol));
  M4       stackIdx{}; // absolute stack index -> physical
  for (int r = 0; r < 3; r++)
  {
    stackIdx[5 * r] = stack.spacing[r];
    stackIdx[4 * r + 3] = stack.origin[r];
  }
  stackIdx[15] = 1.;
  for (int p = 0; p < np; p++)
  {
    const size_t gi = size_t(stack.index[2]) + p;
    const M4     m = Mul(volPPToIdx, Mul(g.ProjectionToFixed(gi), stackIdx));
    for (int r = 0; r < 3; r++)
    {
      // Buffer index (i, j) -> absolute (i + index0, j + index1, gi).
      double *     o = p2v.data() + 12 * size_t(p) + 4 * r;
      const double * row = m.data() + 4 * r;
      o[0] = row[0];
      o[1] = row[1];
      o[2] = 0.;
      o[3] = row[3] + row[0] * stack.index[0] + row[1] * stack.index[1] + row[2] * double(gi);
      const double * s = g.views[gi].source;
      src[3 * size_t(p) + r] = volPPToIdx[4 * r] * s[0] + volPPToIdx[4 * r + 1] * s[1] +
                               volPPToIdx[4 * r + 2] * s[2] + volPPToIdx[4 * r + 3];
    }
  }
}

// Seven weighting values per view (include/rtkFDKWeightProjectionFilter.hxx:40-63).
std::vector<double>
WeightGeometry(const rtkhip_geometry & g, const rtkhip_image & stack)
{
  const auto          gaps = g.AngularGaps();
  const int           np = stack.size[2];
  std::vector<double> out(7 * size_t(np));
  for (int p = 0; p < np; p++)
  {
    const size_t gi = size_t(stack.index[2]) + p;
    const auto & v = g.views[gi];
    double       f = gaps[gi];
    if (v.sdd == 0.)
      f *= 0.5;
    else
    {
      f *= std::abs(v.sdd) / (2. * v.sid * v.sid);
      f *= std::sqrt(v.source[0] * v.source[0] + v.source[1] * v.source[1] + v.source[2] * v.source[2]);
    }
    const double tilt = Wrap2Pi(-v.gantry - v.sourceAngle);
    double *     o = out.data() + 7 * size_t(p);
    o[0] = v.sdd;
    o[1] = v.sid;
    o[2] = v.projX;
    o[3] = v.projY;
    o[4] = v.srcY;
    o[5] = f;
    o[6] = tilt;
  }
  return out;
}

// Forward DFT of any length in FP64, for the ramp kernel's spectrum (once per padded
// size).  Mixed-radix decimation in time: the smallest prime factor p of n splits the
// sequence into p interleaved subsequences whose transforms are combined with the
// twiddles exp(-2 pi i r k / n); a prime length is summed directly.
void
HostFFT(std::vector<std::complex<double>> & a)
{
  const int n = int(a.size());
  if (n <= 1)
    return;
  int p = 2;
  while (p * p <= n && n % p)
    p++;
  if (n % p)
    p = n;
  auto tw = [n](long e) {
    const double ang = -2 * kPi * double(e % n) / n;
    return std::complex<double>(std::cos(ang), std::sin(ang));
  };
  std::vector<std::complex<double>> out(n);
  if (p == n)
  {
    for (int k = 0; k < n; k++)
    {
      std::complex<double> acc = 0.;
      for (int r = 0; r < n; r++)
        acc += a[r] * tw(long(r) * k);
      out[k] = acc;
    }
    a.swap(out);
    return;
  }
  const int                                      m = n / p;
  std::vector<std::vector<std::complex<double>>> sub(p, std::vector<std::complex<double>>(m));
  for (int r = 0; r < p; r++)
  {
    for (int q = 0; q < m; q++)
      sub[r][q] = a[size_t(q) * p + r];
    HostFFT(sub[r]);
  }
  for (int k = 0; k < n; k++)
  {
    std::complex<double> acc = 0.;
    for (int r = 0; r < p; r++)
      acc += sub[r][k % m] * tw(long(r) * k);
    out[k] = acc;
  }
  a.swap(out);
}

// Ramp kernel spectrum (xpad/2+1 complex, interleaved re/im).
const std::vector<double> &
RampKernel(rtkhip_context & c, int xpad, double spacing, const rtkhip_ramp_options & o)
{
  rtkhip_context::RampKey key{ xpad, spacing, { o.hann_cut, o.cosine_cut, o.hamming, o.ramlak_cut, o.shepp_logan_cut, 0. } };
  auto it = c.rampKernels.find(key);
  if (it != c.rampKernels.end())
    return it->second;
  std::vector<std::complex<double>> k(xpad, 0.);
  k[0] = 1. / (4. * spacing);
  for (int ix = 1, jx = xpad - 1; ix < xpad / 2; ix += 2, jx -= 2)
  {
    const double t = ix * kPi;
    k[ix] = k[jx] = -1. / (t * t * spacing);
  }
  HostFFT(k);
  const int n = xpad / 2 + 1;
  auto      rnd = [](double x) { return int(std::floor(x + 0.5)); };
  int       i = 0;
  if (o.hann_cut > 0.)
  {
    const int ncut = rnd(n * std::min(1.0, o.hann_cut));
    for (; i < ncut; i++)
      k[i] *= 0.5 * (1 + std::cos(kPi * i / ncut));
  }
  else if (o.cosine_cut > 0.)
  {
    const int ncut = rnd(n * std::min(1.0, o.cosine_cut));
    for (; i < ncut; i++)
      k[i] *= std::cos(0.5 * kPi * i / ncut);
  }
  else if (o.hamming > 0.)
  {
    const int ncut = rnd(n * std::min(1.0, o.hamming));
    for (; i < ncut; i++)
      k[i] *= 0.54 + 0.46 * std::cos(kPi * i / ncut);
  }
  else if (o.ramlak_cut > 0.)
    i = rnd(n * std::min(1.0, o.ramlak_cut));
  else if (o.shepp_logan_cut > 0.)
  {
    const int ncut = rnd(n * std::min(1.0, o.shepp_logan_cut));
    for (i = 1; i < ncut; i++)
    {
      const double x = 0.5 * kPi * i / ncut;
      k[i] *= std::sin(x) / x;
    }
  }
  else
    i = n;
  for (; i < n; i++)
    k[i] = 0.;
  std::vector<double> out(2 * size_t(n));
  for (int q = 0; q < n; q++)
  {
    out[2 * q] = k[q].real();
    out[2 * q + 1] = k[q].imag();
  }
  return c.rampKernels.emplace(key, std::move(out)).first->second;
}

// GreatestPrimeFactor of include/rtkFFTProjectionsConvolutionImageFilter.hxx:329-340
// (2 for n <= 1).
int
GreatestPrimeFactor(int n)
{
  int v = 2;
  while (v <= n)
  {
    if (n % v == 0)
      n /= v;
    else
      v++;
  }
  return v;
}

// Padded sizes of include/rtkFFTProjectionsConvolutionImageFilter.hxx:280-304: the
// smallest size >= ZeroPadFactor * N whose greatest prime factor is at most the
// configured one, along u; along v the same with the y factor only for a 2-D kernel
// (HannCutFrequencyY), else Nv made smooth.  Unset (zero) options take RTK's defaults:
// zero-pad factors 2, and greatest prime factor 13, the value of RTK's CUDA convolution
// filter (include/rtkCudaFFTProjectionsConvolutionImageFilter.hxx:42) that these
// filters replace (RTK's CPU filter uses 2 unless its ITK FFT is FFTW).
void
PaddedSizes(int nu, int nv, const rtkhip_ramp_options & o, int & xpad, int & ypad)
{
  const int gpf = o.greatest_prime_factor == 0 ? 13 : o.greatest_prime_factor;
  RTKHIP_REQUIRE(gpf >= 2 && gpf <= 13, RTKHIP_ERR_UNSUPPORTED,
                 "ramp: greatest prime factor must lie in [2, 13] (the FFT radices available)");
  const int zpx = o.zero_pad_x <= 0 ? 2 : o.zero_pad_x, zpy = o.zero_pad_y <= 0 ? 2 : o.zero_pad_y;
  RTKHIP_REQUIRE(zpx <= 8 && zpy <= 8, RTKHIP_ERR_INVALID_ARGUMENT, "ramp: zero-pad factors above 8");
  xpad = zpx * nu;
  while (GreatestPrimeFactor(xpad) > gpf)
    xpad++;
  ypad = nv * (o.hann_cut_y > 0. ? zpy : 1);
  while (GreatestPrimeFactor(ypad) > gpf)
    ypad++;
}

// Column spectrum of the 2-D kernel of include/rtkFFTRampImageFilter.hxx:143-176: the
// reference replicates the ramp row over the padded height with a Hann window on the
// first ncut rows and their mirror images, so the 2-D kernel is the outer product of the
// row kernel and this window (ypad complex values, interleaved; imaginary parts 0).
std::vector<double>
HannYSpectrum(int ypad, double cut)
{
  const int           ncut = int(std::floor((ypad / 2 + 1) * std::min(1.0, cut) + 0.5));
  std::vector<double> w(2 * size_t(ypad), 0.);
  for (int j = 0; j < ncut; j++)
    w[2 * j] = 0.5 * (1 + std::cos(kPi * j / ncut));
  for (int j = 1; j < ncut; j++)
    w[2 * (ypad - j)] = 0.5 * (1 + std::cos(kPi * j / ncut));
  return w;
}

__global__ void
multiply_kernel(size_t n, const float * a, const float * b, float * out)
{
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    out[i] = a[i] * b[i];
}

void
Multiply(size_t n, const float * a, const float * b, float * out, hipStream_t s)
{
  const unsigned g = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(multiply_kernel, dim3(g ? g : 1), dim3(256), 0, s, n, a, b, out);
  RTKHIP_CHECK_LAUNCH();
}

// Power-of-two factor that brings max|stack| (as FP32 bits) into (2^13, 2^14]: FP16
// holds values up to 65504, and the CG operator's stacks W FP(P) grow with the
// unnormalised search direction P.  Zero maps to 1; Inf/NaN propagate.
__device__ __forceinline__ int
Fp16ScaleExponent(unsigned maxBits)
{
  const float m = __uint_as_float(maxBits);
  if (m == 0.f)
    return 0;
  int e;
  frexpf(m, &e); // m < 2^e
  return 14 - max(e, -100);
}

// Clears the max word with a kernel rather than a memset node, so a caller may
// capture the sequence in a graph.
__global__ void clear_word_kernel(unsigned * w)
{
  *w = 0u;
}

// *maxBits = max |a * w| over the stack (w NULL = ones), as FP32 bits
// (non-negative floats order like their bit patterns).
__global__ __launch_bounds__(256) void absmax_kernel(size_t n, const float * a, const float * w, unsigned * maxBits)
{
  __shared__ unsigned red[256];
  unsigned            m = 0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    m = max(m, __float_as_uint(fabsf(w ? a[i] * w[i] : a[i])));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1)
  {
    if (int(threadIdx.x) < o)
      red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    atomicMax(maxBits, red[0]);
}

// out = TO(a * w * 2^k) (w NULL = ones; k from maxBits, 0 when maxBits is NULL):
// the weighted stack in its 16-bit storage type.
template <typename TO>
__global__ __launch_bounds__(256) void weight_store_kernel(size_t n, const float * a, const float * w,
                                                            const unsigned * maxBits, TO * out)
{
  const float scale = maxBits ? ldexpf(1.f, Fp16ScaleExponent(*maxBits)) : 1.f;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
  {
    const float v = (w ? a[i] * w[i] : a[i]) * scale;
    if constexpr (std::is_same<TO, __half>::value)
      out[i] = __float2half(v);
    else
      out[i] = __float2bfloat16(v);
  }
}

// vol *= 2^-k: undoes the stack scaling after the (linear) backprojection.
__global__ __launch_bounds__(256) void unscale_kernel(size_t n, float * vol, const unsigned * maxBits)
{
  const float inv = ldexpf(1.f, -Fp16ScaleExponent(*maxBits));
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    vol[i] *= inv;
}

// Stores W a in the 16-bit stack; for FP16 also computes the scaling (maxBits must
// then be a device word) that Unscale16 removes after the backprojection.
void
WeightAndStore(size_t n, const float * a, const float * w, void * out, int dtype, unsigned * maxBits, hipStream_t s)
{
  const unsigned g = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  if (dtype == RTKHIP_F16)
  {
    hipLaunchKernelGGL(clear_word_kernel, dim3(1), dim3(1), 0, s, maxBits);
    RTKHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(absmax_kernel, dim3(g ? g : 1), dim3(256), 0, s, n, a, w, maxBits);
    RTKHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(weight_store_kernel<__half>, dim3(g ? g : 1), dim3(256), 0, s, n, a, w,
                       static_cast<const unsigned *>(maxBits), static_cast<__half *>(out));
  }
  else
    hipLaunchKernelGGL(weight_store_kernel<__hip_bfloat16>, dim3(g ? g : 1), dim3(256), 0, s, n, a, w,
                       static_cast<const unsigned *>(nullptr), static_cast<__hip_bfloat16 *>(out));
  RTKHIP_CHECK_LAUNCH();
}

void
Unscale16(size_t n, float * vol, const unsigned * maxBits, hipStream_t s)
{
  const unsigned g = unsigned(std::min<size_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(unscale_kernel, dim3(g ? g : 1), dim3(256), 0, s, n, vol, maxBits);
  RTKHIP_CHECK_LAUNCH();
}

size_t
DtypeSize(int dtype)
{
  return dtype == RTKHIP_F32 ? 4 : 2;
}

void
Throw(int rc)
{
  if (rc != RTKHIP_OK)
    throw Error(rc, g_lastError);
}

} // namespace
} // namespace rtkhip

// ========================================================================= ABI
using namespace rtkhip;

extern "C" const char *
rtkhip_last_error(void)
{
  return g_lastError.c_str();
}

extern "C" const char *
rtkhip_version(void)
{
  return "rtkhip 0.1.0 (gfx950)";
}

extern "C" int
rtkhip_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess)
  {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

extern "C" rtkhip_geometry *
rtkhip_geometry_create(void)
{
  try
  {
    return new rtkhip_geometry();
  }
  catch (...)
  {
    SetLastError("host allocation failed");
    return nullptr;
  }
}

extern "C" void
rtkhip_geometry_destroy(rtkhip_geometry * g)
{
  delete g;
}

extern "C" int
rtkhip_geometry_add_projection(rtkhip_geometry * g, double sid, double sdd, double gantry_angle, double proj_offset_x,
                               double proj_offset_y, double out_of_plane_angle, double in_plane_angle,
                               double source_offset_x, double source_offset_y)
{
  return Guard([&] {
    RTKHIP_REQUIRE(g, RTKHIP_ERR_INVALID_ARGUMENT, "geometry_add_projection: null geometry");
    g->Add(sid, sdd, gantry_angle, proj_offset_x, proj_offset_y, out_of_plane_angle, in_plane_angle, source_offset_x,
           source_offset_y);
  });
}

extern "C" int
rtkhip_geometry_set_radius_cylindrical_detector(rtkhip_geometry * g, double radius)
{
  return Guard([&] {
    RTKHIP_REQUIRE(g, RTKHIP_ERR_INVALID_ARGUMENT, "null geometry");
    g->radius = radius;
  });
}

extern "C" int
rtkhip_geometry_size(const rtkhip_geometry * g)
{
  return g ? int(g->views.size()) : 0;
}

extern "C" int
rtkhip_geometry_get(const rtkhip_geometry * g, int i, double matrix[12], double source[3], double * source_angle)
{
  return Guard([&] {
    RTKHIP_REQUIRE(g && i >= 0 && i < int(g->views.size()), RTKHIP_ERR_INVALID_ARGUMENT, "geometry_get: bad index");
    if (matrix)
    {
      const auto m = g->Matrix(i);
      std::copy(m.begin(), m.end(), matrix);
    }
    if (source)
      for (int r = 0; r < 3; r++)
        source[r] = g->views[i].source[r];
    if (source_angle)
      *source_angle = g->views[i].sourceAngle;
  });
}

extern "C" rtkhip_context *
rtkhip_context_create(void)
{
  try
  {
    return new rtkhip_context();
  }
  catch (...)
  {
    SetLastError("host allocation failed");
    return nullptr;
  }
}

extern "C" void
rtkhip_context_destroy(rtkhip_context * c)
{
  delete c;
}

extern "C" int
rtkhip_back_project(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * vol, const float * dev_vol_in,
                    float * dev_vol_out, const rtkhip_image * stack, const void * dev_proj, int proj_dtype, int fdk,
                    void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(c, RTKHIP_ERR_INVALID_ARGUMENT, "back_project: null context");
    CheckImage(vol, "back_project volume");
    CheckImage(stack, "back_project projections");
    CheckGeometry(g, *stack, true);
    std::vector<double> mats;
    std::vector<int>    border;
    BackprojectionMatrices(*g, *vol, *stack, fdk != 0, mats, border);
    const int pd[3] = { stack->size[0], stack->size[1], stack->size[2] };
    const int vd[3] = { vol->size[0], vol->size[1], vol->size[2] };
    Throw(fdk ? rtkhip_fdk_backproject(pd, vd, mats.data(), border.data(), dev_vol_in, dev_vol_out, dev_proj,
                                       proj_dtype, stream)
              : rtkhip_voxel_backproject(pd, vd, mats.data(), border.data(), dev_vol_in, dev_vol_out, dev_proj,
                                         proj_dtype, stream));
    RTKHIP_CHECK(hipStreamSynchronize(Stream(stream)));
  });
}

extern "C" int
rtkhip_forward_project(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * stack,
                       const float * dev_proj_in, float * dev_proj_out, const rtkhip_image * vol,
                       const float * dev_vol, long long * samples, void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(c, RTKHIP_ERR_INVALID_ARGUMENT, "forward_project: null context");
    CheckImage(vol, "forward_project volume");
    CheckImage(stack, "forward_project projections");
    CheckGeometry(g, *stack, true);
    RTKHIP_REQUIRE(g->views[0].sdd != 0., RTKHIP_ERR_UNSUPPORTED,
                   "parallel geometry is not supported by the HIP Joseph projector");
    std::vector<double> p2v, src;
    JosephRays(*g, *stack, *vol, p2v, src);
    const int pd[3] = { stack->size[0], stack->size[1], stack->size[2] };
    const int vd[3] = { vol->size[0], vol->size[1], vol->size[2] };
    Throw(rtkhip_joseph_forward_project(pd, vd, p2v.data(), src.data(), vol->spacing, 0., 1., dev_proj_in,
                                        dev_proj_out, dev_vol, samples, stream));
    RTKHIP_CHECK(hipStreamSynchronize(Stream(stream)));
  });
}

extern "C" int
rtkhip_fft_ramp_filter(rtkhip_context * c, const rtkhip_image * stack, const float * dev_in, float * dev_out,
                       const rtkhip_ramp_options * ramp, void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(c && dev_in && dev_out, RTKHIP_ERR_INVALID_ARGUMENT, "fft_ramp_filter: null argument");
    CheckImage(stack, "fft_ramp_filter projections");
    rtkhip_ramp_options o{};
    if (ramp)
      o = *ramp;
    RequireDevice();
    const int nu = stack->size[0], nv = stack->size[1], np = stack->size[2];
    int       xpad = 0, ypad = 0;
    PaddedSizes(nu, nv, o, xpad, ypad);
    const long zeroext = (long(xpad) - nu) / 2;
    // truncation extent floor(TruncationCorrection * Nu), clipped by the zero padding
    // (include/rtkFFTProjectionsConvolutionImageFilter.hxx:210-211,345-355)
    const int           nextFull = int(std::floor(o.truncation_correction * nu));
    const int           next = int(std::min<long>(zeroext, nextFull));
    std::vector<double> mirror(nextFull + 1);
    for (int i = 0; i <= nextFull; i++)
      mirror[i] = std::pow(std::sin((nextFull - i) * kPi / (2 * nextFull - 2)), 0.75);
    const auto & kernel = RampKernel(*c, xpad, stack->spacing[0], o);
    const int    pd[3] = { nu, nv, np };
    Throw(rtkhip_ramp_filter(pd, xpad, kernel.data(), next, mirror.data(), nullptr, nullptr, nullptr, dev_in, dev_out,
                             RTKHIP_F32, stream));
    if (o.hann_cut_y > 0.)
    {
      const std::vector<double> wy = HannYSpectrum(ypad, o.hann_cut_y);
      Throw(rtkhip_column_filter(pd, ypad, wy.data(), dev_out, dev_out, RTKHIP_F32, stream));
    }
    RTKHIP_CHECK(hipStreamSynchronize(Stream(stream)));
  });
}

extern "C" int
rtkhip_fdk_reconstruct(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * vol,
                       const float * dev_vol_in, float * dev_vol_out, const rtkhip_image * stack, const float * dev_proj,
                       const rtkhip_ramp_options * ramp, int subset_size, int stack_dtype, double * bp_ms,
                       void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(c && dev_vol_out && dev_proj, RTKHIP_ERR_INVALID_ARGUMENT, "fdk_reconstruct: null argument");
    CheckImage(vol, "fdk volume");
    CheckImage(stack, "fdk projections");
    CheckGeometry(g, *stack, true);
    RTKHIP_REQUIRE(stack_dtype == RTKHIP_F32 || stack_dtype == RTKHIP_F16 || stack_dtype == RTKHIP_BF16,
                   RTKHIP_ERR_INVALID_ARGUMENT, "fdk_reconstruct: unknown stack dtype");
    rtkhip_ramp_options o{};
    if (ramp)
      o = *ramp;
    RequireDevice();
    hipStream_t s = Stream(stream);
    const int   nu = stack->size[0], nv = stack->size[1], np = stack->size[2];
    int         xpad = 0, ypad = 0;
    PaddedSizes(nu, nv, o, xpad, ypad);
    const bool                twoD = o.hann_cut_y > 0.;
    const std::vector<double> wy = twoD ? HannYSpectrum(ypad, o.hann_cut_y) : std::vector<double>();
    const long zeroext = (long(xpad) - nu) / 2;
    const int  nextFull = int(std::floor(o.truncation_correction * nu));
    const int  next = int(std::min<long>(zeroext, nextFull));
    std::vector<double> mirror(nextFull + 1);
    for (int i = 0; i <= nextFull; i++)
      mirror[i] = std::pow(std::sin((nextFull - i) * kPi / (2 * nextFull - 2)), 0.75);
    const auto & kernel = RampKernel(*c, xpad, stack->spacing[0], o);

    std::vector<double> mats, w7all;
    std::vector<int>    border;
    BackprojectionMatrices(*g, *vol, *stack, true, mats, border);
    w7all = WeightGeometry(*g, *stack);
    const double origin[2] = { stack->origin[0] + stack->index[0] * stack->spacing[0],
                               stack->origin[1] + stack->index[1] * stack->spacing[1] };
    const int    subset = (subset_size <= 0 || subset_size > np) ? np : subset_size;
    const size_t ppix = size_t(nu) * nv;
    const size_t vox = size_t(vol->size[0]) * vol->size[1] * vol->size[2];
    if (np == 0)
    {
      if (dev_vol_in != dev_vol_out)
        RTKHIP_CHECK(hipMemcpyAsync(dev_vol_out, dev_vol_in, vox * sizeof(float), hipMemcpyDeviceToDevice, s));
      RTKHIP_CHECK(hipStreamSynchronize(s));
      if (bp_ms)
        *bp_ms = 0.;
      return;
    }
    c->filtered.Resize(ppix * subset * DtypeSize(stack_dtype));
    if (twoD && stack_dtype != RTKHIP_F32)
      c->rowFiltered.Resize(ppix * subset);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (bp_ms)
    {
      RTKHIP_CHECK(hipEventCreate(&e0));
      RTKHIP_CHECK(hipEventCreate(&e1));
    }
    double       bpTotal = 0.;
    const float * vin = dev_vol_in;
    for (int first = 0; first < np; first += subset)
    {
      const int cnt = std::min(subset, np - first);
      const int pd[3] = { nu, nv, cnt };
      const int vd[3] = { vol->size[0], vol->size[1], vol->size[2] };
      if (twoD)
      {
        // rows into FP32, then the column window into the stack type
        float * rows = stack_dtype == RTKHIP_F32 ? reinterpret_cast<float *>(c->filtered.get()) : c->rowFiltered.get();
        Throw(rtkhip_ramp_filter(pd, xpad, kernel.data(), next, mirror.data(), w7all.data() + 7 * size_t(first),
                                 origin, stack->spacing, dev_proj + ppix * first, rows, RTKHIP_F32, stream));
        Throw(rtkhip_column_filter(pd, ypad, wy.data(), rows, c->filtered.get(), stack_dtype, stream));
      }
      else
        Throw(rtkhip_ramp_filter(pd, xpad, kernel.data(), next, mirror.data(), w7all.data() + 7 * size_t(first),
                                 origin, stack->spacing, dev_proj + ppix * first, c->filtered.get(), stack_dtype,
                                 stream));
      if (bp_ms)
        RTKHIP_CHECK(hipEventRecord(e0, s));
      Throw(rtkhip_fdk_backproject(pd, vd, mats.data() + 12 * size_t(first), border.data() + first, vin, dev_vol_out,
                                   c->filtered.get(), stack_dtype, stream));
      if (bp_ms)
      {
        RTKHIP_CHECK(hipEventRecord(e1, s));
        RTKHIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        RTKHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        bpTotal += ms;
      }
      vin = dev_vol_out;
    }
    RTKHIP_CHECK(hipStreamSynchronize(s));
    if (bp_ms)
    {
      *bp_ms = bpTotal;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
  });
}

extern "C" int
rtkhip_joseph_back_project(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * vol,
                           const float * dev_vol_in, float * dev_vol_out, const rtkhip_image * stack,
                           const void * dev_proj, int proj_dtype, void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(c, RTKHIP_ERR_INVALID_ARGUMENT, "joseph_back_project: null context");
    CheckImage(vol, "joseph_back_project volume");
    CheckImage(stack, "joseph_back_project projections");
    CheckGeometry(g, *stack, true);
    RTKHIP_REQUIRE(g->views[0].sdd != 0., RTKHIP_ERR_UNSUPPORTED,
                   "parallel geometry is not supported by the HIP Joseph projector");
    std::vector<double> p2v, src;
    JosephRays(*g, *stack, *vol, p2v, src);
    const int pd[3] = { stack->size[0], stack->size[1], stack->size[2] };
    const int vd[3] = { vol->size[0], vol->size[1], vol->size[2] };
    Throw(rtkhip_joseph_backproject(pd, vd, p2v.data(), src.data(), vol->spacing, 0., 1., dev_vol_in, dev_vol_out,
                                    dev_proj, proj_dtype, stream));
    RTKHIP_CHECK(hipStreamSynchronize(Stream(stream)));
  });
}

extern "C" int
rtkhip_cg_reconstruct(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * vol, float * dev_x,
                      const rtkhip_image * stack, const float * dev_proj, const float * dev_weights, int niter,
                      double rel_tol, int backprojector, int stack_dtype, double * resid, int * iterations,
                      void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(c && dev_x && dev_proj && niter >= 0, RTKHIP_ERR_INVALID_ARGUMENT, "cg_reconstruct: bad argument");
    RTKHIP_REQUIRE(backprojector == RTKHIP_BP_VOXELBASED || backprojector == RTKHIP_BP_JOSEPH,
                   RTKHIP_ERR_INVALID_ARGUMENT, "cg_reconstruct: unknown backprojector");
    RTKHIP_REQUIRE(stack_dtype == RTKHIP_F32 || stack_dtype == RTKHIP_F16 || stack_dtype == RTKHIP_BF16,
                   RTKHIP_ERR_INVALID_ARGUMENT, "cg_reconstruct: unknown stack dtype");
    CheckImage(vol, "cg volume");
    CheckImage(stack, "cg projections");
    CheckGeometry(g, *stack, true);
    RTKHIP_REQUIRE(g->views[0].sdd != 0., RTKHIP_ERR_UNSUPPORTED,
                   "parallel geometry is not supported by the HIP Joseph projector");
    RequireDevice();
    hipStream_t s = Stream(stream);
    const size_t nvox = size_t(vol->size[0]) * vol->size[1] * vol->size[2];
    const size_t npix = size_t(stack->size[0]) * stack->size[1] * stack->size[2];
    c->vB.Resize(nvox);
    c->vR.Resize(nvox);
    c->vP.Resize(nvox);
    c->vAP.Resize(nvox);
    c->pT.Resize(npix);
    if (stack_dtype != RTKHIP_F32)
    {
      c->pT16.Resize(npix);
      c->maxBits.Resize(1);
    }
    c->state.Resize(8);
    std::vector<double> mats, p2v, src;
    std::vector<int>    border;
    BackprojectionMatrices(*g, *vol, *stack, false, mats, border);
    JosephRays(*g, *stack, *vol, p2v, src);
    const int pd[3] = { stack->size[0], stack->size[1], stack->size[2] };
    const int vd[3] = { vol->size[0], vol->size[1], vol->size[2] };
    float *   T = c->pT.get();
    // BP of W q, with q rounded to the storage type first when the stack is 16-bit
    auto bpw = [&](const float * q, float * out) {
      const void * proj = q;
      if (stack_dtype != RTKHIP_F32)
      {
        WeightAndStore(npix, q, dev_weights, c->pT16.get(), stack_dtype, c->maxBits.get(), s);
        proj = c->pT16.get();
      }
      else if (dev_weights)
      {
        Multiply(npix, q, dev_weights, T, s);
        proj = T;
      }
      if (backprojector == RTKHIP_BP_JOSEPH)
        Throw(rtkhip_joseph_backproject(pd, vd, p2v.data(), src.data(), vol->spacing, 0., 1., nullptr, out, proj,
                                        stack_dtype, stream));
      else
        Throw(rtkhip_voxel_backproject(pd, vd, mats.data(), border.data(), nullptr, out, proj, stack_dtype, stream));
      if (stack_dtype == RTKHIP_F16)
        Unscale16(nvox, out, c->maxBits.get(), s);
    };
    // A x = BP(0, W FP(0, x))   (ReconstructionConjugateGradientOperator)
    auto applyA = [&](const float * x, float * out) {
      Throw(rtkhip_joseph_forward_project(pd, vd, p2v.data(), src.data(), vol->spacing, 0., 1., nullptr, T, x, nullptr,
                                          stream));
      bpw(T, out);
    };
    // B = BP(0, W p)
    bpw(dev_proj, c->vB.get());
    double * st = c->state.get();
    Throw(rtkhip_dot(nvox, c->vB.get(), c->vB.get(), st + 4, stream));
    // R = B - A X0.  A cold start (X0 = 0, so that x.x = 0 exactly) skips the
    // projector pair: A X0 is X0 itself there
    Throw(rtkhip_dot(nvox, dev_x, dev_x, st + 5, stream));
    double h[8];
    RTKHIP_CHECK(hipMemcpyAsync(h + 5, st + 5, sizeof(double), hipMemcpyDeviceToHost, s));
    RTKHIP_CHECK(hipStreamSynchronize(s));
    const float * ax = dev_x;
    if (h[5] != 0.)
    {
      applyA(dev_x, c->vAP.get());
      ax = c->vAP.get();
    }
    Throw(rtkhip_cg_init(nvox, c->vB.get(), ax, c->vR.get(), c->vP.get(), st, stream));
    RTKHIP_CHECK(hipMemcpyAsync(h, st, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
    RTKHIP_CHECK(hipStreamSynchronize(s));
    const double bnorm = std::sqrt(h[4]);
    auto         rel = [&](double rr) { return bnorm > 0. ? std::sqrt(rr) / bnorm : std::sqrt(rr); };
    if (resid)
      resid[0] = rel(h[0]);
    int it = 0;
    bool done = rel_tol > 0. && rel(h[0]) <= rel_tol;
    for (; it < niter && !done; it++)
    {
      applyA(c->vP.get(), c->vAP.get());
      Throw(rtkhip_cg_step(nvox, dev_x, c->vR.get(), c->vP.get(), c->vAP.get(), st, stream));
      RTKHIP_CHECK(hipMemcpyAsync(h, st, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
      RTKHIP_CHECK(hipStreamSynchronize(s));
      const double r = rel(h[0]);
      if (resid)
        resid[it + 1] = r;
      if (rel_tol > 0. && r <= rel_tol)
        done = true;
    }
    if (iterations)
      *iterations = it;
    RTKHIP_CHECK(hipStreamSynchronize(s));
  });
}

extern "C" int
rtkhip_project_quadrics(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * stack,
                        const float * dev_in, float * dev_out, int nshapes, const double * shapes, void * stream)
{
  return Guard([&] {
    RTKHIP_REQUIRE(c && dev_out && nshapes >= 0 && (nshapes == 0 || shapes), RTKHIP_ERR_INVALID_ARGUMENT,
                   "project_quadrics: bad argument");
    CheckImage(stack, "project_quadrics projections");
    CheckGeometry(g, *stack, true);
    RTKHIP_REQUIRE(g->views[0].sdd != 0., RTKHIP_ERR_UNSUPPORTED, "parallel geometry is not supported");
    RequireDevice();
    const int           np = stack->size[2];
    std::vector<double> rays(16 * size_t(np));
    M4                  stackIdx{};
    for (int r = 0; r < 3; r++)
    {
      stackIdx[5 * r] = stack->spacing[r];
      stackIdx[4 * r + 3] = stack->origin[r];
    }
    stackIdx[15] = 1.;
    for (int p = 0; p < np; p++)
    {
      const size_t gi = size_t(stack->index[2]) + p;
      const M4     m = Mul(g->ProjectionToFixed(gi), stackIdx);
      double *     o = rays.data() + 16 * size_t(p);
      for (int r = 0; r < 3; r++)
      {
        o[4 * r + 0] = m[4 * r + 0];
        o[4 * r + 1] = m[4 * r + 1];
        o[4 * r + 2] = 0.;
        o[4 * r + 3] = m[4 * r + 3] + m[4 * r] * stack->index[0] + m[4 * r + 1] * stack->index[1] +
                       m[4 * r + 2] * double(gi);
        o[12 + r] = g->views[gi].source[r];
      }
      o[15] = 1.;
    }
    if (np > 0)
      ProjectQuadricsLaunch(stack->size[0], stack->size[1], np, rays.data(), nshapes, shapes, dev_in, dev_out,
                            Stream(stream));
  });
}

// ------------------------------------------------- short-scan / displaced detector --
namespace rtkhip
{
namespace
{

struct ParkerSetupResult
{
  bool   isShortScan = false;
  double delta = 0., firstAngle = 0.;
  int    warnView = -1;
};

// include/rtkParkerShortScanImageFilter.hxx:41-110.  Corners are those of the
// stack's buffered region (= its largest possible region here).
ParkerSetupResult
ParkerSetup(const rtkhip_geometry & g, const rtkhip_image & stack, double threshold)
{
  ParkerSetupResult         r;
  const std::vector<double> gaps = g.GantryGapsWithNext();
  const int                 n = int(gaps.size());
  if (n == 0)
    return r;
  int maxPos = 0;
  for (int i = 1; i < n; i++)
    if (gaps[i] > gaps[maxPos])
      maxPos = i;
  if (g.views[0].sdd == 0. || gaps[maxPos] < threshold)
    return r;
  r.isShortScan = true;
  // GetUniqueSortedAngles (.cxx:434-444): the angle after the largest gap, cyclically
  std::vector<double> sorted(n);
  for (int i = 0; i < n; i++)
    sorted[i] = g.views[i].gantry;
  std::sort(sorted.begin(), sorted.end());
  sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
  const double lastA = g.views[maxPos].gantry;
  auto         it = std::upper_bound(sorted.begin(), sorted.end(), lastA);
  r.firstAngle = (it == sorted.end()) ? sorted.front() : *it;
  double last = lastA;
  if (last < r.firstAngle)
    last += 2 * kPi;
  r.delta = 0.5 * (last - r.firstAngle - kPi);
  r.delta = r.delta - 2 * kPi * std::floor(r.delta / (2 * kPi));
  const double c1 = stack.origin[0] + stack.spacing[0] * stack.index[0];
  const double c2 = c1 + stack.spacing[0] * (stack.size[0] - 1);
  for (int k = 0; k < stack.size[2]; k++)
  {
    const auto & v = g.views[k];
    const double invsid = 1. / std::sqrt(v.sid * v.sid + v.srcX * v.srcX);
    const double hw = std::min(std::abs(g.Untilted(k, c1)), std::abs(g.Untilted(k, c2)));
    if (r.delta < std::atan(hw * invsid))
    {
      r.warnView = k;
      break;
    }
  }
  return r;
}

struct DisplacedSetupResult
{
  rtkhip_image out;
  int          offset = 0; // output column of the input's column 0
  double       inferior = 0., superior = 0.;
  bool         weighting = false;
};

// include/rtkDisplacedDetectorImageFilter.hxx:81-173 (GenerateOutputInformation).
DisplacedSetupResult
DisplacedSetup(const rtkhip_geometry & g, const rtkhip_image & in, bool disable, bool pad, bool offsetsSet,
               double minOffset, double maxOffset)
{
  DisplacedSetupResult r;
  r.out = in;
  if (disable)
    return r;
  RTKHIP_REQUIRE(g.radius == 0., RTKHIP_ERR_UNSUPPORTED,
                 "Displaced detector cannot handle cylindrical detector. Consider disabling it by setting "
                 "m_Disable=true or using the nodisplaced flag of the application you are running");
  RTKHIP_REQUIRE(!g.views.empty() && g.views.front().sdd != 0., RTKHIP_ERR_UNSUPPORTED,
                 "Displaced detector cannot handle parallel geometry. Consider disabling it by setting "
                 "m_Disable=true or using the nodisplaced flag of the application you are running");
  double inf = in.origin[0] + in.spacing[0] * in.index[0], sup = inf;
  if (in.spacing[0] < 0.)
    inf += in.spacing[0] * (in.size[0] - 1);
  else
    sup += in.spacing[0] * (in.size[0] - 1);
  if (!offsetsSet)
  {
    double minSup = std::numeric_limits<double>::max(), maxInf = -std::numeric_limits<double>::max();
    for (size_t i = 0; i < g.views.size(); i++)
    {
      maxInf = std::max(maxInf, g.Untilted(i, inf));
      minSup = std::min(minSup, g.Untilted(i, sup));
    }
    inf = maxInf;
    sup = minSup;
  }
  else
  {
    inf += maxOffset;
    sup += minOffset;
  }
  r.inferior = inf;
  r.superior = sup;
  RTKHIP_REQUIRE(!(inf > 0. || sup < 0.), RTKHIP_ERR_UNSUPPORTED,
                 "Cannot account for detector displacement larger than 50% of panel size. Corner inf=" +
                   std::to_string(inf) + " and corner sup=" + std::to_string(sup));
  r.weighting = !(std::abs(inf + sup) < 0.1 * std::abs(sup - inf));
  if (!r.weighting || !pad)
    return r;
  r.out.size[0] = 2 * in.size[0];
  if (sup + inf > 0.)
  {
    r.out.index[0] = in.index[0] - in.size[0];
    r.offset = in.size[0];
  }
  return r;
}

// Copies a stack of n floats unless in == out.
void
CopyStack(const float * in, float * out, size_t n, hipStream_t s)
{
  if (in != out && n)
    RTKHIP_CHECK(hipMemcpyAsync(out, in, n * sizeof(float), hipMemcpyDeviceToDevice, s));
}

} // namespace
} // namespace rtkhip

extern "C" int
rtkhip_parker_short_scan(const rtkhip_geometry * g, const rtkhip_image * stack, const float * dev_in, float * dev_out,
                         double threshold, int * is_short_scan, int * warn_view, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(dev_in && dev_out, RTKHIP_ERR_INVALID_ARGUMENT, "parker_short_scan: null buffer");
    CheckImage(stack, "parker_short_scan projections");
    CheckGeometry(g, *stack, false);
    RequireDevice();
    const ParkerSetupResult r = ParkerSetup(*g, *stack, threshold);
    if (is_short_scan)
      *is_short_scan = r.isShortScan;
    if (warn_view)
      *warn_view = r.warnView;
    hipStream_t  s = Stream(stream);
    const size_t n = size_t(stack->size[0]) * stack->size[1] * stack->size[2];
    if (!r.isShortScan)
      CopyStack(dev_in, dev_out, n, s);
    else
    {
      std::vector<double> g5(5 * size_t(stack->size[2]));
      for (int p = 0; p < stack->size[2]; p++)
      {
        const auto & v = g->views[size_t(stack->index[2]) + p];
        const double row[5] = { v.sdd, v.srcX, v.projX, v.sid, v.gantry };
        std::copy(row, row + 5, g5.begin() + 5 * size_t(p));
      }
      const int rc = rtkhip_parker_weight(stack->size, g5.data(), r.delta, r.firstAngle,
                                          stack->origin[0] + stack->spacing[0] * stack->index[0], stack->spacing[0],
                                          dev_in, dev_out, stream);
      RTKHIP_REQUIRE(rc == RTKHIP_OK, rc, rtkhip_last_error());
    }
    RTKHIP_CHECK(hipStreamSynchronize(s));
  });
}

extern "C" int
rtkhip_displaced_detector_region(const rtkhip_geometry * g, const rtkhip_image * in, int disable,
                                 int pad_on_truncated_side, int offsets_set, double min_offset, double max_offset,
                                 rtkhip_image * out, double corners[2], int * weighting)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(out, RTKHIP_ERR_INVALID_ARGUMENT, "displaced_detector_region: null output descriptor");
    CheckImage(in, "displaced_detector projections");
    RTKHIP_REQUIRE(g, RTKHIP_ERR_GEOMETRY, "Geometry has not been set.");
    const DisplacedSetupResult r =
      DisplacedSetup(*g, *in, disable != 0, pad_on_truncated_side != 0, offsets_set != 0, min_offset, max_offset);
    *out = r.out;
    if (corners)
    {
      corners[0] = r.inferior;
      corners[1] = r.superior;
    }
    if (weighting)
      *weighting = r.weighting;
  });
}

extern "C" int
rtkhip_displaced_detector(const rtkhip_geometry * g, const rtkhip_image * in, const float * dev_in,
                          const rtkhip_image * out, float * dev_out, int disable, int pad_on_truncated_side,
                          int offsets_set, double min_offset, double max_offset, void * stream)
{
  using namespace rtkhip;
  return Guard([&] {
    RTKHIP_REQUIRE(dev_in && dev_out && out, RTKHIP_ERR_INVALID_ARGUMENT, "displaced_detector: null argument");
    CheckImage(in, "displaced_detector projections");
    CheckGeometry(g, *in, false);
    RequireDevice();
    const DisplacedSetupResult r =
      DisplacedSetup(*g, *in, disable != 0, pad_on_truncated_side != 0, offsets_set != 0, min_offset, max_offset);
    for (int i = 0; i < 3; i++)
      RTKHIP_REQUIRE(out->size[i] == r.out.size[i] && out->index[i] == r.out.index[i], RTKHIP_ERR_INVALID_ARGUMENT,
                     "displaced_detector: output region differs from rtkhip_displaced_detector_region's");
    hipStream_t s = Stream(stream);
    if (!r.weighting)
      CopyStack(dev_in, dev_out, size_t(in->size[0]) * in->size[1] * in->size[2], s);
    else
    {
      const double theta = std::min(-1 * r.inferior, r.superior);
      std::vector<double> g4(4 * size_t(in->size[2]));
      for (int p = 0; p < in->size[2]; p++)
      {
        const auto & v = g->views[size_t(in->index[2]) + p];
        const double row[4] = { v.sdd, v.srcX, v.projX, v.sid };
        std::copy(row, row + 4, g4.begin() + 4 * size_t(p));
      }
      const int rc = rtkhip_displaced_detector_weight(
        in->size, r.out.size[0], r.offset, g4.data(), theta, (r.superior + r.inferior > 0.) ? 1 : 0,
        in->origin[0] + in->spacing[0] * in->index[0], in->spacing[0], dev_in, dev_out, stream);
      RTKHIP_REQUIRE(rc == RTKHIP_OK, rc, rtkhip_last_error());
    }
    RTKHIP_CHECK(hipStreamSynchronize(s));
  });
}
