// rtkref — CPU oracle for the projector hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is an ITK-free C++ restatement of RTK's CPU
// filters, templated on the pixel type (float = the reference's float path,
// double = the FP64 oracle "same filters on double pixels").  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the correctness reference or the timed CPU baseline -- never as the product
// path.
//
// Every function cites the reference file:line whose behaviour it restates.
// Coordinates, matrices and interpolation weights are computed in double as RTK
// does (CoordinateType = double); pixel arithmetic follows the reference's
// operand types so the float instantiation reproduces the float CPU filters.
//
// Pinning (see tests/test_oracle_*.py):
//   geometry        closed form documentation/docs/Geometry.md:60-140 + invariants
//   ramp            known answers test/rtkrampfiltertest2.cxx:54-83 (1e-6)
//   fdk weight      closed form cos(gamma) = SID-ray geometry, brute force
//   voxel/FDK BP    brute-force counts (constant-1 projections), sum of w^2
//   Joseph FP       constant-1 volume -> analytic voxel-centre box chord; an
//                   independent numpy brute force on tiny random volumes (all three
//                   dominant axes, partial end steps, clamped border taps, 1e-12)
//   CG              SPD invariants, monotone residual, double vs float
//   phantoms        Shepp-Logan table src/rtkSheppLoganPhantom.cxx:23-35,
//                   chord length of a sphere in closed form, SetEllipsoid's y rotation
//   truncation      padding recovered from filtered unit rows by least squares
//   mirror          (Ohnesorge eqs. 3a/3b, RTK's weights); thresholds of
//                   test/rtkrampfiltertest.cxx:121-134 pass with it, fail without
//   ITK generic     brute force with ITK's border (index clamped to [0, n - 1]
//   interpolation   inside [-0.5, n - 0.5)) on out-of-plane/in-plane geometries
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <thread>
#include <vector>

namespace rtkref
{

constexpr double kPi = 3.141592653589793238462643383279502884;

// ---------------------------------------------------------------- matrices --
template <int R, int C>
struct Mat
{
  double a[R][C];
  Mat() { std::memset(a, 0, sizeof(a)); }
  double *       operator[](int i) { return a[i]; }
  const double * operator[](int i) const { return a[i]; }
  static Mat Identity()
  {
    Mat m;
    for (int i = 0; i < R && i < C; i++)
      m.a[i][i] = 1.;
    return m;
  }
};

template <int R, int K, int C>
Mat<R, C>
operator*(const Mat<R, K> & x, const Mat<K, C> & y)
{
  Mat<R, C> z;
  for (int i = 0; i < R; i++)
    for (int j = 0; j < C; j++)
    {
      double s = 0.;
      for (int k = 0; k < K; k++)
        s += x.a[i][k] * y.a[k][j];
      z.a[i][j] = s;
    }
  return z;
}

using M33 = Mat<3, 3>;
using M34 = Mat<3, 4>;
using M44 = Mat<4, 4>;

// General 4x4 inverse by Gauss-Jordan with partial pivoting (vnl_inverse is
// used by RTK; for the orthonormal / diagonal matrices involved this agrees to
// rounding).
inline M44
Inverse(const M44 & m)
{
  double a[4][8];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 8; j++)
      a[i][j] = j < 4 ? m.a[i][j] : (j - 4 == i ? 1. : 0.);
  for (int c = 0; c < 4; c++)
  {
    int p = c;
    for (int r = c + 1; r < 4; r++)
      if (std::abs(a[r][c]) > std::abs(a[p][c]))
        p = r;
    for (int j = 0; j < 8; j++)
      std::swap(a[c][j], a[p][j]);
    double d = a[c][c];
    for (int j = 0; j < 8; j++)
      a[c][j] /= d;
    for (int r = 0; r < 4; r++)
      if (r != c)
      {
        double f = a[r][c];
        for (int j = 0; j < 8; j++)
          a[r][j] -= f * a[c][j];
      }
  }
  M44 out;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      out.a[i][j] = a[i][j + 4];
  return out;
}

// ----------------------------------------------------------- image meta-data --
// x-fastest buffer, identity direction, buffered region starting at index 0.
struct ImageInfo
{
  int    size[3];
  double origin[3];
  double spacing[3];
  size_t N() const { return size_t(size[0]) * size[1] * size[2]; }
};

// include/rtkHomogeneousMatrix.h:38-66 (GetIndexToPhysicalPointMatrix)
inline M44
IndexToPhysical(const ImageInfo & im)
{
  M44 m;
  for (int i = 0; i < 3; i++)
  {
    m.a[i][i] = im.spacing[i];
    m.a[i][3] = im.origin[i];
  }
  m.a[3][3] = 1.;
  return m;
}

// include/rtkHomogeneousMatrix.h:78-84 (GetPhysicalPointToIndexMatrix)
inline M44
PhysicalToIndex(const ImageInfo & im)
{
  return Inverse(IndexToPhysical(im));
}

// ------------------------------------------------------------------ geometry --
// src/rtkThreeDCircularProjectionGeometry.cxx
struct Geometry
{
  std::vector<double> gantry, oop, ip, sid, srcX, srcY, sdd, projX, projY, sourceAngles;
  std::vector<M33>    projTrans;
  std::vector<M34>    mag;
  std::vector<M44>    rot, srcTrans;
  std::vector<M34>    matrices;
  double              radius = 0.;

  static double To0_2Pi(double a) { return a - 2 * kPi * std::floor(a / (2 * kPi)); } // .cxx:41-45

  size_t size() const { return sid.size(); }

  // .cxx:536-549 + documentation/docs/Geometry.md:60-90; ITK Euler3D ZXY order,
  // R = Rz(angleZ) Rx(angleX) Ry(angleY).
  static M44 Rotation(double ax, double ay, double az)
  {
    const double cx = std::cos(ax), sx = std::sin(ax);
    const double cy = std::cos(ay), sy = std::sin(ay);
    const double cz = std::cos(az), sz = std::sin(az);
    M44          Rx = M44::Identity(), Ry = M44::Identity(), Rz = M44::Identity();
    Rx.a[1][1] = cx;
    Rx.a[1][2] = -sx;
    Rx.a[2][1] = sx;
    Rx.a[2][2] = cx;
    Ry.a[0][0] = cy;
    Ry.a[0][2] = sy;
    Ry.a[2][0] = -sy;
    Ry.a[2][2] = cy;
    Rz.a[0][0] = cz;
    Rz.a[0][1] = -sz;
    Rz.a[1][0] = sz;
    Rz.a[1][1] = cz;
    return Rz * Rx * Ry;
  }

  // .cxx:55-76 (degrees) and .cxx:79-153 (radians)
  void AddProjection(double sid_, double sdd_, double gantryDeg, double px = 0., double py = 0.,
                     double oopDeg = 0., double ipDeg = 0., double sx = 0., double sy = 0.)
  {
    const double d2r = std::atan(1.0) / 45.0;
    AddProjectionInRadians(sid_, sdd_, d2r * gantryDeg, px, py, d2r * oopDeg, d2r * ipDeg, sx, sy);
  }

  void AddProjectionInRadians(double sid_, double sdd_, double ga, double px, double py, double oa, double ia,
                              double sx, double sy)
  {
    if (!sdd.empty() && ((sdd_ == 0.) != (sdd[0] == 0.)))
      throw std::runtime_error("Cannot mix parallel and divergent projections");
    gantry.push_back(To0_2Pi(ga));
    oop.push_back(To0_2Pi(oa));
    ip.push_back(To0_2Pi(ia));
    sid.push_back(sid_);
    srcX.push_back(sx);
    srcY.push_back(sy);
    sdd.push_back(sdd_);
    projX.push_back(px);
    projY.push_back(py);

    // .cxx:548-582 sub-matrices
    M33 pt = M33::Identity();
    pt.a[0][2] = sx - px;
    pt.a[1][2] = sy - py;
    M34          mg;
    const double msdd = -sdd_, msid = -sid_;
    mg.a[0][0] = mg.a[1][1] = (msdd == 0.) ? 1. : msdd;
    mg.a[2][2] = (msdd == 0.) ? 0. : 1.;
    mg.a[2][3] = (msdd == 0.) ? 1. : msid;
    M44 st = M44::Identity();
    st.a[0][3] = -sx;
    st.a[1][3] = -sy;
    M44 r = Rotation(-oa, -ga, -ia);
    projTrans.push_back(pt);
    mag.push_back(mg);
    srcTrans.push_back(st);
    rot.push_back(r);
    matrices.push_back(pt * mg * st * r); // .cxx:125-130

    // .cxx:133-144 source angle
    double s[4];
    SourcePosition(size() - 1, s);
    s[1] = 0.;
    double n = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    double a = std::acos(s[2] / n);
    if (s[0] / n > 0.)
      a = 2. * kPi - a;
    sourceAngles.push_back(To0_2Pi(a));
  }

  // .cxx:597-608: R^-1 (srcX, srcY, sid, 1)
  void SourcePosition(size_t i, double out[4]) const
  {
    double v[4] = { srcX[i], srcY[i], sid[i], 1. };
    M44    ri = Inverse(rot[i]);
    for (int r = 0; r < 4; r++)
    {
      out[r] = 0.;
      for (int c = 0; c < 4; c++)
        out[r] += ri.a[r][c] * v[c];
    }
  }

  // .cxx:611-637
  M44 ProjectionCoordinatesToFixedSystem(size_t i) const
  {
    M44 m = M44::Identity();
    m.a[0][3] = projX[i];
    m.a[1][3] = projY[i];
    m.a[2][3] = (sdd[i] == 0.) ? -sid[i] : sid[i] - sdd[i];
    m.a[2][2] = 0.;
    return Inverse(rot[i]) * m;
  }

  // .cxx:407-418
  std::vector<double> TiltAngles() const
  {
    std::vector<double> t;
    for (size_t i = 0; i < size(); i++)
      t.push_back(To0_2Pi(-gantry[i] - sourceAngles[i]));
    return t;
  }

  // .cxx:447-478: gap from each angle to the next one in sorted order (the last wraps).
  std::vector<double> AngularGapsWithNext(const std::vector<double> & angles) const
  {
    const size_t        n = angles.size();
    std::vector<double> g(n, 0.);
    if (n == 1)
      g[0] = 2 * kPi;
    if (n < 2)
      return g;
    std::multimap<double, unsigned> s;
    for (unsigned i = 0; i < n; i++)
      s.insert({ angles[i], i });
    auto curr = s.begin(), next = s.begin();
    ++next;
    for (; next != s.end(); ++curr, ++next)
      g[curr->second] = next->first - curr->first;
    g[curr->second] = s.begin()->first + 2 * kPi - curr->first;
    return g;
  }

  // .cxx:641-663: coordinate on the untilted virtual detector through the isocentre.
  double ToUntiltedCoordinateAtIsocenter(size_t i, double tilted) const
  {
    const double sidv = sid[i], sx = srcX[i];
    const double sidu = std::sqrt(sidv * sidv + sx * sx);
    const double l = (tilted + projX[i] - sx) * sidv / sdd[i] + sx;
    const double cosa = sx / sidu;
    return l * std::abs(sidv) / (sidu - l * cosa);
  }

  // .cxx:481-533
  std::vector<double> AngularGaps(const std::vector<double> & angles) const
  {
    const size_t        n = angles.size();
    std::vector<double> g(n, 0.);
    if (n == 1)
      g[0] = 2 * kPi;
    if (n < 2)
      return g;
    std::multimap<double, unsigned> s;
    for (unsigned i = 0; i < n; i++)
      s.insert({ angles[i], i });
    auto prev = s.begin(), curr = s.begin(), next = s.begin();
    ++next;
    g[curr->second] = 0.5 * (next->first - s.rbegin()->first + 2 * kPi);
    ++curr;
    ++next;
    while (next != s.end())
    {
      g[curr->second] = 0.5 * (next->first - prev->first);
      ++prev;
      ++curr;
      ++next;
    }
    g[curr->second] = 0.5 * (s.begin()->first + 2 * kPi - prev->first);
    if (sdd[0] == 0.)
    {
      double mx = *std::max_element(g.begin(), g.end());
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

// ------------------------------------------------- backprojection matrices --
// include/rtkBackProjectionImageFilter.hxx:535-595
inline M34
IndexToIndexProjectionMatrix(const Geometry & g, size_t iProj, const ImageInfo & vol, const ImageInfo & stack,
                             bool transpose)
{
  M34 volIdxToProjPP = g.mag[iProj] * (g.srcTrans[iProj] * (g.rot[iProj] * IndexToPhysical(vol)));
  M44 sp = PhysicalToIndex(stack);
  M33 mp = M33::Identity();
  for (int i = 0; i < 2; i++)
  {
    mp.a[i][2] = sp.a[i][3];
    for (int j = 0; j < 2; j++)
      mp.a[i][j] = sp.a[i][j];
  }
  M33 flip = M33::Identity();
  if (transpose)
  {
    std::swap(flip.a[0][0], flip.a[0][1]);
    std::swap(flip.a[1][0], flip.a[1][1]);
  }
  M33 ppToIdx = flip * mp * g.projTrans[iProj];
  return ppToIdx * volIdxToProjPP;
}

// ---------------------------------------------------------------- threading --
template <class F>
void
ParallelFor(int n, int nthreads, F && f)
{
  nthreads = std::max(1, std::min(nthreads, n));
  if (nthreads == 1)
  {
    for (int i = 0; i < n; i++)
      f(i);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([&, t]() {
      for (int i = t; i < n; i += nthreads)
        f(i);
    });
  for (auto & x : th)
    x.join();
}

// Transposed copy of one projection (include/rtkBackProjectionImageFilter.hxx:480-531)
template <class T>
std::vector<T>
TransposedProjection(const T * stack, const ImageInfo & s, int iProj)
{
  const int      nu = s.size[0], nv = s.size[1];
  const T *      pi = stack + size_t(iProj) * nu * nv;
  std::vector<T> out(size_t(nu) * nv);
  for (int v = 0; v < nv; v++)
    for (int u = 0; u < nu; u++)
      out[size_t(u) * nv + v] = pi[size_t(v) * nu + u];
  return out;
}

// ITK LinearInterpolateImageFunction on a 2-D image (generic BP path,
// include/rtkBackProjectionImageFilter.hxx:250-277).  ITK is not available;
// restated from ITK 5.4's published 2-D EvaluateOptimized: inside the buffer
// ([-0.5, n - 0.5) per axis) it is bilinear at the index clamped to [0, n - 1].
template <class T>
struct LinearInterp2D
{
  const T * p;
  int       n0, n1;
  bool      IsInside(double x, double y) const
  {
    return x >= -0.5 && x < n0 - 0.5 && y >= -0.5 && y < n1 - 0.5;
  }
  double At(int i, int j) const { return double(p[size_t(j) * n0 + i]); }
  double Eval(double x, double y) const
  {
    int b0 = (int)std::floor(x), b1 = (int)std::floor(y);
    if (b0 < 0)
      b0 = 0;
    if (b1 < 0)
      b1 = 0;
    const double d0 = x - b0, d1 = y - b1;
    const double v00 = At(b0, b1);
    if (d0 <= 0. && d1 <= 0.)
      return v00;
    if (d1 <= 0.)
    {
      if (b0 + 1 > n0 - 1)
        return v00;
      return v00 + (At(b0 + 1, b1) - v00) * d0;
    }
    if (d0 <= 0.)
    {
      if (b1 + 1 > n1 - 1)
        return v00;
      return v00 + (At(b0, b1 + 1) - v00) * d1;
    }
    if (b0 + 1 > n0 - 1)
    {
      if (b1 + 1 > n1 - 1)
        return v00;
      return v00 + (At(b0, b1 + 1) - v00) * d1;
    }
    const double v10 = At(b0 + 1, b1);
    const double vx0 = v00 + (v10 - v00) * d0;
    if (b1 + 1 > n1 - 1)
      return vx0;
    const double v11 = At(b0 + 1, b1 + 1), v01 = At(b0, b1 + 1);
    const double vx1 = v01 + (v11 - v01) * d0;
    return vx0 + (vx1 - vx0) * d1;
  }
};

// ------------------------------------------------------------- backprojection --
// Voxel-driven backprojection, plain (include/rtkBackProjectionImageFilter.hxx:185-475)
// and FDK-weighted (include/rtkFDKBackProjectionImageFilter.hxx:52-306).
// vol is updated in place (out = in + sum over views).  `fdk` selects the FDK
// normalisation and w^2 weight.  `firstProj`/`nProj` select the views of `stack`
// (whose slice p corresponds to geometry entry firstProj+p when stackOffset=firstProj).
template <class T>
void
Backproject(const Geometry & g, const ImageInfo & vi, T * vol, const ImageInfo & si, const T * stack, bool fdk,
            int nthreads, int geomOffset = 0)
{
  const int nProj = si.size[2];
  if ((int)g.size() < geomOffset + nProj)
    throw std::runtime_error("Mismatch between the number of projections and the geometry entries.");
  if (g.radius != 0.)
    throw std::runtime_error("cylindrical detector not supported by the oracle");
  const int nx = vi.size[0], ny = vi.size[1], nz = vi.size[2];

  // Rotation centre index (FDK .hxx:79-85)
  double rc[3];
  for (int i = 0; i < 3; i++)
    rc[i] = (0. - vi.origin[i]) / vi.spacing[i];

  std::vector<M34>            mats(nProj);
  std::vector<std::vector<T>> projs(nProj);
  for (int p = 0; p < nProj; p++)
  {
    M34 m = IndexToIndexProjectionMatrix(g, geomOffset + p, vi, si, true);
    if (fdk)
    {
      double pf = m.a[2][3];
      for (int j = 0; j < 3; j++)
        pf += m.a[2][j] * rc[j];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++)
          m.a[r][c] /= pf;
    }
    mats[p] = m;
    projs[p] = TransposedProjection(stack, si, p);
  }
  // After the transpose, the projection image is (nv x nu): pSize[0]=nv, pSize[1]=nu.
  const int ps0 = si.size[1], ps1 = si.size[0];
  using CT = T; // ComponentType for plain BP (.hxx:391,456); FDK uses double (.hxx:190,270)

  // Each work unit is one k slice; every view is applied in order per voxel.
  ParallelFor(nz, nthreads, [&](int k) {
    for (int p = 0; p < nProj; p++)
    {
      const M34 & m = mats[p];
      const T *   proj = projs[p].data();
      if (std::abs(m.a[1][0]) < 1e-10 && std::abs(m.a[2][0]) < 1e-10)
      {
        // OptimizedBackprojectionX (.hxx:352-413 / FDK .hxx:152-228)
        for (int j = 0; j < ny; j++)
        {
          int    i = 0;
          double u = m.a[0][0] * i + m.a[0][1] * j + m.a[0][2] * k + m.a[0][3];
          double v = m.a[1][1] * j + m.a[1][2] * k + m.a[1][3];
          double w = m.a[2][1] * j + m.a[2][2] * k + m.a[2][3];
          w = 1 / w;
          u = u * w;
          v = v * w;
          double du = w * m.a[0][0];
          w *= w;
          int vi_ = (int)std::floor(v);
          if (vi_ >= 0 && vi_ < ps1 - 1)
          {
            const T * pp = proj + size_t(vi_) * ps0;
            T *       pv = vol + size_t(nx) * (j + size_t(k) * ny);
            for (; i < nx; i++, u += du)
            {
              int ui = (int)std::floor(u);
              if (ui >= 0 && ui < ps0 - 1)
              {
                if (fdk)
                {
                  double u1 = u - ui, u2 = 1.0 - u1, v1 = v - vi_, v2 = 1.0 - v1;
                  pv[i] += w * (v2 * (u2 * pp[ui] + u1 * pp[ui + 1]) + v1 * (u2 * pp[ui + ps0] + u1 * pp[ui + ps0 + 1]));
                }
                else
                {
                  CT v1 = v - vi_, v2 = 1.0 - v1, u1 = u - ui, u2 = 1.0 - u1;
                  pv[i] += v2 * (u2 * pp[ui] + u1 * pp[ui + 1]) + v1 * (u2 * pp[ui + ps0] + u1 * pp[ui + ps0 + 1]);
                }
              }
            }
          }
        }
      }
      else if (std::abs(m.a[1][1]) < 1e-10 && std::abs(m.a[2][1]) < 1e-10)
      {
        // OptimizedBackprojectionY (.hxx:417-475 / FDK .hxx:232-306)
        for (int i = 0; i < nx; i++)
        {
          int    j = 0;
          double u = m.a[0][0] * i + m.a[0][1] * j + m.a[0][2] * k + m.a[0][3];
          double v = m.a[1][0] * i + m.a[1][2] * k + m.a[1][3];
          double w = m.a[2][0] * i + m.a[2][2] * k + m.a[2][3];
          w = 1 / w;
          u = u * w;
          v = v * w;
          double du = w * m.a[0][1];
          w *= w;
          int vi_ = (int)std::floor(v);
          if (vi_ >= 0 && vi_ < ps1 - 1)
          {
            T * pv = vol + i + size_t(nx) * (j + size_t(k) * ny);
            for (; j < ny; j++, pv += nx, u += du)
            {
              int ui = (int)std::floor(u);
              if (ui >= 0 && ui < ps0 - 1)
              {
                const T * pp = proj + size_t(vi_) * ps0 + ui;
                if (fdk)
                {
                  double v1 = v - vi_, v2 = 1.0 - v1, u1 = u - ui, u2 = 1.0 - u1;
                  *pv += w * (v2 * (u2 * pp[0] + u1 * pp[1]) + v1 * (u2 * pp[ps0] + u1 * pp[ps0 + 1]));
                }
                else
                {
                  CT v1 = v - vi_, v2 = 1.0 - v1, u1 = u - ui, u2 = 1.0 - u1;
                  *pv += v2 * (u2 * pp[0] + u1 * pp[1]) + v1 * (u2 * pp[ps0] + u1 * pp[ps0 + 1]);
                }
              }
            }
          }
        }
      }
      else
      {
        // Generic path with ITK linear interpolation (.hxx:250-277, FDK .hxx:117-145)
        LinearInterp2D<T> itp{ proj, ps0, ps1 };
        for (int j = 0; j < ny; j++)
          for (int i = 0; i < nx; i++)
          {
            double x = m.a[0][3] + m.a[0][0] * i + m.a[0][1] * j + m.a[0][2] * k;
            double y = m.a[1][3] + m.a[1][0] * i + m.a[1][1] * j + m.a[1][2] * k;
            double pf = m.a[2][3] + m.a[2][0] * i + m.a[2][1] * j + m.a[2][2] * k;
            pf = 1 / pf;
            x *= pf;
            y *= pf;
            if (itp.IsInside(x, y))
            {
              T & o = vol[i + size_t(nx) * (j + size_t(k) * ny)];
              if (fdk)
                o = o + pf * pf * itp.Eval(x, y);
              else
                o = o + T(itp.Eval(x, y));
            }
          }
      }
    }
  });
}

// -------------------------------------------------------------- FDK weighting --
// include/rtkFDKWeightProjectionFilter.hxx:40-128.  `firstProj` is the index of
// the stack's first slice in the geometry.  In place allowed.
template <class T>
void
FDKWeight(const Geometry & g, const ImageInfo & si, const T * in, T * out, int nthreads, int firstProj = 0)
{
  std::vector<double> factor = g.AngularGaps(g.sourceAngles);
  std::vector<double> tilt = g.TiltAngles();
  for (size_t k = 0; k < factor.size(); k++)
  {
    const double sdd = g.sdd[k];
    if (sdd == 0.)
      factor[k] *= 0.5;
    else
    {
      double sp[4];
      g.SourcePosition(k, sp);
      const double sid = g.sid[k];
      factor[k] *= std::abs(sdd) / (2. * sid * sid);
      factor[k] *= std::sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]);
    }
  }
  const int nu = si.size[0], nv = si.size[1];
  ParallelFor(si.size[2], nthreads, [&](int kk) {
    const int    k = kk + firstProj;
    const T *    pi = in + size_t(kk) * nu * nv;
    T *          po = out + size_t(kk) * nu * nv;
    const double sdd = g.sdd[k];
    if (sdd != 0.)
    {
      double       p1 = si.origin[1] + g.projY[k] - g.srcY[k];
      const double cosa = std::cos(tilt[k]), sina = std::sin(tilt[k]), tana = std::tan(tilt[k]);
      const double sid = g.sid[k], sdd2 = sdd * sdd, RD = sdd - sid;
      const double numpart1 = sdd * (cosa + tana * sina), sddtana = sdd * tana;
      for (int j = 0; j < nv; j++, p1 += si.spacing[1])
      {
        double       p0 = si.origin[0] + g.projX[k] + tana * RD;
        const double sdd2y2 = sdd2 + p1 * p1;
        for (int i = 0; i < nu; i++, p0 += si.spacing[0])
        {
          const double denom = std::sqrt(sdd2y2 + std::pow(p0 - sddtana, 2.));
          const double cosGamma = (numpart1 - p0 * sina) / denom;
          po[size_t(j) * nu + i] = T(pi[size_t(j) * nu + i] * factor[k] * cosGamma);
        }
      }
    }
    else
    {
      for (int i = 0; i < nu * nv; i++)
        po[i] = T(pi[i] * factor[k]);
    }
  });
}

// ------------------------------------------------------- Parker short scan --
// include/rtkParkerShortScanImageFilter.hxx.  The stack's buffer column 0 sits at
// physical u = si.origin[0]; slice kk is geometry entry firstProj + kk.
struct ParkerInfo
{
  bool   isShortScan = false;
  double delta = 0., firstAngle = 0.;
  int    insufficientView = -1; // first view that triggers the reference's warning
};

// .hxx:41-110 (GenerateInputRequestedRegion)
inline ParkerInfo
ParkerSetup(const Geometry & g, const ImageInfo & si, double threshold)
{
  ParkerInfo                r;
  const std::vector<double> gaps = g.AngularGapsWithNext(g.gantry);
  const int                 n = int(gaps.size());
  if (n == 0)
    return r;
  int maxPos = 0;
  for (int i = 1; i < n; i++)
    if (gaps[i] > gaps[maxPos])
      maxPos = i;
  if (g.sdd[0] == 0. || gaps[maxPos] < threshold)
    return r;
  r.isShortScan = true;
  std::map<double, unsigned> sorted; // GetUniqueSortedAngles (.cxx:434-444)
  for (unsigned i = 0; i < g.gantry.size(); i++)
    sorted.insert({ g.gantry[i], i });
  auto itLast = sorted.find(g.gantry[maxPos]);
  auto itFirst = itLast;
  itFirst = (++itFirst == sorted.end()) ? sorted.begin() : itFirst;
  r.firstAngle = itFirst->first;
  double lastAngle = itLast->first;
  if (lastAngle < r.firstAngle)
    lastAngle += 2 * kPi;
  r.delta = 0.5 * (lastAngle - r.firstAngle - kPi);
  r.delta = r.delta - 2 * kPi * std::floor(r.delta / (2 * kPi));
  const double c1 = si.origin[0], c2 = si.origin[0] + si.spacing[0] * (si.size[0] - 1);
  for (int k = 0; k < si.size[2]; k++)
  {
    const double sox = g.srcX[k], sid = g.sid[k];
    const double invsid = 1. / std::sqrt(sid * sid + sox * sox);
    const double hw = std::min(std::abs(g.ToUntiltedCoordinateAtIsocenter(k, c1)),
                               std::abs(g.ToUntiltedCoordinateAtIsocenter(k, c2)));
    if (r.delta < std::atan(hw * invsid))
    {
      r.insufficientView = k;
      break;
    }
  }
  return r;
}

// .hxx:112-209 (DynamicThreadedGenerateData): one row of weights per view in the
// pixel type, multiplied into every row.  Not a short scan: out = in.
template <class T>
ParkerInfo
ParkerShortScan(const Geometry & g, const ImageInfo & si, const T * in, T * out, double threshold, int nthreads,
                int firstProj = 0)
{
  const ParkerInfo r = ParkerSetup(g, si, threshold);
  const int        nu = si.size[0], nv = si.size[1];
  const size_t     perView = size_t(nu) * nv;
  if (!r.isShortScan)
  {
    if (in != out)
      std::copy(in, in + perView * si.size[2], out);
    return r;
  }
  ParallelFor(si.size[2], nthreads, [&](int kk) {
    const int    k = kk + firstProj;
    const double sox = g.srcX[k], sid = g.sid[k];
    const double invsid = 1. / std::sqrt(sid * sid + sox * sox);
    double       beta = g.gantry[k] - r.firstAngle;
    if (beta < 0)
      beta += 2 * kPi;
    std::vector<T> w(nu);
    double         point = si.origin[0];
    for (int i = 0; i < nu; i++, point += si.spacing[0])
    {
      const double l = g.ToUntiltedCoordinateAtIsocenter(k, point);
      const double alpha = std::atan(-1 * l * invsid);
      const double d = r.delta;
      if (beta <= 2 * d - 2 * alpha)
        w[i] = T(2. * std::pow(std::sin((kPi * beta) / (4 * (d - alpha))), 2.));
      else if (beta <= kPi - 2 * alpha)
        w[i] = T(2.);
      else if (beta <= kPi + 2 * d)
        w[i] = T(2. * std::pow(std::sin((kPi * (kPi + 2 * d - beta)) / (4 * (d + alpha))), 2.));
      else
        w[i] = T(0.);
    }
    const T * pi = in + kk * perView;
    T *       po = out + kk * perView;
    for (int j = 0; j < nv; j++)
      for (int i = 0; i < nu; i++)
        po[size_t(j) * nu + i] = pi[size_t(j) * nu + i] * w[i];
  });
  return r;
}

// -------------------------------------------------------- displaced detector --
// include/rtkDisplacedDetectorImageFilter.hxx.  The padded output is described by
// its own ImageInfo (origin of its buffer column 0); input column i lands in
// output column i + offset.
struct DisplacedInfo
{
  ImageInfo out;
  int       offset = 0;
  double    inferior = 0., superior = 0.;
  bool      weighting = false; // displaced and not disabled
};

// .hxx:81-173 (GenerateOutputInformation).  Throws where the reference throws.
inline DisplacedInfo
DisplacedDetectorSetup(const Geometry & g, const ImageInfo & si, bool disable, bool padOnTruncatedSide,
                       bool offsetsSet, double minOffset, double maxOffset, double radius = 0.)
{
  DisplacedInfo r;
  r.out = si;
  if (disable)
    return r;
  if (radius != 0.)
    throw std::runtime_error("Displaced detector cannot handle cylindrical detector.");
  if (g.sdd.front() == 0.)
    throw std::runtime_error("Displaced detector cannot handle parallel geometry.");
  double inf = si.origin[0], sup = inf;
  if (si.spacing[0] < 0.)
    inf += si.spacing[0] * (si.size[0] - 1);
  else
    sup += si.spacing[0] * (si.size[0] - 1);
  if (!offsetsSet)
  {
    double minSup = std::numeric_limits<double>::max();
    double maxInf = -std::numeric_limits<double>::max();
    for (size_t i = 0; i < g.size(); i++)
    {
      maxInf = std::max(maxInf, g.ToUntiltedCoordinateAtIsocenter(i, inf));
      minSup = std::min(minSup, g.ToUntiltedCoordinateAtIsocenter(i, sup));
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
  if (inf > 0. || sup < 0.)
    throw std::runtime_error("Cannot account for detector displacement larger than 50% of panel size.");
  r.weighting = !(std::abs(inf + sup) < 0.1 * std::abs(sup - inf));
  if (!r.weighting || !padOnTruncatedSide)
    return r;
  r.out.size[0] = 2 * si.size[0];
  if (sup + inf > 0.)
  {
    r.offset = si.size[0];
    r.out.origin[0] = si.origin[0] - si.spacing[0] * si.size[0];
  }
  return r;
}

// .hxx:175-323 (DynamicThreadedGenerateData): weights in double, zeros outside the
// input's columns.  `out` has r.out's size.
template <class T>
void
DisplacedDetector(const Geometry & g, const DisplacedInfo & r, const ImageInfo & si, const T * in, T * out,
                  int nthreads, int firstProj = 0)
{
  const int    nu = si.size[0], nv = si.size[1], nuo = r.out.size[0];
  const double theta = std::min(-1 * r.inferior, r.superior);
  ParallelFor(si.size[2], nthreads, [&](int kk) {
    const int           k = kk + firstProj;
    std::vector<double> w(nu, 1.);
    if (r.weighting)
    {
      const double sx = g.srcX[k];
      double       sid = g.sid[k];
      sid = std::sqrt(sid * sid + sx * sx);
      double invsid = 0., invden = 0.;
      if (sid != 0.)
      {
        invsid = 1. / sid;
        invden = 1. / (2. * std::atan(theta * invsid));
      }
      double point = si.origin[0];
      for (int i = 0; i < nu; i++, point += si.spacing[0])
      {
        const double l = g.ToUntiltedCoordinateAtIsocenter(k, point);
        if (r.superior + r.inferior > 0.)
          w[i] = (l <= -1 * theta) ? 0.0 : (l >= theta) ? 2.0 : std::sin(kPi * std::atan(l * invsid) * invden) + 1;
        else
          w[i] = (l <= -1 * theta) ? 2.0 : (l >= theta) ? 0.0 : 1 - std::sin(kPi * std::atan(l * invsid) * invden);
      }
    }
    const T * pi = in + size_t(kk) * nu * nv;
    T *       po = out + size_t(kk) * nuo * nv;
    for (int j = 0; j < nv; j++)
    {
      T * row = po + size_t(j) * nuo;
      std::fill(row, row + nuo, T(0));
      for (int i = 0; i < nu; i++)
        row[i + r.offset] = r.weighting ? T(pi[size_t(j) * nu + i] * w[i]) : pi[size_t(j) * nu + i];
    }
  });
}

// ----------------------------------------------------------------------- FFT --
using cplx = std::complex<double>;

inline bool
IsPow2(int n)
{
  return n > 0 && (n & (n - 1)) == 0;
}

// In-place iterative radix-2 DIT FFT; sign=-1 forward (unnormalised), +1 inverse (unnormalised).
inline void
FFTPow2(cplx * a, int n, int sign)
{
  for (int i = 1, j = 0; i < n; i++)
  {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1)
      j ^= bit;
    j ^= bit;
    if (i < j)
      std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1)
  {
    const double ang = sign * 2 * kPi / len;
    for (int i = 0; i < n; i += len)
      for (int k = 0; k < len / 2; k++)
      {
        cplx w(std::cos(ang * k), std::sin(ang * k));
        cplx x = a[i + k], y = a[i + k + len / 2] * w;
        a[i + k] = x + y;
        a[i + k + len / 2] = x - y;
      }
  }
}

// Arbitrary-length DFT (Bluestein over a power-of-two FFT); same sign convention.
inline void
FFT(std::vector<cplx> & a, int sign)
{
  const int n = (int)a.size();
  if (IsPow2(n))
  {
    FFTPow2(a.data(), n, sign);
    return;
  }
  int m = 1;
  while (m < 2 * n - 1)
    m <<= 1;
  std::vector<cplx> w(n), A(m, 0.), B(m, 0.);
  for (int k = 0; k < n; k++)
  {
    double ang = sign * kPi * double((long long)k * k % (2LL * n)) / n;
    w[k] = cplx(std::cos(ang), std::sin(ang));
  }
  for (int k = 0; k < n; k++)
    A[k] = a[k] * w[k];
  B[0] = std::conj(w[0]);
  for (int k = 1; k < n; k++)
    B[k] = B[m - k] = std::conj(w[k]);
  FFTPow2(A.data(), m, -1);
  FFTPow2(B.data(), m, -1);
  for (int k = 0; k < m; k++)
    A[k] *= B[k];
  FFTPow2(A.data(), m, +1);
  for (int k = 0; k < n; k++)
    a[k] = A[k] * w[k] / double(m);
}

// --------------------------------------------------------------- ramp filter --
struct RampOptions
{
  double hann = 0., hannY = 0., cosine = 0., hamming = 0., ramlak = 0., sheppLogan = 0.;
  double truncation = 0.;
  int    greatestPrimeFactor = 2; // 2 without FFTW, 13 with (include/rtkFFTProjectionsConvolutionImageFilter.hxx:35-48)
  int    zeroPadX = 2, zeroPadY = 2;
};

// include/rtkFFTProjectionsConvolutionImageFilter.hxx:318-341
inline int
GreatestPrimeFactor(int n)
{
  int v = 2;
  while (v <= n)
  {
    bool prime = true;
    for (int x = 2; x <= (int)std::sqrt(double(v)); x++)
      if (v % x == 0)
      {
        prime = false;
        break;
      }
    if (n % v == 0 && prime)
      n /= v;
    else
      v += 1;
  }
  return v;
}

// include/rtkFFTProjectionsConvolutionImageFilter.hxx:280-306
inline void
PaddedSize(int nu, int nv, const RampOptions & o, int & xpad, int & ypad, long & zeroext)
{
  xpad = o.zeroPadX * nu;
  while (GreatestPrimeFactor(xpad) > o.greatestPrimeFactor)
    xpad++;
  zeroext = ((long)xpad - nu) / 2;
  ypad = nv;
  if (o.hannY != 0.)
    ypad *= o.zeroPadY;
  while (GreatestPrimeFactor(ypad) > o.greatestPrimeFactor)
    ypad++;
}

// Kernel spectrum (xpad/2+1 columns x kh rows, kh = 1 or ypad) of
// include/rtkFFTRampImageFilter.hxx:47-178.
inline std::vector<cplx>
RampKernelSpectrum(int xpad, int ypad, double spacing, const RampOptions & o, int & kh)
{
  std::vector<cplx> k(xpad, 0.);
  k[0] = 1. / (4. * spacing);
  for (int ix = 1, jx = xpad - 1; ix < xpad / 2; ix += 2, jx -= 2)
  {
    double v = ix * kPi;
    v = -1. / (v * v * spacing);
    k[ix] = v;
    k[jx] = v;
  }
  FFT(k, -1);
  const int         n = xpad / 2 + 1;
  std::vector<cplx> K(k.begin(), k.begin() + n);
  int               i = 0;
  auto              rnd = [](double x) { return (unsigned)std::floor(x + 0.5); }; // itk::Math::Round (half up)
  if (o.hann > 0.)
  {
    unsigned ncut = rnd(n * std::min(1.0, o.hann));
    for (; i < (int)ncut; i++)
      K[i] *= 0.5 * (1 + std::cos(kPi * i / ncut));
  }
  else if (o.cosine > 0.)
  {
    unsigned ncut = rnd(n * std::min(1.0, o.cosine));
    for (; i < (int)ncut; i++)
      K[i] *= std::cos(0.5 * kPi * i / ncut);
  }
  else if (o.hamming > 0.)
  {
    unsigned ncut = rnd(n * std::min(1.0, o.hamming));
    for (; i < (int)ncut; i++)
      K[i] *= 0.54 + 0.46 * std::cos(kPi * i / ncut);
  }
  else if (o.ramlak > 0.)
  {
    i = (int)rnd(n * std::min(1.0, o.ramlak));
  }
  else if (o.sheppLogan > 0.)
  {
    unsigned ncut = rnd(n * std::min(1.0, o.sheppLogan));
    i = 1;
    for (; i < (int)ncut; i++)
    {
      double x = 0.5 * kPi * i / ncut;
      K[i] *= std::sin(x) / x;
    }
  }
  else
    i = n;
  for (; i < n; i++)
    K[i] = 0.;
  kh = 1;
  if (o.hannY > 0.)
  {
    kh = ypad;
    std::vector<cplx> K2(size_t(n) * ypad, 0.);
    unsigned          ncut = rnd((ypad / 2 + 1) * std::min(1.0, o.hannY));
    for (unsigned j = 0; j < ncut; j++)
    {
      double win = 0.5 * (1 + std::cos(kPi * j / ncut));
      for (int x = 0; x < n; x++)
        K2[size_t(j) * n + x] = win * K[x];
    }
    for (unsigned j = 1; j < ncut; j++)
    {
      double win = 0.5 * (1 + std::cos(kPi * j / ncut));
      for (int x = 0; x < n; x++)
        K2[size_t(ypad - j) * n + x] = win * K[x];
    }
    return K2;
  }
  return K;
}

// FFTRampImageFilter<float/double in, TFFTPrecision=double>: pad (incl. Ohnesorge
// truncation mirror), R2C, multiply, C2R (1/N), crop.
// include/rtkFFTProjectionsConvolutionImageFilter.hxx:121-276,345-355.
template <class T>
void
RampFilter(const ImageInfo & si, const T * in, T * out, const RampOptions & o, int nthreads)
{
  const int nu = si.size[0], nv = si.size[1], np = si.size[2];
  int       xpad, ypad, kh;
  long      zeroext;
  PaddedSize(nu, nv, o, xpad, ypad, zeroext);
  const std::vector<cplx> K = RampKernelSpectrum(xpad, ypad, si.spacing[0], o, kh);
  const int               n = xpad / 2 + 1;
  // UpdateTruncationMirrorWeights (.hxx:345-355) sizes the weights with the full
  // extent; the mirrored extent is clipped by the zero padding (.hxx:213-214).
  const int               nextFull = (int)std::floor(o.truncation * nu);
  const int               next = (int)std::min<long>(zeroext, nextFull);
  std::vector<double>     mw(nextFull + 1);
  for (int i = 0; i < nextFull + 1; i++)
    mw[i] = std::pow(std::sin((nextFull - i) * kPi / (2 * nextFull - 2)), 0.75);

  ParallelFor(np, nthreads, [&](int p) {
    const T *           pi = in + size_t(p) * nu * nv;
    T *                 po = out + size_t(p) * nu * nv;
    std::vector<double> pad(size_t(xpad) * ypad, 0.);
    for (int j = 0; j < nv; j++)
    {
      double *  row = pad.data() + size_t(j) * xpad;
      const T * r = pi + size_t(j) * nu;
      if (next)
      {
        // Left mirror (eq. 3a, Ohnesorge 2000): reference pixel is input index 1
        // (.hxx:222-224); padded position zeroext-b mirrors input index b.
        const double sa = double(r[1]);
        for (int b = 1; b <= next; b++)
          row[zeroext - b] = mw[b] * (2.0 * sa - double(r[b]));
        // Right mirror (eq. 3b): reference pixel is the last input pixel (.hxx:243-245).
        const double se = double(r[nu - 1]);
        for (int b = 1; b <= next; b++)
          row[zeroext + nu - 1 + b] = mw[b] * (2.0 * se - double(r[nu - 1 - b]));
      }
      for (int i = 0; i < nu; i++)
        row[zeroext + i] = double(r[i]);
    }
    // Forward R2C along x, full complex along y when the kernel is 2-D.
    std::vector<cplx> spec(size_t(n) * ypad);
    std::vector<cplx> line(xpad);
    for (int j = 0; j < ypad; j++)
    {
      for (int i = 0; i < xpad; i++)
        line[i] = pad[size_t(j) * xpad + i];
      FFT(line, -1);
      for (int i = 0; i < n; i++)
        spec[size_t(j) * n + i] = line[i];
    }
    if (kh > 1)
    {
      // 2-D kernel: forward along y, multiply, inverse along y (1/ypad).
      std::vector<cplx> col(ypad);
      for (int i = 0; i < n; i++)
      {
        for (int j = 0; j < ypad; j++)
          col[j] = spec[size_t(j) * n + i];
        FFT(col, -1);
        for (int j = 0; j < ypad; j++)
          col[j] *= K[size_t(j) * n + i];
        FFT(col, +1);
        for (int j = 0; j < ypad; j++)
          spec[size_t(j) * n + i] = col[j] / double(ypad);
      }
    }
    else
    {
      for (int j = 0; j < ypad; j++)
        for (int i = 0; i < n; i++)
          spec[size_t(j) * n + i] *= K[i];
    }
    // C2R along x (Hermitian extension), normalised by 1/xpad.
    for (int j = 0; j < nv; j++)
    {
      for (int i = 0; i < n; i++)
        line[i] = spec[size_t(j) * n + i];
      for (int i = n; i < xpad; i++)
        line[i] = std::conj(spec[size_t(j) * n + (xpad - i)]);
      FFT(line, +1);
      for (int i = 0; i < nu; i++)
        po[size_t(j) * nu + i] = T(line[zeroext + i].real() / double(xpad));
    }
  });
}

// ----------------------------------------------------------------- Joseph FP --
// include/rtkJosephForwardProjectionImageFilter.hxx:182-543 with the default
// functors (.h:42-146), BoxShape::IsIntersectedByRay (src/rtkBoxShape.cxx:48-97)
// and the flat-panel ray iterator
// (include/rtkProjectionsRegionConstIteratorRayBasedWithFlatPanel.hxx:40-71).
// out = in + sum * |stepMM|.  Returns the number of Joseph samples
// (sum over rays of |fs-ns|+1).
template <class T>
long long
JosephForward(const Geometry & g, const ImageInfo & si, const T * projIn, T * projOut, const ImageInfo & vi,
              const T * vol, int nthreads, int geomOffset = 0, double inferiorClipIn = 0., double superiorClipIn = 1.)
{
  if (g.radius != 0.)
    throw std::runtime_error("cylindrical detector not supported by the oracle");
  const int np = si.size[2], nu = si.size[0], nv = si.size[1];
  if ((int)g.size() < geomOffset + np)
    throw std::runtime_error("Mismatch between the number of projections and the geometry entries.");
  const long offsets[3] = { 1, vi.size[0], long(vi.size[0]) * vi.size[1] };
  const M44  volPPToIndex = PhysicalToIndex(vi);
  double     boxMin[3], boxMax[3];
  for (int i = 0; i < 3; i++)
  {
    boxMin[i] = 0.;
    boxMax[i] = (vi.size[i] - 1) * (1. - std::numeric_limits<double>::epsilon());
  }
  const double inferiorClip = 1. - superiorClipIn, superiorClip = 1. - inferiorClipIn;
  std::vector<long long> counts(np, 0);

  ParallelFor(np, nthreads, [&](int p) {
    const int gi = p + geomOffset;
    double    s4[4], src[3];
    g.SourcePosition(gi, s4);
    for (int r = 0; r < 3; r++)
    {
      src[r] = 0.;
      for (int c = 0; c < 4; c++)
        src[r] += volPPToIndex.a[r][c] * s4[c];
    }
    // The stack index of slice p is p (region index 0); its z does not matter.
    const M44 pm = volPPToIndex * (g.ProjectionCoordinatesToFixedSystem(gi) * IndexToPhysical(si));
    long long cnt = 0;
    for (int jv = 0; jv < nv; jv++)
      for (int iu = 0; iu < nu; iu++)
      {
        const size_t pix = size_t(p) * nu * nv + size_t(jv) * nu + iu;
        double       pos[3], dir[3];
        for (int r = 0; r < 3; r++)
        {
          pos[r] = pm.a[r][3] + pm.a[r][0] * iu + pm.a[r][1] * jv + pm.a[r][2] * p;
          dir[r] = -(pos[r] - src[r]);
        }
        int mainDir = 0;
        for (int i = 0; i < 3; i++)
          if (std::abs(dir[i]) > std::abs(dir[mainDir]))
            mainDir = i;
        // Box intersection (src/rtkBoxShape.cxx:48-97)
        double infDist = -std::numeric_limits<double>::max(), supDist = std::numeric_limits<double>::max();
        bool   hit = true;
        for (int i = 0; i < 3 && hit; i++)
        {
          if (dir[i] == 0. && (pos[i] < boxMin[i] || pos[i] > boxMax[i]))
          {
            hit = false;
            break;
          }
          double inv = 1 / dir[i];
          double T1 = (boxMin[i] - pos[i]) * inv, T2 = (boxMax[i] - pos[i]) * inv;
          if (T1 > T2)
            std::swap(T1, T2);
          infDist = std::max(T1, infDist);
          supDist = std::min(T2, supDist);
          if (infDist > supDist || supDist < 0.)
            hit = false;
        }
        infDist = std::max(infDist, inferiorClip);
        supDist = std::min(supDist, superiorClip);
        if (!(hit && supDist > 0. && infDist <= 1. && supDist > infDist))
        {
          projOut[pix] = projIn[pix];
          continue;
        }
        double npnt[3], fpnt[3];
        for (int r = 0; r < 3; r++)
        {
          npnt[r] = pos[r] + infDist * dir[r];
          fpnt[r] = pos[r] + supDist * dir[r];
        }
        auto      rnd = [](double x) { return (int)std::lround(x); }; // half away from zero (vnl_math::rnd)
        const int ns = rnd(npnt[mainDir]), fs = rnd(fpnt[mainDir]);
        int       dInf = (mainDir + 1) % 3, dSup = (mainDir + 2) % 3;
        if (dInf > dSup)
          std::swap(dInf, dSup);
        const double minx = boxMin[dInf], miny = boxMin[dSup], maxx = boxMax[dInf], maxy = boxMax[dSup];
        const long   ox = offsets[dInf], oy = offsets[dSup];
        long         oz = offsets[mainDir];
        const T *    pxiyi = vol + ns * oz;
        double       residualB = ns - npnt[mainDir], residualE = fpnt[mainDir] - fs;
        const double norm = 1. / dir[mainDir];
        double       stepx = dir[dInf] * norm, stepy = dir[dSup] * norm;
        if (npnt[mainDir] > fpnt[mainDir])
        {
          residualB *= -1;
          residualE *= -1;
          oz *= -1;
          stepx *= -1;
          stepy *= -1;
        }
        double currentx = npnt[dInf] + residualB * stepx;
        double currenty = npnt[dSup] + residualB * stepy;
        double stepMM[3];
        stepMM[dInf] = vi.spacing[dInf] * stepx;
        stepMM[dSup] = vi.spacing[dSup] * stepy;
        stepMM[mainDir] = vi.spacing[mainDir];

        auto bil = [&](const T * b, double x, double y) -> T {
          int    ix = (int)std::floor(x), iy = (int)std::floor(y);
          long   idx = ix * ox + iy * oy;
          double lx = x - ix, ly = y - iy, lxc = 1. - lx, lyc = 1. - ly;
          // InterpolationWeightMultiplication<T, ValueType<T>>: the weight is
          // converted to the pixel component type before the product (.h:42-69, .h:165-167).
          return T(T(T(lxc * lyc) * b[idx]) + T(T(lx * lyc) * b[idx + ox]) + T(T(lxc * ly) * b[idx + oy]) +
                   T(T(lx * ly) * b[idx + ox + oy]));
        };
        auto bilb = [&](double step, const T * b, double x, double y) -> T {
          int    ix = (int)std::floor(x), iy = (int)std::floor(y);
          long   idx = ix * ox + iy * oy;
          double lx = x - ix, ly = y - iy, lxc = 1. - lx, lyc = 1. - ly;
          long   xi = 0, yi = 0, xs = 0, ys = 0;
          if (ix < minx)
            xi = ox;
          if (iy < miny)
            yi = oy;
          if (ix >= maxx)
            xs = -ox;
          if (iy >= maxy)
            ys = -oy;
          T r = 0;
          r += T(T(lxc * lyc) * b[idx + xi + yi]);
          r += T(T(lxc * ly) * b[idx + oy + xi + ys]);
          r += T(T(lx * lyc) * b[idx + ox + xs + yi]);
          r += T(T(lx * ly) * b[idx + ox + oy + xs + ys]);
          r *= step;
          return r;
        };
        T sum = 0;
        if (fs == ns)
          sum += bilb(std::abs(fpnt[mainDir] - npnt[mainDir]), pxiyi, currentx, currenty);
        else
        {
          sum += bilb(residualB + 0.5, pxiyi, currentx, currenty);
          pxiyi += oz;
          currentx += stepx;
          currenty += stepy;
          for (int i = 0; i < std::abs(fs - ns) - 1; ++i)
          {
            sum += bil(pxiyi, currentx, currenty);
            pxiyi += oz;
            currentx += stepx;
            currenty += stepy;
          }
          sum += bilb(residualE + 0.5, pxiyi, currentx, currenty);
        }
        cnt += std::abs(fs - ns) + 1;
        const double nrm = std::sqrt(stepMM[0] * stepMM[0] + stepMM[1] * stepMM[1] + stepMM[2] * stepMM[2]);
        projOut[pix] = T(projIn[pix] + sum * nrm);
      }
    counts[p] = cnt;
  });
  long long total = 0;
  for (auto c : counts)
    total += c;
  return total;
}

// ------------------------------------------------------------------ Joseph BP --
// include/rtkJosephBackProjectionImageFilter.hxx:58-412 with the default functors
// (ValueAlongRay returns the pixel value; SplatWeightMultiplication adds
// value * weight * |stepMM| * stepLength, .h:110-136): the splat transpose of the
// Joseph forward projector.  vol is updated in place.  RTK runs it single-threaded;
// here views are split over threads, each with a private accumulator, summed in
// thread order.
template <class T>
void
JosephBack(const Geometry & g, const ImageInfo & vi, T * vol, const ImageInfo & si, const T * proj, int nthreads,
           int geomOffset = 0, double inferiorClipIn = 0., double superiorClipIn = 1.)
{
  if (g.radius != 0.)
    throw std::runtime_error("cylindrical detector not supported by the oracle");
  const int np = si.size[2], nu = si.size[0], nv = si.size[1];
  if ((int)g.size() < geomOffset + np)
    throw std::runtime_error("Mismatch between the number of projections and the geometry entries.");
  const long offsets[3] = { 1, vi.size[0], long(vi.size[0]) * vi.size[1] };
  const M44  volPPToIndex = PhysicalToIndex(vi);
  double     boxMin[3], boxMax[3];
  for (int i = 0; i < 3; i++)
  {
    boxMin[i] = 0.;
    boxMax[i] = (vi.size[i] - 1) * (1. - std::numeric_limits<double>::epsilon());
  }
  const double inferiorClip = 1. - superiorClipIn, superiorClip = 1. - inferiorClipIn;
  nthreads = std::max(1, std::min(nthreads, np));
  std::vector<std::vector<T>> acc(nthreads, std::vector<T>(vi.N(), T(0)));
  std::vector<std::thread>    th;
  for (int tid = 0; tid < nthreads; tid++)
    th.emplace_back([&, tid]() {
      T * out = acc[tid].data();
      for (int p = tid; p < np; p += nthreads)
      {
        const int gi = p + geomOffset;
        double    s4[4], src[3];
        g.SourcePosition(gi, s4);
        for (int r = 0; r < 3; r++)
        {
          src[r] = 0.;
          for (int c = 0; c < 4; c++)
            src[r] += volPPToIndex.a[r][c] * s4[c];
        }
        const M44 pm = volPPToIndex * (g.ProjectionCoordinatesToFixedSystem(gi) * IndexToPhysical(si));
        for (int jv = 0; jv < nv; jv++)
          for (int iu = 0; iu < nu; iu++)
          {
            const T rayValue = proj[size_t(p) * nu * nv + size_t(jv) * nu + iu];
            double  pos[3], dir[3];
            for (int r = 0; r < 3; r++)
            {
              pos[r] = pm.a[r][3] + pm.a[r][0] * iu + pm.a[r][1] * jv + pm.a[r][2] * p;
              dir[r] = -(pos[r] - src[r]);
            }
            int mainDir = 0;
            for (int i = 0; i < 3; i++)
              if (std::abs(dir[i]) > std::abs(dir[mainDir]))
                mainDir = i;
            double infDist = -std::numeric_limits<double>::max(), supDist = std::numeric_limits<double>::max();
            bool   hit = true;
            for (int i = 0; i < 3 && hit; i++)
            {
              if (dir[i] == 0. && (pos[i] < boxMin[i] || pos[i] > boxMax[i]))
              {
                hit = false;
                break;
              }
              double inv = 1 / dir[i];
              double T1 = (boxMin[i] - pos[i]) * inv, T2 = (boxMax[i] - pos[i]) * inv;
              if (T1 > T2)
                std::swap(T1, T2);
              infDist = std::max(T1, infDist);
              supDist = std::min(T2, supDist);
              if (infDist > supDist || supDist < 0.)
                hit = false;
            }
            // .hxx:148-155: test before clipping
            if (!(hit && supDist >= 0. && infDist <= 1.))
              continue;
            infDist = std::max(infDist, inferiorClip);
            supDist = std::min(supDist, superiorClip);
            double npnt[3], fpnt[3];
            for (int r = 0; r < 3; r++)
            {
              npnt[r] = pos[r] + infDist * dir[r];
              fpnt[r] = pos[r] + supDist * dir[r];
            }
            const int ns = (int)std::lround(npnt[mainDir]), fs = (int)std::lround(fpnt[mainDir]);
            int       dInf = (mainDir + 1) % 3, dSup = (mainDir + 2) % 3;
            if (dInf > dSup)
              std::swap(dInf, dSup);
            const double maxx = boxMax[dInf], maxy = boxMax[dSup];
            const long   ox = offsets[dInf], oy = offsets[dSup];
            long         oz = offsets[mainDir];
            T *          pxiyi = out + ns * oz;
            double       residualB = ns - npnt[mainDir], residualE = fpnt[mainDir] - fs;
            const double norm = 1. / dir[mainDir];
            double       stepx = dir[dInf] * norm, stepy = dir[dSup] * norm;
            if (npnt[mainDir] > fpnt[mainDir])
            {
              residualB *= -1;
              residualE *= -1;
              oz *= -1;
              stepx *= -1;
              stepy *= -1;
            }
            double currentx = npnt[dInf] + residualB * stepx;
            double currenty = npnt[dSup] + residualB * stepy;
            double stepMM[3];
            stepMM[dInf] = vi.spacing[dInf] * stepx;
            stepMM[dSup] = vi.spacing[dSup] * stepy;
            stepMM[mainDir] = vi.spacing[mainDir];
            const double voxelSize = std::sqrt(stepMM[0] * stepMM[0] + stepMM[1] * stepMM[1] + stepMM[2] * stepMM[2]);
            // BilinearSplat (.hxx:296-318) and BilinearSplatOnBorders (.hxx:326-370)
            auto splat = [&](T * b, double step, double x, double y, bool borders) {
              int    ix = (int)std::floor(x), iy = (int)std::floor(y);
              long   idx = ix * ox + iy * oy;
              double lx = x - ix, ly = y - iy, lxc = 1. - lx, lyc = 1. - ly;
              long   xi = 0, yi = 0, xs = 0, ys = 0;
              if (borders)
              {
                if (ix < 0)
                  xi = ox;
                if (iy < 0)
                  yi = oy;
                if (ix >= maxx)
                  xs = -ox;
                if (iy >= maxy)
                  ys = -oy;
              }
              b[idx + xi + yi] += rayValue * (lxc * lyc) * voxelSize * step;
              b[idx + oy + xi + ys] += rayValue * (lxc * ly) * voxelSize * step;
              b[idx + ox + xs + yi] += rayValue * (lx * lyc) * voxelSize * step;
              b[idx + ox + oy + xs + ys] += rayValue * (lx * ly) * voxelSize * step;
            };
            if (fs == ns)
              splat(pxiyi, std::abs(fpnt[mainDir] - npnt[mainDir]), currentx, currenty, true);
            else
            {
              splat(pxiyi, residualB + 0.5, currentx, currenty, true);
              pxiyi += oz;
              currentx += stepx;
              currenty += stepy;
              for (int i = 0; i < std::abs(fs - ns) - 1; ++i)
              {
                splat(pxiyi, 1.0, currentx, currenty, false);
                pxiyi += oz;
                currentx += stepx;
                currenty += stepy;
              }
              splat(pxiyi, residualE + 0.5, currentx, currenty, true);
            }
          }
      }
    });
  for (auto & x : th)
    x.join();
  for (int t = 0; t < nthreads; t++)
    for (size_t i = 0; i < vi.N(); i++)
      vol[i] += acc[t][i];
}

// ------------------------------------------------------------------------ CG --
// ConjugateGradientConeBeamReconstructionFilter (include/rtkConjugateGradientConeBeamReconstructionFilter.hxx:195-305)
// with A x = BP_voxel(0, W * FP_joseph(0, x)) (include/rtkReconstructionConjugateGradientOperator.hxx:194-286)
// and the CG loop of include/rtkConjugateGradientImageFilter.hxx:108-258 with
// scalars in the pixel type.  `weights` may be null (ones).  `x` holds X0 on
// entry and the result on exit.  resid[it] (if non-null, niter+1 entries) gets
// ||R_k||^2 as tracked by the loop (numerator), computed in the pixel type.
template <class T>
void
ConjugateGradient(const Geometry & g, const ImageInfo & vi, T * x, const ImageInfo & si, const T * proj,
                  const T * weights, int niter, int nthreads, double * resid = nullptr, bool josephBP = false)
{
  const size_t   nv = vi.N(), ns = si.N();
  std::vector<T> pw(ns), tmp(ns), B(nv, 0), AX(nv), R(nv), P(nv);
  for (size_t i = 0; i < ns; i++)
    pw[i] = weights ? T(proj[i] * weights[i]) : proj[i];
  // BP_VOXELBASED (default) or BP_JOSEPH (include/rtkIterativeConeBeamReconstructionFilter.hxx:35-103)
  auto bp = [&](T * out, const T * p) {
    if (josephBP)
      JosephBack(g, vi, out, si, p, nthreads);
    else
      Backproject(g, vi, out, si, p, false, nthreads);
  };
  bp(B.data(), pw.data());
  auto applyA = [&](const T * in, T * out) {
    std::vector<T> zp(ns, 0);
    JosephForward(g, si, zp.data(), tmp.data(), vi, in, nthreads);
    if (weights)
      for (size_t i = 0; i < ns; i++)
        tmp[i] = T(tmp[i] * weights[i]);
    std::fill(out, out + nv, T(0));
    bp(out, tmp.data());
  };
  applyA(x, AX.data());
  for (size_t i = 0; i < nv; i++)
  {
    R[i] = B[i] - AX[i];
    P[i] = R[i];
  }
  const T eps = std::numeric_limits<T>::min();
  T       num = 0, den = 0, alpha, beta;
  if (resid)
  {
    T r2 = 0;
    for (size_t i = 0; i < nv; i++)
      r2 += R[i] * R[i];
    resid[0] = double(r2);
  }
  for (int it = 0; it < niter; it++)
  {
    applyA(P.data(), AX.data());
    num = 0;
    den = 0;
    for (size_t i = 0; i < nv; i++)
    {
      num += R[i] * R[i];
      den += P[i] * AX[i];
    }
    alpha = num / (den + eps);
    for (size_t i = 0; i < nv; i++)
      x[i] = x[i] + alpha * P[i];
    den = num;
    num = 0;
    for (size_t i = 0; i < nv; i++)
    {
      R[i] = R[i] - alpha * AX[i];
      num += R[i] * R[i];
    }
    beta = num / (den + eps);
    for (size_t i = 0; i < nv; i++)
      P[i] = R[i] + beta * P[i];
    if (resid)
      resid[it + 1] = double(num);
  }
}

// -------------------------------------------------------------- FDK pipeline --
// include/rtkFDKConeBeamReconstructionFilter.hxx:116-157: weight -> ramp -> FDK BP
// per projection subset; the result does not depend on the subset size because
// every voxel accumulates the views in order.
template <class T>
void
FDK(const Geometry & g, const ImageInfo & vi, T * vol, const ImageInfo & si, const T * proj, const RampOptions & o,
    int nthreads, int subset = 16)
{
  const int    nu = si.size[0], nv = si.size[1], np = si.size[2];
  const size_t ppix = size_t(nu) * nv;
  for (int first = 0; first < np; first += subset)
  {
    const int      cnt = std::min(subset, np - first);
    ImageInfo      sub = si;
    sub.size[2] = cnt;
    std::vector<T> w(ppix * cnt), r(ppix * cnt);
    FDKWeight(g, sub, proj + ppix * first, w.data(), nthreads, first);
    RampFilter(sub, w.data(), r.data(), o, nthreads);
    Backproject(g, vi, vol, sub, r.data(), true, nthreads, first);
  }
}

// ---------------------------------------------------------------- phantoms --
// Quadric x^T Q x + G.x + J <= 0 with density (src/rtkQuadricShape.cxx).  The
// oracle keeps the symmetric-matrix form; RTK keeps A..J (A=Q00, B=Q11, C=Q22,
// D=2Q01, E=2Q02, F=2Q12), which is the same polynomial.
struct Quadric
{
  double A, B, C, D, E, F, G, H, I, J, density;
};

// QuadricShape::IsIntersectedByRay (src/rtkQuadricShape.cxx:43-186) restricted
// to ellipsoids (Aq > 0, no clip planes).
inline bool
QuadricIntersect(const Quadric & q, const double o[3], const double d[3], double & tin, double & tout)
{
  const double Aq = q.A * d[0] * d[0] + q.B * d[1] * d[1] + q.C * d[2] * d[2] + q.D * d[0] * d[1] +
                    q.E * d[0] * d[2] + q.F * d[1] * d[2];
  const double Bq = 2 * (q.A * o[0] * d[0] + q.B * o[1] * d[1] + q.C * o[2] * d[2]) +
                    q.D * (o[0] * d[1] + o[1] * d[0]) + q.E * (o[2] * d[0] + o[0] * d[2]) +
                    q.F * (o[1] * d[2] + o[2] * d[1]) + q.G * d[0] + q.H * d[1] + q.I * d[2];
  const double Cq = q.A * o[0] * o[0] + q.B * o[1] * o[1] + q.C * o[2] * o[2] + q.D * o[0] * o[1] +
                    q.E * o[0] * o[2] + q.F * o[1] * o[2] + q.G * o[0] + q.H * o[1] + q.I * o[2] + q.J;
  if (Aq == 0.)
    return false;
  const double disc = Bq * Bq - 4 * Aq * Cq;
  const double eps = 1e5 * std::numeric_limits<double>::epsilon();
  if (std::abs(disc) <= eps)
  {
    tin = tout = -Bq / (2 * Aq);
    return true;
  }
  if (disc < 0.)
    return false;
  tin = (-Bq - std::sqrt(disc)) / (2 * Aq);
  tout = (-Bq + std::sqrt(disc)) / (2 * Aq);
  if (tin > tout)
    std::swap(tin, tout);
  return true;
}

inline bool
QuadricInside(const Quadric & q, const double p[3])
{
  const double v = q.A * p[0] * p[0] + q.B * p[1] * p[1] + q.C * p[2] * p[2] + q.D * p[0] * p[1] +
                   q.E * p[0] * p[2] + q.F * p[1] * p[2] + q.G * p[0] + q.H * p[1] + q.I * p[2] + q.J;
  return v <= 0.;
}

// RayConvexIntersectionImageFilter chained over shapes
// (include/rtkRayConvexIntersectionImageFilter.hxx:57-84 and
// include/rtkProjectGeometricPhantomImageFilter.hxx:51-101):
// out = in + sum_shapes density * (tsup - tinf) along the unit source->pixel line.
template <class T>
void
ProjectQuadrics(const Geometry & g, const ImageInfo & si, const T * in, T * out, const std::vector<Quadric> & qs,
                int nthreads)
{
  const int nu = si.size[0], nv = si.size[1], np = si.size[2];
  ParallelFor(np, nthreads, [&](int p) {
    double s4[4];
    g.SourcePosition(p, s4);
    const M44 pm = g.ProjectionCoordinatesToFixedSystem(p) * IndexToPhysical(si);
    for (int jv = 0; jv < nv; jv++)
      for (int iu = 0; iu < nu; iu++)
      {
        double pos[3], d[3];
        for (int r = 0; r < 3; r++)
        {
          pos[r] = pm.a[r][3] + pm.a[r][0] * iu + pm.a[r][1] * jv + pm.a[r][2] * p;
          d[r] = pos[r] - s4[r];
        }
        const double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int r = 0; r < 3; r++)
          d[r] /= n;
        const size_t pix = size_t(p) * nu * nv + size_t(jv) * nu + iu;
        T            acc = in[pix];
        for (const auto & q : qs)
        {
          double a, b;
          if (QuadricIntersect(q, s4, d, a, b))
            acc = T(acc + q.density * (b - a));
        }
        out[pix] = acc;
      }
  });
}

// DrawConvexImageFilter (include/rtkDrawConvexImageFilter.hxx:41-61): voxel
// centres inside a shape receive its density.
template <class T>
void
DrawQuadrics(const ImageInfo & vi, const T * in, T * out, const std::vector<Quadric> & qs, int nthreads)
{
  const int nx = vi.size[0], ny = vi.size[1], nz = vi.size[2];
  ParallelFor(nz, nthreads, [&](int k) {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
      {
        const double p[3] = { vi.origin[0] + i * vi.spacing[0], vi.origin[1] + j * vi.spacing[1],
                              vi.origin[2] + k * vi.spacing[2] };
        const size_t idx = i + size_t(nx) * (j + size_t(ny) * k);
        T            v = in[idx];
        for (const auto & q : qs)
          if (QuadricInside(q, p))
            v = T(v + q.density);
        out[idx] = v;
      }
  });
}

} // namespace rtkref
