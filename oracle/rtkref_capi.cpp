// C entry points of the CPU oracle (test infrastructure only; see rtkref.hpp).
// Geometry is passed as n rows of the nine ThreeDCircularProjectionGeometry::
// AddProjection arguments (sid, sdd, gantry deg, projOffsetX, projOffsetY,
// outOfPlane deg, inPlane deg, sourceOffsetX, sourceOffsetY),
// src/rtkThreeDCircularProjectionGeometry.cxx:55-76.
#include "rtkref.hpp"

#include <cstdio>
#include <string>

using namespace rtkref;

extern "C"
{
  struct rtkref_image
  {
    int    size[3];
    double origin[3];
    double spacing[3];
  };
}

namespace
{
thread_local std::string g_err;

Geometry
MakeGeometry(int n, const double * g9)
{
  Geometry g;
  for (int i = 0; i < n; i++)
  {
    const double * r = g9 + 9 * i;
    g.AddProjection(r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8]);
  }
  return g;
}

ImageInfo
Info(const rtkref_image * im)
{
  ImageInfo o;
  for (int i = 0; i < 3; i++)
  {
    o.size[i] = im->size[i];
    o.origin[i] = im->origin[i];
    o.spacing[i] = im->spacing[i];
  }
  return o;
}

RampOptions
Ramp(const double * o)
{
  RampOptions r;
  if (o)
  {
    r.hann = o[0];
    r.hannY = o[1];
    r.cosine = o[2];
    r.hamming = o[3];
    r.ramlak = o[4];
    r.sheppLogan = o[5];
    r.truncation = o[6];
    r.greatestPrimeFactor = (int)o[7];
  }
  return r;
}

std::vector<Quadric>
Quads(int nq, const double * q)
{
  std::vector<Quadric> v(nq);
  for (int i = 0; i < nq; i++)
  {
    const double * r = q + 11 * i;
    v[i] = Quadric{ r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10] };
  }
  return v;
}

template <class F>
int
Guard(F && f)
{
  try
  {
    f();
    return 0;
  }
  catch (const std::exception & e)
  {
    g_err = e.what();
    return 1;
  }
}
} // namespace

extern "C"
{
  const char * rtkref_last_error() { return g_err.c_str(); }

  // Per view: 3x4 geometry matrix, source position (3), source angle, angular gap, tilt angle.
  int rtkref_geometry(int n, const double * g9, double * mat12, double * src3, double * srcAngle, double * gaps,
                      double * tilt)
  {
    return Guard([&] {
      Geometry g = MakeGeometry(n, g9);
      auto     ga = g.AngularGaps(g.sourceAngles);
      auto     ta = g.TiltAngles();
      for (int i = 0; i < n; i++)
      {
        for (int j = 0; j < 12; j++)
          mat12[12 * i + j] = g.matrices[i].a[j / 4][j % 4];
        double s[4];
        g.SourcePosition(i, s);
        for (int j = 0; j < 3; j++)
          src3[3 * i + j] = s[j];
        srcAngle[i] = g.sourceAngles[i];
        gaps[i] = ga[i];
        tilt[i] = ta[i];
      }
    });
  }

  // Index-to-index backprojection matrices (optionally FDK-normalised).
  int rtkref_bp_matrices(int n, const double * g9, const rtkref_image * vol, const rtkref_image * stack,
                         int transpose, int fdk, double * out12)
  {
    return Guard([&] {
      Geometry  g = MakeGeometry(n, g9);
      ImageInfo vi = Info(vol), si = Info(stack);
      for (int p = 0; p < n; p++)
      {
        M34 m = IndexToIndexProjectionMatrix(g, p, vi, si, transpose != 0);
        if (fdk)
        {
          double pf = m.a[2][3];
          for (int j = 0; j < 3; j++)
            pf += m.a[2][j] * (0. - vi.origin[j]) / vi.spacing[j];
          for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++)
              m.a[r][c] /= pf;
        }
        for (int j = 0; j < 12; j++)
          out12[12 * p + j] = m.a[j / 4][j % 4];
      }
    });
  }

  int rtkref_ramp_kernel(int xpad, int ypad, double spacing, const double * opts, double * out_cplx, int * kh)
  {
    return Guard([&] {
      RampOptions o = Ramp(opts);
      auto        K = RampKernelSpectrum(xpad, ypad, spacing, o, *kh);
      for (size_t i = 0; i < K.size(); i++)
      {
        out_cplx[2 * i] = K[i].real();
        out_cplx[2 * i + 1] = K[i].imag();
      }
    });
  }

  int rtkref_padded_size(int nu, int nv, const double * opts, int * xpad, int * ypad)
  {
    return Guard([&] {
      long z;
      PaddedSize(nu, nv, Ramp(opts), *xpad, *ypad, z);
    });
  }

  // Untilted coordinate at the isocentre of a physical detector coordinate.
  int rtkref_untilted(int n, const double * g9, int view, double tilted, double * out)
  {
    return Guard([&] { *out = MakeGeometry(n, g9).ToUntiltedCoordinateAtIsocenter(view, tilted); });
  }

  // Output descriptor of the displaced-detector filter: out image, offset of the
  // input's column 0 in it, corners {inferior, superior}, weighting flag.
  int rtkref_displaced_setup(int n, const double * g9, const rtkref_image * stack, int disable, int pad,
                             int offsetsSet, double minOffset, double maxOffset, rtkref_image * out, int * offset,
                             double * corners, int * weighting)
  {
    return Guard([&] {
      Geometry      g = MakeGeometry(n, g9);
      DisplacedInfo r = DisplacedDetectorSetup(g, Info(stack), disable != 0, pad != 0, offsetsSet != 0, minOffset,
                                               maxOffset);
      for (int i = 0; i < 3; i++)
      {
        out->size[i] = r.out.size[i];
        out->origin[i] = r.out.origin[i];
        out->spacing[i] = r.out.spacing[i];
      }
      *offset = r.offset;
      corners[0] = r.inferior;
      corners[1] = r.superior;
      *weighting = r.weighting;
    });
  }

#define RTKREF_TYPED(T, S)                                                                                          \
  int rtkref_fdk_weight_##S(int n, const double * g9, const rtkref_image * stack, int first, const T * in, T * out, \
                            int threads)                                                                            \
  {                                                                                                                 \
    return Guard([&] { FDKWeight<T>(MakeGeometry(n, g9), Info(stack), in, out, threads, first); });                 \
  }                                                                                                                 \
  int rtkref_parker_##S(int n, const double * g9, const rtkref_image * stack, const T * in, T * out,              \
                        double threshold, int threads, int * isShort, double * deltaFirst, int * warnView)          \
  {                                                                                                                 \
    return Guard([&] {                                                                                              \
      ParkerInfo r = ParkerShortScan<T>(MakeGeometry(n, g9), Info(stack), in, out, threshold, threads);             \
      *isShort = r.isShortScan;                                                                                     \
      deltaFirst[0] = r.delta;                                                                                      \
      deltaFirst[1] = r.firstAngle;                                                                                 \
      *warnView = r.insufficientView;                                                                               \
    });                                                                                                             \
  }                                                                                                                 \
  int rtkref_displaced_##S(int n, const double * g9, const rtkref_image * stack, const T * in, T * out,           \
                           int disable, int pad, int offsetsSet, double minOffset, double maxOffset, int threads)   \
  {                                                                                                                 \
    return Guard([&] {                                                                                              \
      Geometry      g = MakeGeometry(n, g9);                                                                        \
      ImageInfo     si = Info(stack);                                                                               \
      DisplacedInfo r = DisplacedDetectorSetup(g, si, disable != 0, pad != 0, offsetsSet != 0, minOffset,           \
                                               maxOffset);                                                          \
      DisplacedDetector<T>(g, r, si, in, out, threads);                                                             \
    });                                                                                                             \
  }                                                                                                                 \
  int rtkref_ramp_##S(const rtkref_image * stack, const double * opts, const T * in, T * out, int threads)         \
  {                                                                                                                 \
    return Guard([&] { RampFilter<T>(Info(stack), in, out, Ramp(opts), threads); });                               \
  }                                                                                                                 \
  int rtkref_backproject_##S(int n, const double * g9, const rtkref_image * vol, T * volume,                       \
                             const rtkref_image * stack, const T * proj, int fdk, int threads, int geomOffset)      \
  {                                                                                                                 \
    return Guard([&] {                                                                                              \
      Backproject<T>(MakeGeometry(n, g9), Info(vol), volume, Info(stack), proj, fdk != 0, threads, geomOffset);    \
    });                                                                                                             \
  }                                                                                                                 \
  int rtkref_joseph_##S(int n, const double * g9, const rtkref_image * stack, const T * projIn, T * projOut,       \
                        const rtkref_image * vol, const T * volume, int threads, int geomOffset,                    \
                        long long * samples)                                                                        \
  {                                                                                                                 \
    return Guard([&] {                                                                                              \
      long long c = JosephForward<T>(                                                                               \
        MakeGeometry(n, g9), Info(stack), projIn, projOut, Info(vol), volume, threads, geomOffset);                 \
      if (samples)                                                                                                  \
        *samples = c;                                                                                               \
    });                                                                                                             \
  }                                                                                                                 \
  int rtkref_joseph_back_##S(int n, const double * g9, const rtkref_image * vol, T * volume,                     \
                             const rtkref_image * stack, const T * proj, int threads, int geomOffset)               \
  {                                                                                                                 \
    return Guard(                                                                                                   \
      [&] { JosephBack<T>(MakeGeometry(n, g9), Info(vol), volume, Info(stack), proj, threads, geomOffset); });      \
  }                                                                                                                 \
  int rtkref_fdk_##S(int n, const double * g9, const rtkref_image * vol, T * volume, const rtkref_image * stack,   \
                     const T * proj, const double * opts, int threads, int subset)                                  \
  {                                                                                                                 \
    return Guard([&] {                                                                                              \
      FDK<T>(MakeGeometry(n, g9), Info(vol), volume, Info(stack), proj, Ramp(opts), threads, subset);               \
    });                                                                                                             \
  }                                                                                                                 \
  int rtkref_cg_##S(int n, const double * g9, const rtkref_image * vol, T * x, const rtkref_image * stack,         \
                    const T * proj, const T * weights, int niter, int threads, double * resid, int josephBP)        \
  {                                                                                                                 \
    return Guard([&] {                                                                                              \
      ConjugateGradient<T>(                                                                                         \
        MakeGeometry(n, g9), Info(vol), x, Info(stack), proj, weights, niter, threads, resid, josephBP != 0);       \
    });                                                                                                             \
  }                                                                                                                 \
  int rtkref_project_quadrics_##S(int n, const double * g9, const rtkref_image * stack, const T * in, T * out,     \
                                  int nq, const double * q, int threads)                                           \
  {                                                                                                                 \
    return Guard([&] { ProjectQuadrics<T>(MakeGeometry(n, g9), Info(stack), in, out, Quads(nq, q), threads); });   \
  }                                                                                                                 \
  int rtkref_draw_quadrics_##S(const rtkref_image * vol, const T * in, T * out, int nq, const double * q,          \
                               int threads)                                                                         \
  {                                                                                                                 \
    return Guard([&] { DrawQuadrics<T>(Info(vol), in, out, Quads(nq, q), threads); });                             \
  }

  RTKREF_TYPED(float, f)
  RTKREF_TYPED(double, d)
}
