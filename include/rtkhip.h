/*
 * rtkhip — C ABI of the MI355X (gfx950) projector hot path of RTK.
 *
 * Two levels:
 *   1. Kernel level: the data each of RTK's CUDA back-end functions receives
 *      (raw device pointers, int sizes, per-view matrices), so the reference's
 *      CUDA filter classes could bind these in place of their CUDA_* calls.
 *   2. Filter level: geometry handle + image descriptors, the work done by
 *      RTK's filter classes on top of the kernels (matrix construction, FDK
 *      subsets, ramp kernel, CG loop).  These are what rtk::Hip* filters and
 *      the Python mirror call.
 *
 * Conventions
 *   - Images are x-fastest (u-fastest for projection stacks: u, v, view), with
 *     identity direction.  `index` is the buffered region's start index (ITK).
 *   - Every function returns 0 (RTKHIP_OK) or an error code, never throws or
 *     aborts; rtkhip_last_error() returns this thread's last message.
 *   - `stream` is a hipStream_t (NULL = legacy default stream).  Kernels are
 *     stream-ordered; filter-level entry points synchronise `stream` before
 *     returning (drop-in semantics of the synchronous CUDA filters).
 *   - Device pointers may alias where noted (in-place accumulation, like the
 *     reference's InPlace filters).
 */
#ifndef RTKHIP_H
#define RTKHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
enum rtkhip_status
{
  RTKHIP_OK = 0,
  RTKHIP_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size, bad enum */
  RTKHIP_ERR_GEOMETRY = 2,         /* geometry missing / fewer entries than views */
  RTKHIP_ERR_UNSUPPORTED = 3,      /* e.g. cylindrical detector, parallel beam, > max views */
  RTKHIP_ERR_HIP = 4,              /* HIP runtime error (launch, copy, ...) */
  RTKHIP_ERR_OUT_OF_MEMORY = 5,    /* device allocation failed */
  RTKHIP_ERR_NO_DEVICE = 6         /* no usable HIP device */
};

/* Pixel storage type of a projection stack read by a backprojector. */
enum rtkhip_dtype
{
  RTKHIP_F32 = 0,
  RTKHIP_F16 = 1,
  RTKHIP_BF16 = 2
};

/* Border rule per view, chosen by the host exactly like RTK's CPU filters:
 * RTKHIP_BORDER_OPTIMIZED = OptimizedBackprojectionX/Y (sample kept iff
 * 0 <= floor(u) < Nu-1 and 0 <= floor(v) < Nv-1, plain bilinear);
 * RTKHIP_BORDER_ITK_LINEAR = generic path with itk::LinearInterpolateImageFunction
 * (inside iff -0.5 <= u < Nu-0.5, neighbours clamped). */
enum rtkhip_border
{
  RTKHIP_BORDER_OPTIMIZED = 0,
  RTKHIP_BORDER_ITK_LINEAR = 1
};

const char * rtkhip_last_error(void);
const char * rtkhip_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int rtkhip_device_count(void);

/* ================================================================ kernel level */

/* FDK voxel backprojection.  Replaces CUDA_reconstruct_conebeam
 * (include/rtkCudaFDKBackProjectionImageFilter.hcu:22-28) and the CPU hot loop
 * of include/rtkFDKBackProjectionImageFilter.hxx:232-306.
 *   proj_dim   {Nu, Nv, Nviews}; vol_dim {Nx, Ny, Nz}
 *   matrices   12 doubles per view: the 3x4 volume-index -> projection-index
 *              matrix (u row, v row, w row; no transpose, no texture 0.5 offset),
 *              already divided by the perspective factor at the rotation centre.
 *   border     Nviews entries of enum rtkhip_border (NULL = all OPTIMIZED)
 *   dev_vol_out = dev_vol_in + sum_views w^2 * bilinear(proj, u/w, v/w); in == out allowed;
 *              dev_vol_in NULL = zeros.
 *   proj_dtype enum rtkhip_dtype of dev_proj (FP32 math and accumulation in every case).
 * Any number of views (the CUDA filter's 1024-view limit does not apply). */
int rtkhip_fdk_backproject(const int proj_dim[3], const int vol_dim[3], const double * matrices,
                           const int * border, const float * dev_vol_in, float * dev_vol_out, const void * dev_proj,
                           int proj_dtype, void * stream);

/* Plain voxel-driven backprojection (no weight).  Replaces CUDA_back_project
 * (include/rtkCudaBackProjectionImageFilter.hcu:24-34) for a flat panel and the
 * CPU loops of include/rtkBackProjectionImageFilter.hxx:352-475.  Same argument
 * meaning as rtkhip_fdk_backproject, matrices not normalised. */
int rtkhip_voxel_backproject(const int proj_dim[3], const int vol_dim[3], const double * matrices,
                             const int * border, const float * dev_vol_in, float * dev_vol_out,
                             const void * dev_proj, int proj_dtype, void * stream);

/* Joseph forward projection.  Takes the data CUDA_forward_project receives
 * (include/rtkCudaForwardProjectionImageFilter.hcu:24-38) but computes RTK's CPU
 * Joseph algorithm (include/rtkJosephForwardProjectionImageFilter.hxx:182-543):
 *   proj_to_vol  12 doubles per view: rows 0-2 of the 4x4 projection-index ->
 *                volume-index matrix (third column unused)
 *   source       3 doubles per view: source position in volume index coordinates
 *   vol_spacing  3 doubles (mm per voxel; sets |stepMM|)
 *   clip         inferior/superior clip along source->pixel (RTK defaults 0, 1)
 *   dev_proj_out = dev_proj_in + integral; in == out allowed; dev_proj_in NULL = zeros.
 *   samples      optional host pointer (NULL = not counted): receives the number
 *                of Joseph samples (sum over rays of |fs-ns|+1); requesting it
 *                synchronises `stream`. */
int rtkhip_joseph_forward_project(const int proj_dim[3], const int vol_dim[3], const double * proj_to_vol,
                                  const double * source, const double vol_spacing[3], double inferior_clip,
                                  double superior_clip, const float * dev_proj_in, float * dev_proj_out,
                                  const float * dev_vol, long long * samples, void * stream);

/* Joseph backprojection: the splat transpose of rtkhip_joseph_forward_project
 * (include/rtkJosephBackProjectionImageFilter.hxx:58-412; the backprojector of
 * BP_JOSEPH, include/rtkIterativeConeBeamReconstructionFilter.hxx:35-103).  Same
 * geometry arguments as the forward projector.  dev_vol_out = dev_vol_in +
 * BP(dev_proj); dev_vol_in NULL = zeros; in == out allowed.  Accumulates with
 * FP32 atomics (results reproducible to rounding, not bitwise).
 *   proj_dtype enum rtkhip_dtype of dev_proj (FP32 math and accumulation in every case). */
int rtkhip_joseph_backproject(const int proj_dim[3], const int vol_dim[3], const double * proj_to_vol,
                              const double * source, const double vol_spacing[3], double inferior_clip,
                              double superior_clip, const float * dev_vol_in, float * dev_vol_out,
                              const void * dev_proj, int proj_dtype, void * stream);

/* FDK cosine/angular weighting.  Replaces CUDA_weight_projection
 * (src/rtkCudaFDKWeightProjectionFilter.cu:86-129) with the CPU filter's math
 * (include/rtkFDKWeightProjectionFilter.hxx:67-128).
 *   geom7  7 doubles per view: sdd, sid, projOffsetX, projOffsetY, sourceOffsetY,
 *          weight factor (angular gap * |sdd|/(2 sid^2) * |source|), tilt angle
 *   proj_origin/proj_spacing: physical position of pixel (0,0) of the slab and pitch.
 *   in == out allowed. */
int rtkhip_fdk_weight(const int proj_dim[3], const double * geom7, const double proj_origin[2],
                      const double proj_spacing[2], const float * dev_in, float * dev_out, void * stream);

/* Parker short-scan weighting.  Replaces CUDA_parker_weight
 * (src/rtkCudaParkerShortScanImageFilter.cu:98-142) with the CPU filter's FP64
 * weight math (include/rtkParkerShortScanImageFilter.hxx:162-188; weights rounded
 * to FP32 like the reference's float weight image).
 *   geom5          5 doubles per view: sdd, sourceOffsetX, projOffsetX, sid, gantry
 *                  angle (rad), as src/rtkCudaParkerShortScanImageFilter.cxx:89-97 builds
 *   delta, first_angle  the scan's Parker delta and first angle (rad)
 *   proj_origin_u  physical u of buffer column 0; proj_spacing_u its pitch
 *   in == out allowed. */
int rtkhip_parker_weight(const int proj_dim[3], const double * geom5, double delta, double first_angle,
                         double proj_origin_u, double proj_spacing_u, const float * dev_in, float * dev_out,
                         void * stream);

/* Displaced-detector weighting with zero padding.  Replaces CUDA_displaced_weight
 * (src/rtkCudaDisplacedDetectorImageFilter.cu:142-185) with the CPU filter's FP64
 * weights (include/rtkDisplacedDetectorImageFilter.hxx:250-290).
 *   in_dim         {Nu, Nv, Nviews} of the input; the output is out_nu x Nv x Nviews
 *   in_offset      output column of input column 0; columns outside the input are 0
 *   geom4          4 doubles per view: sdd, sourceOffsetX, projOffsetX, sid
 *                  (src/rtkCudaDisplacedDetectorImageFilter.cxx:112-119)
 *   theta          min(-inferior corner, superior corner); positive_case: superior +
 *                  inferior > 0 (weights rise toward +u)
 *   proj_origin_u  physical u of input column 0
 *   in == out allowed only without padding (out_nu == Nu, in_offset == 0). */
int rtkhip_displaced_detector_weight(const int in_dim[3], int out_nu, int in_offset, const double * geom4,
                                     double theta, int positive_case, double proj_origin_u, double proj_spacing_u,
                                     const float * dev_in, float * dev_out, void * stream);

/* Ramp filtering of every row of a projection slab: zero padding to `xpad`
 * (centred, optional Ohnesorge truncation mirror), R2C FFT, multiplication by
 * the kernel spectrum, C2R with 1/xpad, crop.  Replaces CUDA_padding +
 * CUDA_fft_convolution + CudaCropImageFilter
 * (src/rtkCudaFFTProjectionsConvolutionImageFilter.cu:70-181) for 1-D kernels, and
 * their row part for the 2-D kernel (then rtkhip_column_filter).
 *   xpad       Nu <= xpad <= 8192, no prime factor above 13 (RTK pads to 2 Nu made
 *              2- or 13-smooth)
 *   kernel     (xpad/2+1) complex values as interleaved doubles (re, im)
 *   trunc_next number of mirrored pixels per side (0 = none); mirror_weights has
 *              trunc_next+1 doubles (weights[b] for border distance b)
 *   weight_geom7  optional (NULL = none): fuse rtkhip_fdk_weight into the load
 *   out_dtype  enum rtkhip_dtype of dev_out (F16/BF16 halve what the BP reads) */
int rtkhip_ramp_filter(const int proj_dim[3], int xpad, const double * kernel, int trunc_next,
                       const double * mirror_weights, const double * weight_geom7, const double proj_origin[2],
                       const double proj_spacing[2], const float * dev_in, void * dev_out, int out_dtype,
                       void * stream);

/* Filtering of every detector column (along v) of a projection slab: each column is
 * zero padded at its end to `ypad` samples, transformed, multiplied by the spectrum,
 * inverse transformed with 1/ypad and cropped to Nv.  RTK's 2-D ramp kernel
 * (HannCutFrequencyY, include/rtkFFTRampImageFilter.hxx:143-176; on CUDA
 * multiply_kernel2D, src/rtkCudaFFTProjectionsConvolutionImageFilter.cu:49-68) is the
 * outer product of the 1-D ramp row and a window along v, so the 2-D convolution is
 * rtkhip_ramp_filter (row kernel) followed by this filter (the window).
 *   ypad       Nv <= ypad <= 8192, no prime factor above 13
 *   kernel     ypad complex values (re, im interleaved doubles): the full column
 *              spectrum, Hermitian (K[ypad - k] = conj K[k]) for a real kernel
 *   dev_in     FP32; dev_out of enum rtkhip_dtype out_dtype; in == out allowed for FP32 */
int rtkhip_column_filter(const int proj_dim[3], int ypad, const double * kernel, const float * dev_in,
                         void * dev_out, int out_dtype, void * stream);

/* One CG step of include/rtkConjugateGradientImageFilter.hxx:166-250 on device
 * vectors of n floats, given AP = A*P (replaces CUDA_conjugate_gradient,
 * src/rtkCudaConjugateGradientImageFilter.cu:80-122).  Scalars are FP64 and stay
 * on the device: dev_state holds 4 doubles {rr, pAp, alpha, beta}; rr must hold
 * ||R||^2 on entry (rtkhip_cg_init sets it) and holds ||R_new||^2 on exit.
 *   alpha = rr / (pAp + eps); X += alpha P; R -= alpha AP; beta = rr_new / (rr + eps);
 *   P = R + beta P.  eps = FLT_MIN (the reference's float eps). */
int rtkhip_cg_step(size_t n, float * dev_x, float * dev_r, float * dev_p, const float * dev_ap,
                   double * dev_state, void * stream);

/* rtkhip_cg_step without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * reconstruction loops, harnesses that forbid hipMalloc).  Same arguments, kernels,
 * grid and results (bit for bit) as rtkhip_cg_step; the FP64 partial sums of its two
 * reductions live in a static 8 KiB device array (1024 doubles) instead of a buffer
 * allocated by the library.  The kernels stride over the vectors with at most 1024
 * workgroups, so every n fits; a grid above the array's size is refused with
 * RTKHIP_ERR_UNSUPPORTED.  The array's device address is looked up on the first call
 * on each device; each device keeps its copy until the process ends.  All calls on one
 * device share that copy: calls that overlap on one device (other streams, other host
 * threads) race and must be serialised by the caller. */
int rtkhip_cg_step_noalloc(size_t n, float * dev_x, float * dev_r, float * dev_p, const float * dev_ap,
                           double * dev_state, void * stream);

/* R = B - AX; P = R; dev_state[0] = ||R||^2 (FP64).  dev_r may alias dev_ax. */
int rtkhip_cg_init(size_t n, const float * dev_b, const float * dev_ax, float * dev_r, float * dev_p,
                   double * dev_state, void * stream);

/* FP64 dot product of two device vectors (deterministic tree order); result written
 * to *dev_result (device double). */
int rtkhip_dot(size_t n, const float * dev_a, const float * dev_b, double * dev_result, void * stream);

/* dev_out = dev_a * dev_b element-wise over n floats (aliasing allowed): the projection
 * weighting step of the CG operator, done by an itk::MultiplyImageFilter in RTK
 * (include/rtkReconstructionConjugateGradientOperator.hxx:216-218). */
int rtkhip_multiply(size_t n, const float * dev_a, const float * dev_b, float * dev_out, void * stream);

/* ------------------------------------------------------------ image buffers */
/* Device-resident image buffer with an optional host mirror: the image data manager
 * behind RTK's CUDA filters (itk::CudaImage / CudaDataManager of ITKCudaCommon, whose
 * device pointers the filters take with GetCudaDataManager()->GetGPUBufferPointer(),
 * src/rtkCudaFDKBackProjectionImageFilter.cxx:71-73).  The buffer owns one device
 * allocation of `bytes`; the caller may attach host memory of the same size (a CPU
 * image's pixels, still owned by the caller).  Each access makes the accessed side
 * current first (one copy if the other side was written since) and a write access
 * marks the other side stale, so images passed between filters stay in HBM.  Calls on
 * one buffer are serialised; transfers are ordered on `stream`. */
typedef struct rtkhip_buffer rtkhip_buffer;
rtkhip_buffer * rtkhip_buffer_create(size_t bytes); /* NULL on failure (rtkhip_last_error) */
void            rtkhip_buffer_destroy(rtkhip_buffer * b);
/* Attaches host memory; its contents become the image (the device copy is stale).
 * NULL detaches (the device copy stays the image). */
int rtkhip_buffer_set_host(rtkhip_buffer * b, void * host);
/* Device address of the image, uploaded first if the host copy is newer; write != 0
 * marks the host copy stale. */
int rtkhip_buffer_device(rtkhip_buffer * b, int write, void ** dev, void * stream);
/* The attached host memory, downloaded first (and `stream` synchronised) if the device
 * copy is newer; write != 0 marks the device copy stale. */
int rtkhip_buffer_host(rtkhip_buffer * b, int write, void ** host, void * stream);
/* Host-to-device and device-to-host copies made so far. */
int rtkhip_buffer_transfers(const rtkhip_buffer * b, long long * uploads, long long * downloads);

/* ================================================================ filter level */

typedef struct rtkhip_image
{
  int    size[3];
  int    index[3];
  double origin[3];
  double spacing[3];
} rtkhip_image;

/* Ramp options of rtk::FFTRampImageFilter (include/rtkFFTRampImageFilter.h) and its
 * base rtk::FFTProjectionsConvolutionImageFilter (include/rtkFFTProjectionsConvolutionImageFilter.h).
 * Only one window applies, in the order hann, cosine, hamming, ramlak, shepp_logan. */
typedef struct rtkhip_ramp_options
{
  double hann_cut;
  double hann_cut_y; /* > 0: the 2-D kernel (ramp row x Hann window along v) */
  double cosine_cut;
  double hamming;
  double ramlak_cut;
  double shepp_logan_cut;
  double truncation_correction;
  int    zero_pad_x; /* ZeroPadFactors[0]; <= 0 means RTK's default 2 */
  int    zero_pad_y; /* ZeroPadFactors[1] (2-D kernel only); <= 0 means 2 */
  /* GreatestPrimeFactor of the padded sizes, 2..13; 0 means 13, the value of RTK's CUDA
   * convolution filter (include/rtkCudaFFTProjectionsConvolutionImageFilter.hxx:42).
   * RTK's CPU filter uses 2 unless ITK's FFT is FFTW. */
  int greatest_prime_factor;
} rtkhip_ramp_options;

typedef struct rtkhip_geometry rtkhip_geometry;
typedef struct rtkhip_context  rtkhip_context;

/* rtk::ThreeDCircularProjectionGeometry
 * (src/rtkThreeDCircularProjectionGeometry.cxx:55-153). Angles in degrees. */
rtkhip_geometry * rtkhip_geometry_create(void);
void              rtkhip_geometry_destroy(rtkhip_geometry * g);
int rtkhip_geometry_add_projection(rtkhip_geometry * g, double sid, double sdd, double gantry_angle,
                                   double proj_offset_x, double proj_offset_y, double out_of_plane_angle,
                                   double in_plane_angle, double source_offset_x, double source_offset_y);
int rtkhip_geometry_set_radius_cylindrical_detector(rtkhip_geometry * g, double radius);
int rtkhip_geometry_size(const rtkhip_geometry * g);
/* 3x4 matrix (row-major), source position (3) and source angle of view i. */
int rtkhip_geometry_get(const rtkhip_geometry * g, int i, double matrix[12], double source[3], double * source_angle);

/* A context owns cached device workspaces (matrix uploads, ramp kernels, CG
 * vectors).  Use one context per host thread. */
rtkhip_context * rtkhip_context_create(void);
void             rtkhip_context_destroy(rtkhip_context * c);

/* rtk::HipFDKConeBeamReconstructionFilter: weighting + ramp + FDK BP over
 * projection subsets (include/rtkFDKConeBeamReconstructionFilter.hxx:116-157).
 * vol_out = vol_in + FDK(projections).  subset_size <= 0 means all views at once.
 * stack_dtype: storage of the filtered projections read by the BP (enum rtkhip_dtype).
 * bp_ms (optional): milliseconds spent in the backprojection kernels. */
int rtkhip_fdk_reconstruct(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * vol,
                           const float * dev_vol_in, float * dev_vol_out, const rtkhip_image * stack,
                           const float * dev_proj, const rtkhip_ramp_options * ramp, int subset_size,
                           int stack_dtype, double * bp_ms, void * stream);

/* rtk::HipFFTRampImageFilter (rtk::FFTRampImageFilter, include/rtkFFTRampImageFilter.hxx:47-178,
 * on the padding of include/rtkFFTProjectionsConvolutionImageFilter.hxx:203-304):
 * dev_out = ramp-filtered dev_in (FP32 both; in == out allowed).  `ramp` NULL = defaults. */
int rtkhip_fft_ramp_filter(rtkhip_context * c, const rtkhip_image * stack, const float * dev_in, float * dev_out,
                           const rtkhip_ramp_options * ramp, void * stream);

/* rtk::HipForwardProjectionImageFilter (Joseph).  proj_out = proj_in + FP(vol). */
int rtkhip_forward_project(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * stack,
                           const float * dev_proj_in, float * dev_proj_out, const rtkhip_image * vol,
                           const float * dev_vol, long long * samples, void * stream);

/* rtk::HipBackProjectionImageFilter (voxel-based, fdk = 0) and the FDK
 * backprojector (fdk = 1).  vol_out = vol_in + BP(proj). */
int rtkhip_back_project(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * vol,
                        const float * dev_vol_in, float * dev_vol_out, const rtkhip_image * stack,
                        const void * dev_proj, int proj_dtype, int fdk, void * stream);

/* Joseph backprojector at filter level (rtk::JosephBackProjectionImageFilter).
 * vol_out = vol_in + BP_Joseph(proj); proj_dtype as for rtkhip_back_project. */
int rtkhip_joseph_back_project(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * vol,
                               const float * dev_vol_in, float * dev_vol_out, const rtkhip_image * stack,
                               const void * dev_proj, int proj_dtype, void * stream);

/* Backprojector of the CG operator (RTK's BackProjectionType subset). */
enum rtkhip_backprojector
{
  RTKHIP_BP_VOXELBASED = 0, /* RTK's default: A is not symmetric */
  RTKHIP_BP_JOSEPH = 1      /* matched adjoint of the Joseph FP: A symmetric */
};

/* rtk::HipConjugateGradientConeBeamReconstructionFilter with the Joseph FP and the
 * chosen backprojector (include/rtkConjugateGradientConeBeamReconstructionFilter.hxx:195-305):
 * solves BP(W FP(x)) = BP(W p) from dev_x (X0 on entry, result on exit).
 *   dev_weights  optional projection weights (NULL = ones)
 *   niter        maximum iterations
 *   rel_tol      stop once ||B - A X_k|| / ||B|| <= rel_tol (0 = run niter); the
 *                norm is the CG recursion's R_k, FP64
 *   stack_dtype  storage of the stacks the backprojector reads (enum rtkhip_dtype): the
 *                weighted measured projections and each iteration's W FP(P) are
 *                rounded to FP16/BF16 once, all arithmetic stays FP32/FP64; FP16
 *                stacks are scaled by a power of two that maps their max |value|
 *                into (2^13, 2^14] (W FP(P) grows with the unnormalised direction P),
 *                undone exactly after the backprojection
 *   resid        optional host array of niter+1 doubles: ||R_k|| / ||B||
 *   iterations   optional: iterations done */
int rtkhip_cg_reconstruct(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * vol, float * dev_x,
                          const rtkhip_image * stack, const float * dev_proj, const float * dev_weights, int niter,
                          double rel_tol, int backprojector, int stack_dtype, double * resid, int * iterations,
                          void * stream);

/* Phantom generators (input generation for tests and benchmarks).  Each shape is
 * 11 doubles: quadric coefficients A..J of A x^2 + B y^2 + C z^2 + D xy + E xz +
 * F yz + G x + H y + I z + J <= 0 (src/rtkQuadricShape.cxx:30-41) and a density.
 * draw: DrawConvexImageFilter (include/rtkDrawConvexImageFilter.hxx:41-61);
 * project: RayConvexIntersectionImageFilter chain (FP64 analytic chords,
 * include/rtkRayConvexIntersectionImageFilter.hxx:57-84). */
/* rtk::ParkerShortScanImageFilter (include/rtkParkerShortScanImageFilter.hxx:41-209;
 * CUDA: src/rtkCudaParkerShortScanImageFilter.cxx:28-116).  The scan is short when the
 * largest gap between sorted gantry angles is >= threshold (rad; the reference's default
 * is pi/9).  Not short: dev_out = dev_in (copied unless in == out).
 *   is_short_scan  optional out
 *   warn_view      optional out: first view whose half fan angle exceeds delta (the
 *                  reference warns "not enough data"), -1 if none.  in == out allowed. */
int rtkhip_parker_short_scan(const rtkhip_geometry * g, const rtkhip_image * stack, const float * dev_in,
                             float * dev_out, double threshold, int * is_short_scan, int * warn_view, void * stream);

/* Output region of rtk::DisplacedDetectorImageFilter
 * (include/rtkDisplacedDetectorImageFilter.hxx:81-173): twice as wide, extended on the
 * truncated side, when the detector is displaced and pad_on_truncated_side is set;
 * otherwise the input's.  offsets_set: use min/max_offset instead of the geometry's
 * projection offsets (SetOffsets).  corners (optional) receives the inferior and
 * superior corners, weighting (optional) whether weights apply.  Errors as the
 * reference: RTKHIP_ERR_UNSUPPORTED for cylindrical or parallel geometry and for
 * displacements over 50% of the panel. */
int rtkhip_displaced_detector_region(const rtkhip_geometry * g, const rtkhip_image * in, int disable,
                                     int pad_on_truncated_side, int offsets_set, double min_offset,
                                     double max_offset, rtkhip_image * out, double corners[2], int * weighting);

/* rtk::DisplacedDetectorImageFilter (.hxx:175-323; CUDA:
 * src/rtkCudaDisplacedDetectorImageFilter.cxx:28-137).  `out` must be the region
 * rtkhip_displaced_detector_region returns for the same arguments.  Not displaced or
 * disabled: dev_out = dev_in (copied unless in == out). */
int rtkhip_displaced_detector(const rtkhip_geometry * g, const rtkhip_image * in, const float * dev_in,
                              const rtkhip_image * out, float * dev_out, int disable, int pad_on_truncated_side,
                              int offsets_set, double min_offset, double max_offset, void * stream);

int rtkhip_draw_quadrics(const rtkhip_image * vol, const float * dev_in, float * dev_out, int nshapes,
                         const double * shapes, void * stream);
int rtkhip_project_quadrics(rtkhip_context * c, const rtkhip_geometry * g, const rtkhip_image * stack,
                            const float * dev_in, float * dev_out, int nshapes, const double * shapes,
                            void * stream);

/* ============================================================ learned warm start */
/* A 3D residual CNN mapping an FDK volume to a CG warm start.  No counterpart in
 * the reference; it feeds the warm-start input of CG (RTK resumes CG from any
 * input volume, applications/rtkconjugategradient/rtkconjugategradient.cxx:61-77).
 * Fixed architecture: six 3x3x3 convolutions, 32 channels, ReLU after the first
 * five, y = x + conv6(...) / scale.  BF16 weights/activations (MFMA), FP32
 * accumulation, FP32 master parameters and Adam state.  Parameter vector layout:
 * W1[27][32], b1[32], then for layers 2..5 W[27][32 out][32 in], b[32], then
 * W6[27][32], b6[1]; tap t = (dz*3 + dy)*3 + dx, offset (dx-1, dy-1, dz-1). */
typedef struct rtkhip_cnn rtkhip_cnn;
int          rtkhip_cnn_num_params(void);
rtkhip_cnn * rtkhip_cnn_create(void);
void         rtkhip_cnn_destroy(rtkhip_cnn * net);
/* Upload parameters (host, FP32) and reset the Adam moments. */
int rtkhip_cnn_set_params(rtkhip_cnn * net, const float * host_params, void * stream);
/* Download parameters and/or the last gradients (either pointer may be NULL). */
int rtkhip_cnn_get_params(rtkhip_cnn * net, float * host_params, float * host_grads, void * stream);
/* y = net(x) for `batch` volumes of dim {X, Y, Z} (x-fastest FP32, batch slowest).
 * Keeps the activations (and a reference to dev_in) for rtkhip_cnn_backward. */
int rtkhip_cnn_forward(rtkhip_cnn * net, int batch, const int dim[3], const float * dev_in, float * dev_out,
                       float scale, void * stream);
/* Loss = mean((scale (y - target))^2) of the last forward and its parameter
 * gradient (kept in the handle); *loss (optional) receives the loss. */
int rtkhip_cnn_backward(rtkhip_cnn * net, const float * dev_out, const float * dev_target, double * loss,
                        void * stream);
/* One Adam update with bias correction (step >= 1) from the last gradient. */
int rtkhip_cnn_adam_step(rtkhip_cnn * net, float lr, float beta1, float beta2, float eps, int step, void * stream);

#ifdef __cplusplus
}
#endif

#endif /* RTKHIP_H */
