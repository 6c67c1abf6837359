"""Python mirror of RTK's projector filter API on the HIP back end.

Class and method names follow RTK's C++ API (``rtk::ThreeDCircularProjectionGeometry``,
``rtk::FDKConeBeamReconstructionFilter``, ``rtk::JosephForwardProjectionImageFilter``,
``rtk::BackProjectionImageFilter``, ``rtk::ConjugateGradientConeBeamReconstructionFilter``)
so that code written against the reference reads the same.  Every filter calls the
filter-level C ABI of ``librtkhip.so`` (include/rtkhip.h); images are device-resident
torch tensors (the image data manager), indexed ``[z, y, x]`` / ``[view, v, u]``.
"""
import ctypes
import threading

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

_tls = threading.local()


def _context():
    c = getattr(_tls, "ctx", None)
    if c is None:
        c = lib().rtkhip_context_create()
        if not c:
            raise _lib.RtkHipError(_lib.RTKHIP_ERR_OUT_OF_MEMORY, "context allocation failed")
        _tls.ctx = c
    return c


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


_DTYPES = {"f32": (_lib.F32, torch.float32), "f16": (_lib.F16, torch.float16),
           "bf16": (_lib.BF16, torch.bfloat16)}


class ThreeDCircularProjectionGeometry:
    """rtk::ThreeDCircularProjectionGeometry (src/rtkThreeDCircularProjectionGeometry.cxx:55-153)."""

    def __init__(self):
        # the library is held by the instance: at interpreter exit the module's own
        # globals may already be cleared when the destructor runs
        self._dll = lib()
        self._h = self._dll.rtkhip_geometry_create()
        if not self._h:
            raise _lib.RtkHipError(_lib.RTKHIP_ERR_OUT_OF_MEMORY, "geometry allocation failed")
        self._rows = []
        self._radius = 0.
        self.collimation = []  # per view: UInf, USup, VInf, VSup (field-of-view filters only)

    def __del__(self):
        h = getattr(self, "_h", None)
        dll = getattr(self, "_dll", None)
        if h and dll is not None:
            dll.rtkhip_geometry_destroy(h)
            self._h = None

    def AddProjection(self, sid, sdd, gantryAngle, projOffsetX=0., projOffsetY=0., outOfPlaneAngle=0.,
                      inPlaneAngle=0., sourceOffsetX=0., sourceOffsetY=0.):
        row = [float(x) for x in (sid, sdd, gantryAngle, projOffsetX, projOffsetY, outOfPlaneAngle,
                                  inPlaneAngle, sourceOffsetX, sourceOffsetY)]
        check(lib().rtkhip_geometry_add_projection(self._h, *row))
        self._rows.append(row)
        self.collimation.append([np.finfo(np.float64).max] * 4)

    def SetCollimationOfLastProjection(self, uinf, usup, vinf, vsup):
        """Stored and written with the geometry; the projectors of this path do not use it
        (the reference reads it in its field-of-view and collimation-mask filters)."""
        if not self.collimation:
            raise ValueError("no projection to set the collimation of")
        self.collimation[-1] = [float(uinf), float(usup), float(vinf), float(vsup)]

    def SetRadiusCylindricalDetector(self, r):
        check(lib().rtkhip_geometry_set_radius_cylindrical_detector(self._h, float(r)))
        self._radius = float(r)

    def GetRadiusCylindricalDetector(self):
        return self._radius

    def __len__(self):
        return lib().rtkhip_geometry_size(self._h)

    def GetMatrix(self, i):
        m = (ctypes.c_double * 12)()
        check(lib().rtkhip_geometry_get(self._h, int(i), m, None, None))
        return np.array(m[:]).reshape(3, 4)

    def GetMatrices(self):
        return np.stack([self.GetMatrix(i) for i in range(len(self))]) if len(self) else np.zeros((0, 3, 4))

    def GetSourcePosition(self, i):
        s = (ctypes.c_double * 3)()
        check(lib().rtkhip_geometry_get(self._h, int(i), None, s, None))
        return np.array(s[:])

    def GetSourceAngles(self):
        out = []
        for i in range(len(self)):
            a = ctypes.c_double()
            check(lib().rtkhip_geometry_get(self._h, i, None, None, ctypes.byref(a)))
            out.append(a.value)
        return np.array(out)

    def rows(self):
        """The AddProjection arguments, one row per view (n x 9)."""
        return np.array(self._rows, dtype=np.float64).reshape(-1, 9)

    @property
    def handle(self):
        return self._h


class Image:
    """Device-resident image: a contiguous torch tensor plus ITK meta-data.

    ``size``/``origin``/``spacing``/``index`` are in ITK order (x, y, z); the tensor is
    ``[z, y, x]``.  ``index`` is the buffered-region start (ITK image index)."""

    def __init__(self, tensor, origin=(0., 0., 0.), spacing=(1., 1., 1.), index=(0, 0, 0)):
        if tensor.dim() != 3:
            raise ValueError("Image expects a 3-D tensor [z, y, x]")
        self.tensor = tensor.contiguous()
        self.origin = tuple(float(x) for x in origin)
        self.spacing = tuple(float(x) for x in spacing)
        self.index = tuple(int(x) for x in index)

    @property
    def size(self):
        return tuple(int(x) for x in reversed(self.tensor.shape))

    def descriptor(self):
        d = _lib.Image()
        for i in range(3):
            d.size[i] = self.size[i]
            d.index[i] = self.index[i]
            d.origin[i] = self.origin[i]
            d.spacing[i] = self.spacing[i]
        return d

    def like(self, tensor):
        return Image(tensor, self.origin, self.spacing, self.index)

    def numpy(self):
        return self.tensor.detach().float().cpu().numpy()

    def GetOrigin(self):
        return self.origin

    def GetSpacing(self):
        return self.spacing

    def GetLargestPossibleRegion(self):
        return self.index, self.size


def ConstantImageSource(size, spacing=(1., 1., 1.), origin=(0., 0., 0.), constant=0., device="cuda",
                        dtype=torch.float32):
    """rtk::ConstantImageSource (include/rtkConstantImageSource.h)."""
    t = torch.full((int(size[2]), int(size[1]), int(size[0])), float(constant), dtype=dtype, device=device)
    return Image(t, origin, spacing)


def _require_cuda(img, what, allow16=False):
    """Every image the library reads or writes is a contiguous CUDA tensor of float32
    pixels; only the projection stacks the backprojectors read may be float16 or
    bfloat16 (the C entry points take FP32 buffers everywhere else, so a 16-bit
    image there would be read and written at twice its size)."""
    ok = (torch.float32, torch.float16, torch.bfloat16) if allow16 else (torch.float32,)
    if not isinstance(img, Image) or not img.tensor.is_cuda or img.tensor.dtype not in ok:
        raise ValueError(f"{what}: expected a CUDA {'float32/float16/bfloat16' if allow16 else 'float32'} image")


class _Filter:
    def __init__(self):
        self._inputs = {}
        self._geometry = None
        self._output = None
        self._inplace = False

    def SetInput(self, idx, img=None):
        if img is None:
            idx, img = 0, idx
        self._inputs[int(idx)] = img

    def GetInput(self, idx=0):
        return self._inputs.get(idx)

    def SetGeometry(self, g):
        self._geometry = g

    def GetGeometry(self):
        return self._geometry

    def SetInPlace(self, v):
        self._inplace = bool(v)

    def InPlaceOn(self):
        self._inplace = True

    def InPlaceOff(self):
        self._inplace = False

    def GetOutput(self):
        return self._output

    def _geom(self):
        if self._geometry is None:
            raise _lib.RtkHipError(_lib.RTKHIP_ERR_GEOMETRY, "Geometry has not been set.")
        return self._geometry.handle

    def _out_for(self, img):
        return img.tensor if self._inplace else torch.empty_like(img.tensor)


class RampFilterOptions:
    """Options of rtk::FFTRampImageFilter (include/rtkFFTRampImageFilter.h) and of its
    base rtk::FFTProjectionsConvolutionImageFilter (zero-pad factors, greatest prime
    factor of the padded sizes, truncation correction)."""

    def __init__(self):
        self.hann = self.hannY = self.cosine = self.hamming = self.ramlak = self.shepplogan = 0.
        self.truncation = 0.
        self.zero_pad = (2, 2)
        # 13 as RTK's CUDA convolution filter (rtkCudaFFTProjectionsConvolutionImageFilter.hxx:42);
        # RTK's CPU filter uses 2 unless ITK's FFT is FFTW
        self.gpf = 13

    def SetHannCutFrequency(self, f):
        self.hann = float(f)

    def SetHannCutFrequencyY(self, f):
        self.hannY = float(f)

    def SetCosineCutFrequency(self, f):
        self.cosine = float(f)

    def SetHammingFrequency(self, f):
        self.hamming = float(f)

    def SetRamLakCutFrequency(self, f):
        self.ramlak = float(f)

    def SetSheppLoganCutFrequency(self, f):
        self.shepplogan = float(f)

    def SetTruncationCorrection(self, f):
        self.truncation = float(f)

    def SetZeroPadFactors(self, f):
        """(x, y) factors, or one value for both (itk::FixedArray<int, 2> in RTK)."""
        if hasattr(f, "__len__"):
            self.zero_pad = (int(f[0]), int(f[1]) if len(f) > 1 else int(f[0]))
        else:
            self.zero_pad = (int(f), int(f))

    def SetGreatestPrimeFactor(self, n):
        if not 2 <= int(n) <= 13:
            raise ValueError("GreatestPrimeFactor must lie in [2, 13]")
        self.gpf = int(n)

    def GetGreatestPrimeFactor(self):
        return self.gpf

    def struct(self):
        o = _lib.RampOptions()
        o.hann_cut, o.hann_cut_y, o.cosine_cut = self.hann, self.hannY, self.cosine
        o.hamming, o.ramlak_cut, o.shepp_logan_cut = self.hamming, self.ramlak, self.shepplogan
        o.truncation_correction, o.zero_pad_x, o.zero_pad_y = self.truncation, self.zero_pad[0], self.zero_pad[1]
        o.greatest_prime_factor = self.gpf
        return o


class HipFFTRampImageFilter(_Filter, RampFilterOptions):
    """rtk::FFTRampImageFilter on HIP (rtk::CudaFFTRampImageFilter's counterpart): ramp
    filtering of every projection row, with the optional 2-D kernel (HannCutFrequencyY)
    and truncation correction (include/rtkFFTRampImageFilter.hxx:47-178,
    include/rtkFFTProjectionsConvolutionImageFilter.hxx:121-304).  Input 0 = projections."""

    def __init__(self):
        _Filter.__init__(self)
        RampFilterOptions.__init__(self)

    def Update(self):
        proj = self._inputs[0]
        _require_cuda(proj, "ramp projections")
        out = self._out_for(proj)
        pd = proj.descriptor()
        ro = self.struct()
        check(lib().rtkhip_fft_ramp_filter(_context(), ctypes.byref(pd), _ptr(proj.tensor), _ptr(out), ctypes.byref(ro),
                                           _stream()))
        self._output = proj.like(out)


class HipFDKConeBeamReconstructionFilter(_Filter):
    """rtk::HipFDKConeBeamReconstructionFilter: input 0 = volume to add to, input 1 =
    projections (include/rtkFDKConeBeamReconstructionFilter.h:62-156)."""

    def __init__(self):
        super().__init__()
        self._ramp = RampFilterOptions()
        self._subset = 0  # all views at once: the whole stack fits in HBM
        self._dtype = "f32"
        self._bp_ms = 0.

    def GetRampFilter(self):
        return self._ramp

    def SetProjectionSubsetSize(self, n):
        self._subset = int(n)

    def SetStackDtype(self, name):
        if name not in _DTYPES:
            raise ValueError(f"unknown stack dtype {name}")
        self._dtype = name

    def GetBackProjectionTimeMs(self):
        return self._bp_ms

    def Update(self):
        vol, proj = self._inputs[0], self._inputs[1]
        _require_cuda(vol, "FDK volume")
        _require_cuda(proj, "FDK projections")
        out = self._out_for(vol)
        ms = ctypes.c_double()
        vd, pd = vol.descriptor(), proj.descriptor()
        ro = self._ramp.struct()
        check(lib().rtkhip_fdk_reconstruct(_context(), self._geom(), ctypes.byref(vd), _ptr(vol.tensor), _ptr(out),
                                           ctypes.byref(pd), _ptr(proj.tensor), ctypes.byref(ro), self._subset,
                                           _DTYPES[self._dtype][0], ctypes.byref(ms), _stream()))
        self._bp_ms = ms.value
        self._output = vol.like(out)


class HipForwardProjectionImageFilter(_Filter):
    """Joseph forward projector: input 0 = projections to add to, input 1 = volume
    (include/rtkForwardProjectionImageFilter.h:36-85,
    include/rtkJosephForwardProjectionImageFilter.hxx:182-427)."""

    def __init__(self):
        super().__init__()
        self._samples = 0

    def GetNumberOfSamples(self):
        return self._samples

    def Update(self, count_samples=False):
        proj, vol = self._inputs[0], self._inputs[1]
        _require_cuda(proj, "FP projections")
        _require_cuda(vol, "FP volume")
        out = self._out_for(proj)
        n = ctypes.c_longlong()
        pd, vd = proj.descriptor(), vol.descriptor()
        check(lib().rtkhip_forward_project(_context(), self._geom(), ctypes.byref(pd), _ptr(proj.tensor), _ptr(out),
                                           ctypes.byref(vd), _ptr(vol.tensor),
                                           ctypes.byref(n) if count_samples else None, _stream()))
        self._samples = n.value
        self._output = proj.like(out)


class HipBackProjectionImageFilter(_Filter):
    """Voxel-based backprojector: input 0 = volume to add to, input 1 = projections
    (include/rtkBackProjectionImageFilter.h:49-161)."""

    _fdk = 0

    def Update(self):
        vol, proj = self._inputs[0], self._inputs[1]
        _require_cuda(vol, "BP volume")
        _require_cuda(proj, "BP projections", allow16=True)
        dt = {torch.float32: _lib.F32, torch.float16: _lib.F16, torch.bfloat16: _lib.BF16}[proj.tensor.dtype]
        out = self._out_for(vol)
        vd, pd = vol.descriptor(), proj.descriptor()
        check(lib().rtkhip_back_project(_context(), self._geom(), ctypes.byref(vd), _ptr(vol.tensor), _ptr(out),
                                        ctypes.byref(pd), _ptr(proj.tensor), dt, self._fdk, _stream()))
        self._output = vol.like(out)


class HipJosephBackProjectionImageFilter(_Filter):
    """Joseph backprojector (splat transpose of the Joseph FP): input 0 = volume to
    add to, input 1 = projections (include/rtkJosephBackProjectionImageFilter.h)."""

    def Update(self):
        vol, proj = self._inputs[0], self._inputs[1]
        _require_cuda(vol, "BP volume")
        _require_cuda(proj, "BP projections", allow16=True)
        dt = {torch.float32: _lib.F32, torch.float16: _lib.F16, torch.bfloat16: _lib.BF16}[proj.tensor.dtype]
        out = self._out_for(vol)
        vd, pd = vol.descriptor(), proj.descriptor()
        check(lib().rtkhip_joseph_back_project(_context(), self._geom(), ctypes.byref(vd), _ptr(vol.tensor),
                                               _ptr(out), ctypes.byref(pd), _ptr(proj.tensor), dt, _stream()))
        self._output = vol.like(out)


class HipFDKBackProjectionImageFilter(HipBackProjectionImageFilter):
    """rtk::FDKBackProjectionImageFilter on HIP (w^2-weighted, normalised at the centre)."""

    _fdk = 1


class HipParkerShortScanImageFilter(_Filter):
    """rtk::ParkerShortScanImageFilter / CudaParkerShortScanImageFilter: weights a short
    scan's projections (include/rtkParkerShortScanImageFilter.hxx:41-209).  A scan is
    short when the largest gap between sorted gantry angles reaches the threshold."""

    def __init__(self):
        super().__init__()
        self._threshold = np.pi / 9  # reference default (rtkParkerShortScanImageFilter.h:113)
        self._short = False
        self._warn = -1

    def SetAngularGapThreshold(self, t):
        self._threshold = float(t)

    def GetAngularGapThreshold(self):
        return self._threshold

    def IsShortScan(self):
        return self._short

    def Update(self):
        import warnings
        proj = self._inputs[0]
        _require_cuda(proj, "Parker projections")
        if proj.tensor.dtype != torch.float32:
            raise ValueError("Parker weighting expects an FP32 stack")
        out = self._out_for(proj)
        pd = proj.descriptor()
        s, w = ctypes.c_int(), ctypes.c_int()
        check(lib().rtkhip_parker_short_scan(self._geom(), ctypes.byref(pd), _ptr(proj.tensor), _ptr(out),
                                             self._threshold, ctypes.byref(s), ctypes.byref(w), _stream()))
        self._short, self._warn = bool(s.value), w.value
        if self._warn >= 0:
            warnings.warn("You do not have enough data for proper Parker weighting (short scan) according to "
                          f"projection #{self._warn}")
        self._output = proj.like(out)


class HipDisplacedDetectorImageFilter(_Filter):
    """rtk::DisplacedDetectorImageFilter / CudaDisplacedDetectorImageFilter: weights
    projections of a laterally displaced detector and, by default, zero-pads them on
    the truncated side to twice their width
    (include/rtkDisplacedDetectorImageFilter.hxx:81-323)."""

    def __init__(self):
        super().__init__()
        self._disable = False
        self._pad = True
        self._offsets = None
        self._corners = (0., 0.)
        self._weighting = False

    def SetDisable(self, v):
        self._disable = bool(v)

    def GetDisable(self):
        return self._disable

    def SetPadOnTruncatedSide(self, v):
        self._pad = bool(v)

    def GetPadOnTruncatedSide(self):
        return self._pad

    def SetOffsets(self, min_offset, max_offset):
        self._offsets = (float(min_offset), float(max_offset))

    def GetInferiorCorner(self):
        return self._corners[0]

    def GetSuperiorCorner(self):
        return self._corners[1]

    def _args(self):
        mn, mx = self._offsets if self._offsets is not None else (0., 0.)
        return int(self._disable), int(self._pad), int(self._offsets is not None), mn, mx

    def OutputRegion(self, proj):
        """Output descriptor (size, index) and whether weights apply (GenerateOutputInformation)."""
        pd = proj.descriptor()
        od = _lib.Image()
        c = (ctypes.c_double * 2)()
        wt = ctypes.c_int()
        check(lib().rtkhip_displaced_detector_region(self._geom(), ctypes.byref(pd), *self._args(), ctypes.byref(od),
                                                     c, ctypes.byref(wt)))
        self._corners = (c[0], c[1])
        self._weighting = bool(wt.value)
        return od

    def Update(self):
        proj = self._inputs[0]
        _require_cuda(proj, "displaced-detector projections")
        if proj.tensor.dtype != torch.float32:
            raise ValueError("displaced-detector weighting expects an FP32 stack")
        od = self.OutputRegion(proj)
        size = tuple(od.size)
        if size == proj.size and self._inplace:
            out = proj.tensor
        else:
            out = torch.empty((size[2], size[1], size[0]), dtype=torch.float32, device=proj.tensor.device)
        pd = proj.descriptor()
        check(lib().rtkhip_displaced_detector(self._geom(), ctypes.byref(pd), _ptr(proj.tensor), ctypes.byref(od),
                                              _ptr(out), *self._args(), _stream()))
        self._output = Image(out, proj.origin, proj.spacing, tuple(od.index))


class HipConjugateGradientConeBeamReconstructionFilter(_Filter):
    """rtk::ConjugateGradientConeBeamReconstructionFilter with FP_JOSEPH and BP_VOXELBASED or BP_JOSEPH:
    input 0 = X0, input 1 = projections; optional weights
    (include/rtkConjugateGradientConeBeamReconstructionFilter.hxx:195-305)."""

    BP_VOXELBASED, BP_JOSEPH = _lib.BP_VOXELBASED, _lib.BP_JOSEPH

    def __init__(self):
        super().__init__()
        self._bp = _lib.BP_VOXELBASED  # reference default (rtkIterativeConeBeamReconstructionFilter.hxx:29-30)
        self._niter = 3  # reference default (.hxx:36)
        self._tol = 0.
        self._weights = None
        self._resid = np.zeros(0)
        self._iters = 0
        self._dtype = "f32"
        self._disable_dd = False

    def SetDisableDisplacedDetectorFilter(self, v):
        """The weights are displaced-detector weighted (without padding) unless disabled
        (rtkConjugateGradientConeBeamReconstructionFilter.hxx:41,54,222-242)."""
        self._disable_dd = bool(v)

    def GetDisableDisplacedDetectorFilter(self):
        return self._disable_dd

    def SetStackDtype(self, name):
        """Storage of the stacks the backprojector reads: "f32", "f16" or "bf16"
        (FP32/FP64 arithmetic in every case)."""
        if name not in _DTYPES:
            raise ValueError(f"unknown stack dtype {name}")
        self._dtype = name

    def SetNumberOfIterations(self, n):
        self._niter = int(n)

    def SetBackProjectionFilter(self, bp):
        """BP_VOXELBASED (RTK's default) or BP_JOSEPH (matched adjoint)."""
        if bp not in (_lib.BP_VOXELBASED, _lib.BP_JOSEPH):
            raise ValueError("unsupported backprojector")
        self._bp = int(bp)

    def SetInputWeights(self, w):
        """Projection weights W: a float32 CUDA image of exactly the projections' size
        (checked again at Update, against the projections then set)."""
        _require_cuda(w, "CG weights")
        self._weights = w

    def SetResidualTolerance(self, tol):
        """Stop once ||B - A X_k|| / ||B|| <= tol (extension; 0 = fixed iterations)."""
        self._tol = float(tol)

    def GetResiduals(self):
        return self._resid

    def GetNumberOfIterationsDone(self):
        return self._iters

    def Update(self):
        vol, proj = self._inputs[0], self._inputs[1]
        _require_cuda(vol, "CG volume")
        _require_cuda(proj, "CG projections")
        if self._weights is not None:
            _require_cuda(self._weights, "CG weights")
            if self._weights.size != proj.size:
                raise ValueError(f"CG weights: size {self._weights.size} differs from the projections' "
                                 f"{proj.size}")
        x = vol.tensor.clone()
        res = np.zeros(self._niter + 1)
        it = ctypes.c_int()
        vd, pd = vol.descriptor(), proj.descriptor()
        w = self._weights.tensor if self._weights is not None else None
        ddf = HipDisplacedDetectorImageFilter()
        ddf.SetGeometry(self._geometry)
        ddf.SetDisable(self._disable_dd)
        ddf.SetPadOnTruncatedSide(False)
        ddf.OutputRegion(proj)
        if ddf._weighting:
            ddf.SetInput(self._weights if self._weights is not None else
                         proj.like(torch.ones_like(proj.tensor, dtype=torch.float32)))
            ddf.Update()
            w = ddf.GetOutput().tensor
        check(lib().rtkhip_cg_reconstruct(_context(), self._geom(), ctypes.byref(vd), _ptr(x), ctypes.byref(pd),
                                          _ptr(proj.tensor), _ptr(w), self._niter, self._tol, self._bp,
                                          _DTYPES[self._dtype][0], res.ctypes.data_as(_lib.PD), ctypes.byref(it),
                                          _stream()))
        self._iters = it.value
        self._resid = res[: it.value + 1]
        self._output = vol.like(x)


def draw_quadrics(img, quadrics, inplace=False):
    """DrawGeometricPhantom-style voxelisation of quadric shapes (11 numbers each)."""
    _require_cuda(img, "phantom volume")
    q = np.ascontiguousarray(np.asarray(quadrics, dtype=np.float64).reshape(-1, 11))
    out = img.tensor if inplace else torch.empty_like(img.tensor)
    d = img.descriptor()
    check(lib().rtkhip_draw_quadrics(ctypes.byref(d), _ptr(img.tensor), _ptr(out), q.shape[0],
                                     q.ctypes.data_as(_lib.PD), _stream()))
    return img.like(out)


def project_quadrics(geometry, proj, quadrics, inplace=False):
    """ProjectGeometricPhantom-style analytic line integrals (FP64 chords)."""
    _require_cuda(proj, "phantom projections")
    q = np.ascontiguousarray(np.asarray(quadrics, dtype=np.float64).reshape(-1, 11))
    out = proj.tensor if inplace else torch.empty_like(proj.tensor)
    d = proj.descriptor()
    check(lib().rtkhip_project_quadrics(_context(), geometry.handle, ctypes.byref(d), _ptr(proj.tensor), _ptr(out),
                                        q.shape[0], q.ctypes.data_as(_lib.PD), _stream()))
    return proj.like(out)
