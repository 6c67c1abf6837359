"""rtk_amd — MI355X (gfx950) HIP implementation of RTK's projector hot path.

Joseph forward projection, voxel-driven plain/FDK backprojection, FDK weighting
with an FFT ramp filter, and conjugate-gradient reconstruction with FP64
scalars, behind a C ABI (include/rtkhip.h) and this Python mirror of RTK's
filter API.  The native library ``librtkhip.so`` must be built
(``make -C rtk_amd``); there is no CPU fallback.
"""
from ._lib import RtkHipError, lib  # noqa: F401
from .filters import (  # noqa: F401
    ConstantImageSource,
    HipBackProjectionImageFilter,
    HipConjugateGradientConeBeamReconstructionFilter,
    HipDisplacedDetectorImageFilter,
    HipFDKBackProjectionImageFilter,
    HipFDKConeBeamReconstructionFilter,
    HipFFTRampImageFilter,
    HipForwardProjectionImageFilter,
    HipJosephBackProjectionImageFilter,
    HipParkerShortScanImageFilter,
    Image,
    RampFilterOptions,
    ThreeDCircularProjectionGeometry,
    draw_quadrics,
    project_quadrics,
)
from . import phantoms  # noqa: F401

__version__ = "0.1.0"
