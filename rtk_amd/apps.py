"""Command-line applications on the HIP back end, with the reference's option names:
rtkfdk (applications/rtkfdk/rtkfdk.ggo, rtkfdk.py) and rtkconjugategradient
(applications/rtkconjugategradient/rtkconjugategradient.ggo, rtkconjugategradient.py),
with the shared sections for input projections (applications/rtkinputprojections_section.ggo),
the output volume (applications/rtk3Doutputimage_section.ggo) and the projectors
(applications/rtkprojectors_section.ggo).

Geometry is RTK's XML, images are MetaImage.  Options that select processing outside
this path (scanner raw-data corrections, cropping/binning, motion compensation,
regularisation, masks) are accepted with their defaults and rejected with a message
when set.  As in the reference application with --hardware cuda
(applications/rtkfdk/rtkfdk.cxx:83-113), rtkfdk runs the displaced-detector
weighting (DisplacedDetectorImageFilter, the CUDA filter's semantics) and the Parker
short-scan weighting before FDK; rtkconjugategradient weights the projections with
the displaced-detector filter without padding, unless --nodisplaced is given.
"""
import argparse
import math
import sys
import time

import numpy as np
import torch

from . import io
from .filters import (ConstantImageSource, HipConjugateGradientConeBeamReconstructionFilter,
                      HipDisplacedDetectorImageFilter, HipFDKConeBeamReconstructionFilter,
                      HipParkerShortScanImageFilter, Image)


def _add_input_projections(p):
    g = p.add_argument_group("Input projections and their pre-processing")
    g.add_argument("--path", "-p", required=True, help="Path containing projections")
    g.add_argument("--regexp", "-r", required=True, help="Regular expression to select projection files in path")
    g.add_argument("--nsort", action="store_true", help="Numeric sort for regular expression matches")
    g.add_argument("--submatch", type=int, default=0, help="Index of the submatch that will be used to sort matches")
    g.add_argument("--nolineint", action="store_true", help="Disable raw to line integral conversion, just casts to float")
    g.add_argument("--neworigin", type=float, nargs="+", help="New origin of input projections (before pre-processing)")
    g.add_argument("--newspacing", type=float, nargs="+", help="New spacing of input projections (before pre-processing)")
    # accepted with their defaults only (processing outside this path)
    for name, kw in [("newdirection", dict(type=float, nargs="+")), ("lowercrop", dict(type=int, nargs="+")),
                     ("uppercrop", dict(type=int, nargs="+")), ("binning", dict(type=int, nargs="+")),
                     ("wpc", dict(type=float, nargs="+")), ("spr", dict(type=float)), ("nonneg", dict(type=float)),
                     ("airthres", dict(type=float)), ("i0", dict(type=float)), ("idark", dict(type=float)),
                     ("component", dict(type=int)), ("radius", dict(type=int, nargs="+")),
                     ("multiplier", dict(type=float))]:
        g.add_argument("--" + name, **kw, help=argparse.SUPPRESS)


def _add_output_image(p):
    g = p.add_argument_group("Output 3D image properties")
    g.add_argument("--origin", type=float, nargs="+", help="Origin (default=centered)")
    g.add_argument("--size", "--dimension", dest="size", type=int, nargs="+", default=[256], help="Size")
    g.add_argument("--spacing", type=float, nargs="+", default=[1.], help="Spacing")
    g.add_argument("--direction", type=float, nargs="+", help=argparse.SUPPRESS)
    g.add_argument("--like", help="Copy information from this image (origin, size, spacing)")


def _reject(args, names, defaults=None):
    defaults = defaults or {}
    for n in names:
        v = getattr(args, n, None)
        if v is not None and v != defaults.get(n) and not (isinstance(v, list) and all(x == defaults.get(n, 0)
                                                                                       for x in v)):
            raise SystemExit(f"--{n}: not supported by rtkhip (processing outside the projector path)")


def read_projections(args, verbose=False):
    """The stack [view][v][u] as an FP32 device image (files sorted as the reference's
    projections reader sorts them; 2-D files are views, 3-D files stacks of views)."""
    _reject(args, ["newdirection", "lowercrop", "uppercrop", "binning", "wpc", "spr", "nonneg", "airthres",
                   "i0", "idark", "component", "radius", "multiplier"],
            {"lowercrop": 0, "uppercrop": 0, "binning": 1, "spr": 0., "idark": 0., "component": 0, "radius": 0})
    files = io.projection_files(args.path, args.regexp, args.nsort, args.submatch)
    if not files:
        raise SystemExit(f"no file of {args.path} matches {args.regexp}")
    parts, origin, spacing = [], None, None
    for f in files:
        a, o, s = io.read_image(f)
        if not np.issubdtype(a.dtype, np.floating) and not args.nolineint:
            raise SystemExit(f"{f}: integer projections need the scanner's raw-to-line-integral conversion, which "
                             "rtkhip does not implement; pass --nolineint to use the values as line integrals")
        if a.ndim == 2:
            a = a[None]
            o, s = tuple(o) + (0.,), tuple(s) + (1.,)
        if origin is None:
            origin, spacing = list(o), list(s)
        parts.append(a.astype(np.float32))
    stack = np.concatenate(parts, axis=0)
    if args.neworigin:
        origin[:len(args.neworigin)] = args.neworigin
    if args.newspacing:
        spacing[:len(args.newspacing)] = args.newspacing
    if verbose:
        print(f"Read {stack.shape[0]} projections of {stack.shape[2]}x{stack.shape[1]} from {len(files)} file(s)")
    return Image(torch.from_numpy(stack).cuda(), tuple(origin), tuple(spacing))


def output_volume(args):
    """rtk::SetConstantImageSourceFromGgo: one value applies to every axis, origin centred by default."""
    _reject(args, ["direction"])
    if args.like:
        a, o, s = io.read_image(args.like)
        return ConstantImageSource(a.shape[::-1], s, o)
    size = (args.size * 3)[:3] if len(args.size) == 1 else args.size
    spacing = (args.spacing * 3)[:3] if len(args.spacing) == 1 else args.spacing
    if args.origin:
        origin = (args.origin * 3)[:3] if len(args.origin) == 1 else args.origin
    else:
        origin = [-0.5 * (n - 1) * s for n, s in zip(size, spacing)]
    return ConstantImageSource(tuple(size), tuple(spacing), tuple(origin))


def _write(args, img, verbose):
    t = img.tensor.float().cpu().numpy()
    io.write_image(args.output, t, img.origin, img.spacing)
    if verbose:
        print(f"Wrote {args.output}")


def rtkfdk(argv=None):
    p = argparse.ArgumentParser(prog="rtkfdk", description="Reconstructs a 3D volume from a sequence of projections "
                                "[Feldkamp, David, Kress, 1984] on the HIP back end.")
    p.add_argument("--verbose", "-v", action="store_true")
    p.add_argument("--geometry", "-g", required=True, help="XML geometry file name")
    p.add_argument("--output", "-o", required=True, help="Output file name")
    p.add_argument("--hardware", choices=["hip"], default="hip", help="Hardware used for computation")
    p.add_argument("--lowmem", "-l", action="store_true", help="Accepted; the stack is always resident")
    p.add_argument("--divisions", "-d", type=int, default=1, help="Accepted; the volume is reconstructed at once")
    p.add_argument("--subsetsize", type=int, default=16, help="Number of projections processed at a time")
    p.add_argument("--nodisplaced", action="store_true", help="Disable the displaced detector filter")
    p.add_argument("--short", type=float, default=20., help="Minimum angular gap to detect a short scan (degree)")
    r = p.add_argument_group("Ramp filter")
    r.add_argument("--pad", type=float, default=0., help="Data padding parameter to correct for truncation")
    r.add_argument("--hann", type=float, default=0., help="Cut frequency for hann window in ]0,1] (0.0 disables it)")
    r.add_argument("--hannY", type=float, default=0., help="Cut frequency for hann window in ]0,1] (0.0 disables it)")
    p.add_argument("--signal", help=argparse.SUPPRESS)
    p.add_argument("--dvf", help=argparse.SUPPRESS)
    _add_input_projections(p)
    _add_output_image(p)
    args = p.parse_args(argv)
    if args.signal or args.dvf:
        raise SystemExit("--signal/--dvf: motion compensation is not part of rtkhip")
    proj = read_projections(args, args.verbose)
    g = io.read_geometry(args.geometry)
    ddf = HipDisplacedDetectorImageFilter()
    ddf.SetInput(proj)
    ddf.SetGeometry(g)
    ddf.SetDisable(args.nodisplaced)
    ddf.InPlaceOn()
    ddf.Update()
    pssf = HipParkerShortScanImageFilter()
    pssf.SetInput(ddf.GetOutput())
    pssf.SetGeometry(g)
    pssf.SetAngularGapThreshold(args.short * math.pi / 180.)
    pssf.InPlaceOn()  # the stack read from disk is not used again
    pssf.Update()
    f = HipFDKConeBeamReconstructionFilter()
    f.SetInput(0, output_volume(args))
    f.SetInput(1, pssf.GetOutput())
    f.SetGeometry(g)
    f.SetProjectionSubsetSize(args.subsetsize)
    f.GetRampFilter().SetTruncationCorrection(args.pad)
    f.GetRampFilter().SetHannCutFrequency(args.hann)
    f.GetRampFilter().SetHannCutFrequencyY(args.hannY)
    t0 = time.perf_counter()
    f.Update()
    torch.cuda.synchronize()
    if args.verbose:
        print(f"Reconstruction took {time.perf_counter() - t0:.3f} s")
    _write(args, f.GetOutput(), args.verbose)
    return 0


_FP = {"Joseph": "joseph", "HipJoseph": "joseph"}
_BP = {"VoxelBasedBackProjection": HipConjugateGradientConeBeamReconstructionFilter.BP_VOXELBASED,
       "HipVoxelBased": HipConjugateGradientConeBeamReconstructionFilter.BP_VOXELBASED,
       "Joseph": HipConjugateGradientConeBeamReconstructionFilter.BP_JOSEPH,
       "HipJoseph": HipConjugateGradientConeBeamReconstructionFilter.BP_JOSEPH}


def rtkconjugategradient(argv=None):
    p = argparse.ArgumentParser(prog="rtkconjugategradient", description="Reconstructs a 3D volume from a sequence "
                                "of projections with a conjugate gradient technique, on the HIP back end.")
    p.add_argument("--verbose", "-v", action="store_true")
    p.add_argument("--geometry", "-g", required=True, help="XML geometry file name")
    p.add_argument("--output", "-o", required=True, help="Output file name")
    p.add_argument("--niterations", "-n", type=int, default=5, help="Number of iterations")
    p.add_argument("--time", "-t", action="store_true", help="Records elapsed time during the process")
    p.add_argument("--input", "-i", help="Input volume")
    p.add_argument("--weights", "-w", help="Weights file for Weighted Least Squares (WLS)")
    p.add_argument("--regweights", help=argparse.SUPPRESS)
    p.add_argument("--gamma", type=float, default=0., help=argparse.SUPPRESS)
    p.add_argument("--tikhonov", type=float, default=0., help=argparse.SUPPRESS)
    p.add_argument("--nocudacg", action="store_true", help="Accepted; CG vector operations always run on the GPU")
    p.add_argument("--mask", "-m", help=argparse.SUPPRESS)
    p.add_argument("--nodisplaced", action="store_true", help="Disable the displaced detector filter")
    p.add_argument("--tolerance", type=float, default=0.,
                   help="Stop once ||b - A x|| / ||b|| <= tolerance (an rtkhip extension; 0 = n iterations)")
    p.add_argument("--stackdtype", choices=["f32", "f16", "bf16"], default="f32",
                   help="Storage of the stacks the backprojector reads (an rtkhip extension)")
    pr = p.add_argument_group("Projectors")
    pr.add_argument("--fp", "-f", choices=sorted(_FP), default="Joseph", help="Forward projection method")
    pr.add_argument("--bp", "-b", choices=sorted(_BP), default="VoxelBasedBackProjection",
                    help="Back projection method")
    _add_input_projections(p)
    _add_output_image(p)
    args = p.parse_args(argv)
    if args.regweights or args.mask or args.gamma or args.tikhonov:
        raise SystemExit("--regweights/--mask/--gamma/--tikhonov: regularised CG is not part of rtkhip")
    proj = read_projections(args, args.verbose)
    g = io.read_geometry(args.geometry)
    if args.input:
        a, o, s = io.read_image(args.input)
        x0 = Image(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda(), o, s)
    else:
        x0 = output_volume(args)
    f = HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, x0)
    f.SetInput(1, proj)
    f.SetGeometry(g)
    f.SetBackProjectionFilter(_BP[args.bp])
    f.SetNumberOfIterations(args.niterations)
    f.SetResidualTolerance(args.tolerance)
    f.SetStackDtype(args.stackdtype)
    f.SetDisableDisplacedDetectorFilter(args.nodisplaced)
    if args.weights:
        a, o, s = io.read_image(args.weights)
        if a.ndim == 2:
            a = a[None]
        if tuple(a.shape) != tuple(proj.tensor.shape):
            # RTK's pipeline fails on the mismatched requested region of the weights input
            raise SystemExit(f"--weights: image of size {tuple(reversed(a.shape))} does not match the projections' "
                             f"size {proj.size}")
        f.SetInputWeights(Image(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda(), proj.origin,
                                proj.spacing))
    t0 = time.perf_counter()
    f.Update()
    torch.cuda.synchronize()
    if args.time or args.verbose:
        print(f"Conjugate gradient: {f.GetNumberOfIterationsDone()} iterations in {time.perf_counter() - t0:.3f} s, "
              f"relative residual {f.GetResiduals()[-1]:.3g}")
    _write(args, f.GetOutput(), args.verbose)
    return 0
