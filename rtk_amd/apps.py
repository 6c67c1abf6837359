"""Command-line applications on the HIP back end, with the reference's option names:
rtkfdk (applications/rtkfdk/rtkfdk.ggo, rtkfdk.py) and rtkconjugategradient
(applications/rtkconjugategradient/rtkconjugategradient.ggo, rtkconjugategradient.py),
with the shared sections for input projections (applications/rtkinputprojections_section.ggo),
the output volume (applications/rtk3Doutputimage_section.ggo) and the projectors
(applications/rtkprojectors_section.ggo).

Geometry is RTK's XML, images are MetaImage.  Options that select processing outside
this path (scanner raw-data corrections, cropping/binning, motion compensation,
regularisation, masks) are accepted with their defaults and rejected with a message
when set.  As in the reference application, the displaced-detector and Parker
short-scan weightings run before FDK; they are identities for a full scan with a
centred detector, which is what this build reconstructs: any other acquisition is
rejected (see check_full_scan_centred), or run without that weighting when
--nodisplaced is given, as the reference does.
"""
import argparse
import math
import sys
import time

import numpy as np
import torch

from . import io
from .filters import (ConstantImageSource, HipConjugateGradientConeBeamReconstructionFilter,
                      HipFDKConeBeamReconstructionFilter, Image)


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
            raise SystemExit(f"--{n}: not supported by this build (processing outside the projector path)")


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
                             "this build does not implement; pass --nolineint to use the values as line integrals")
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


def check_full_scan_centred(geometry, stack, short_deg=20., nodisplaced=False):
    """The decisions of rtk::ParkerShortScanImageFilter (max gap between consecutive
    gantry angles below the short-scan threshold: no weighting,
    include/rtkParkerShortScanImageFilter.hxx:46-55) and of
    rtk::DisplacedDetectorImageFilter (untilted detector corners at the isocentre
    symmetric within 10%: no weighting, include/rtkDisplacedDetectorImageFilter.hxx:
    120-160, ToUntiltedCoordinateAtIsocenter at src/rtkThreeDCircularProjectionGeometry.cxx:
    641-663).  Raises if either weighting would change the data."""
    rows = geometry.rows()
    if rows.size == 0:
        raise SystemExit("empty geometry")
    if np.any(rows[:, 1] == 0.):
        raise SystemExit("parallel geometry is not supported by this build")
    ang = np.sort(np.mod(np.radians(rows[:, 2]), 2 * math.pi))
    gaps = np.diff(np.concatenate([ang, [ang[0] + 2 * math.pi]]))
    if gaps.max() >= math.radians(short_deg):
        raise SystemExit(f"short scan (largest angular gap {math.degrees(gaps.max()):.1f} deg >= --short "
                         f"{short_deg}): Parker weighting is not implemented in this build")
    if nodisplaced:
        return
    if geometry.GetRadiusCylindricalDetector() != 0.:
        raise SystemExit("Displaced detector cannot handle cylindrical detector; use --nodisplaced")
    nu = stack.tensor.shape[2]
    inf = stack.origin[0]
    sup = inf + stack.spacing[0] * (nu - 1)
    if stack.spacing[0] < 0:
        inf, sup = sup, inf

    def untilt(r, c):
        sid, sdd, sx, px = r[0], r[1], r[7], r[3]
        sidu = math.sqrt(sid * sid + sx * sx)
        l = (c + px - sx) * sid / sdd + sx
        return l * abs(sid) / (sidu - l * sx / sidu)

    inf = max(untilt(r, inf) for r in rows)
    sup = min(untilt(r, sup) for r in rows)
    if inf > 0. or sup < 0.:
        raise SystemExit(f"Cannot account for detector displacement larger than 50% of panel size "
                         f"(corner inf={inf}, sup={sup})")
    if abs(inf + sup) >= 0.1 * abs(sup - inf):
        raise SystemExit("displaced detector: its weighting is not implemented in this build; "
                         "pass --nodisplaced to reconstruct without it")


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
        raise SystemExit("--signal/--dvf: motion compensation is not part of this build")
    proj = read_projections(args, args.verbose)
    g = io.read_geometry(args.geometry)
    check_full_scan_centred(g, proj, args.short, args.nodisplaced)
    f = HipFDKConeBeamReconstructionFilter()
    f.SetInput(0, output_volume(args))
    f.SetInput(1, proj)
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
                   help="Stop once ||b - A x|| / ||b|| <= tolerance (this build's extension; 0 = n iterations)")
    p.add_argument("--stackdtype", choices=["f32", "f16", "bf16"], default="f32",
                   help="Storage of the stacks the backprojector reads (this build's extension)")
    pr = p.add_argument_group("Projectors")
    pr.add_argument("--fp", "-f", choices=sorted(_FP), default="Joseph", help="Forward projection method")
    pr.add_argument("--bp", "-b", choices=sorted(_BP), default="VoxelBasedBackProjection",
                    help="Back projection method")
    _add_input_projections(p)
    _add_output_image(p)
    args = p.parse_args(argv)
    if args.regweights or args.mask or args.gamma or args.tikhonov:
        raise SystemExit("--regweights/--mask/--gamma/--tikhonov: regularised CG is not part of this build")
    proj = read_projections(args, args.verbose)
    g = io.read_geometry(args.geometry)
    check_full_scan_centred(g, proj, 180., args.nodisplaced)
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
    if args.weights:
        a, o, s = io.read_image(args.weights)
        if a.ndim == 2:
            a = a[None]
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
