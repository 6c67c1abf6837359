"""Edge cases of the projectors against the float oracle: non-finite inputs (a NaN or
infinite pixel or voxel reaches exactly the outputs whose taps read it, as in RTK's
CPU filters, including through the Joseph backprojector's per-ray path for tiles
holding a non-finite value; the voxel backprojectors return NaN where RTK returns
+-Inf, see _compare), single-row detectors and single-view stacks, and zero-view
calls at the kernel level (nothing written)."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import hip_geometry, rel_l2, to_device

pytestmark = pytest.mark.gpu

ROWS = np.array([[700., 1100., 360. * i / 12 + 3, 2. - i * 0.3, 1., 0, 0, 3., -2.] for i in range(12)])
VSIZE, VORG, VSP = (40, 36, 32), (-60., -52., -47.), (3., 3., 3.)


def _stack_img(nu, nv, n):
    return (nu, nv, n), (-(nu - 1) * 2.5, -(nv - 1) * 2.5 if nv > 1 else 0., 0.), (5., 5., 1.)


def _compare(a, b, tol, same_kind=True):
    """a (HIP) and b (oracle) are non-finite at the same voxels and agree elsewhere.
    same_kind: NaN and +-Inf at the same voxels too.  The voxel backprojectors
    interpolate as a + f (b - a), which turns an infinite tap into NaN where RTK's
    (1 - f) a + f b keeps +-Inf; NaN inputs propagate identically."""
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb), f"non-finite sets differ: {(~fa).sum()} vs {(~fb).sum()}"
    assert np.all(np.isnan(a[np.isnan(b)]))
    if same_kind:
        assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[np.isinf(b)], b[np.isinf(b)])
    assert rel_l2(np.where(fa, a, 0.), np.where(fb, b, 0.)) <= tol


@pytest.mark.parametrize("which", ["voxel", "fdk", "joseph"])
def test_nan_pixel_backprojection(oracle, hip, which):
    n = len(ROWS)
    size, org, sp = _stack_img(48, 40, n)
    rng = np.random.default_rng(21)
    p = rng.normal(size=(n, 40, 48)).astype(np.float32)
    p[3, 20, 24] = np.nan
    p[7, 5, 30] = np.inf
    g = hip_geometry(hip, ROWS)
    ov, os_ = oracle.image(VSIZE, VORG, VSP), oracle.image(size, org, sp)
    z = np.zeros(VSIZE[::-1], np.float32)
    if which == "joseph":
        ref = oracle.joseph_back(ROWS, ov, z, os_, p)
        f = hip.HipJosephBackProjectionImageFilter()
    else:
        ref = oracle.backproject(ROWS, ov, z, os_, p, fdk=which == "fdk")
        f = hip.HipFDKBackProjectionImageFilter() if which == "fdk" else hip.HipBackProjectionImageFilter()
    f.SetInput(0, hip.ConstantImageSource(VSIZE, VSP, VORG, 0.))
    f.SetInput(1, to_device(hip, p, org, sp))
    f.SetGeometry(g)
    f.Update()
    out = f.GetOutput().numpy()
    assert np.isnan(ref).any() and np.isinf(ref).any()
    _compare(out, ref, 1e-5, same_kind=which == "joseph")


def test_nan_voxel_forward_projection(oracle, hip):
    n = len(ROWS)
    size, org, sp = _stack_img(48, 40, n)
    rng = np.random.default_rng(22)
    v = rng.normal(size=VSIZE[::-1]).astype(np.float32)
    v[16, 18, 20] = np.nan
    ref = oracle.joseph(ROWS, oracle.image(size, org, sp), np.zeros((n, 40, 48), np.float32),
                        oracle.image(VSIZE, VORG, VSP), v)
    f = hip.HipForwardProjectionImageFilter()
    f.SetInput(0, hip.ConstantImageSource(size, sp, org, 0.))
    f.SetInput(1, to_device(hip, v, VORG, VSP))
    f.SetGeometry(hip_geometry(hip, ROWS))
    f.Update()
    assert np.isnan(ref).any()
    _compare(f.GetOutput().numpy(), ref, 1e-5)


@pytest.mark.parametrize("nv,n", [(1, 12), (40, 1), (1, 1)])
def test_single_row_and_single_view(oracle, hip, nv, n):
    rows = ROWS[:n]
    size, org, sp = _stack_img(48, nv, n)
    rng = np.random.default_rng(23)
    p = rng.normal(size=(n, nv, 48)).astype(np.float32)
    v = rng.normal(size=VSIZE[::-1]).astype(np.float32)
    g = hip_geometry(hip, rows)
    ov, os_ = oracle.image(VSIZE, VORG, VSP), oracle.image(size, org, sp)
    z = np.zeros(VSIZE[::-1], np.float32)
    for name, cls, ref in [
            ("voxel", hip.HipBackProjectionImageFilter, oracle.backproject(rows, ov, z, os_, p)),
            ("fdk", hip.HipFDKBackProjectionImageFilter, oracle.backproject(rows, ov, z, os_, p, fdk=True)),
            ("joseph_bp", hip.HipJosephBackProjectionImageFilter, oracle.joseph_back(rows, ov, z, os_, p))]:
        f = cls()
        f.SetInput(0, hip.ConstantImageSource(VSIZE, VSP, VORG, 0.))
        f.SetInput(1, to_device(hip, p, org, sp))
        f.SetGeometry(g)
        f.Update()
        out = f.GetOutput().numpy()
        e = rel_l2(out, ref) if np.abs(ref).max() > 0 else float(np.abs(out).max())
        assert e <= 1e-5, f"{name} nv={nv} n={n}: {e:.3g}"
    ref = oracle.joseph(rows, os_, np.zeros((n, nv, 48), np.float32), ov, v)
    f = hip.HipForwardProjectionImageFilter()
    f.SetInput(0, hip.ConstantImageSource(size, sp, org, 0.))
    f.SetInput(1, to_device(hip, v, VORG, VSP))
    f.SetGeometry(g)
    f.Update()
    assert rel_l2(f.GetOutput().numpy(), ref) <= 1e-5


def test_zero_views_at_kernel_level(hip):
    """np = 0: every kernel-level entry point returns OK and writes nothing."""
    from rtk_amd import _lib
    L = _lib.lib()
    vol = torch.full((8, 8, 8), 7., device="cuda")
    proj = torch.full((1, 8, 8), 3., device="cuda")
    m, mp = _lib.dbl(np.zeros(12))
    s, sp = _lib.dbl(np.zeros(3))
    vs, vsp = _lib.dbl(np.ones(3))
    g5, g5p = _lib.dbl(np.zeros(5))
    g4, g4p = _lib.dbl(np.zeros(4))
    pv, vv = ctypes.c_void_p(proj.data_ptr()), ctypes.c_void_p(vol.data_ptr())
    pd, vd = _lib.int3((8, 8, 0)), _lib.int3((8, 8, 8))
    _lib.check(L.rtkhip_fdk_backproject(pd, vd, mp, None, vv, vv, pv, _lib.F32, None))
    _lib.check(L.rtkhip_voxel_backproject(pd, vd, mp, None, vv, vv, pv, _lib.F32, None))
    _lib.check(L.rtkhip_joseph_forward_project(pd, vd, mp, sp, vsp, 0., 1., pv, pv, vv, None, None))
    _lib.check(L.rtkhip_joseph_backproject(pd, vd, mp, sp, vsp, 0., 1., vv, vv, pv, _lib.F32, None))
    _lib.check(L.rtkhip_parker_weight(pd, g5p, 0.1, 0., -10., 1., pv, pv, None))
    # a padded output may not alias the input, even when there is nothing to write
    _lib.check(L.rtkhip_displaced_detector_weight(pd, 16, 8, g4p, 5., 1, -10., 1., pv, vv, None))
    torch.cuda.synchronize()
    assert torch.all(vol == 7.) and torch.all(proj == 3.)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_filters_reject_16bit_images_where_the_library_takes_fp32(hip, dtype):
    """Volumes, and the stacks of the FDK, the forward projector, CG and the weighting
    filters, are FP32 buffers in the C ABI; only the backprojectors read 16-bit stacks.
    A 16-bit image elsewhere raises ValueError before any library call."""
    rows = ROWS[:4]
    g = hip_geometry(hip, rows)
    size, org, sp = _stack_img(24, 20, 4)
    p32 = hip.ConstantImageSource(size, sp, org, constant=1.)
    v32 = hip.ConstantImageSource(VSIZE, VSP, VORG)
    p16, v16 = p32.like(p32.tensor.to(dtype)), v32.like(v32.tensor.to(dtype))
    cases = [(hip.HipFDKConeBeamReconstructionFilter, (v16, p32)), (hip.HipFDKConeBeamReconstructionFilter, (v32, p16)),
             (hip.HipForwardProjectionImageFilter, (p16, v32)), (hip.HipForwardProjectionImageFilter, (p32, v16)),
             (hip.HipBackProjectionImageFilter, (v16, p32)), (hip.HipJosephBackProjectionImageFilter, (v16, p32)),
             (hip.HipConjugateGradientConeBeamReconstructionFilter, (v16, p32)),
             (hip.HipConjugateGradientConeBeamReconstructionFilter, (v32, p16))]
    for cls, (a, b) in cases:
        f = cls()
        f.SetInput(0, a)
        f.SetInput(1, b)
        f.SetGeometry(g)
        with pytest.raises(ValueError):
            f.Update()
    for cls in (hip.HipParkerShortScanImageFilter, hip.HipDisplacedDetectorImageFilter):
        f = cls()
        f.SetInput(p16)
        f.SetGeometry(g)
        with pytest.raises(ValueError):
            f.Update()
    with pytest.raises(ValueError):
        hip.draw_quadrics(v16, hip.phantoms.shepp_logan(20.))
    with pytest.raises(ValueError):
        hip.project_quadrics(g, p16, hip.phantoms.shepp_logan(20.))
    # the backprojectors do read 16-bit stacks
    for cls in (hip.HipBackProjectionImageFilter, hip.HipJosephBackProjectionImageFilter):
        f = cls()
        f.SetInput(0, v32)
        f.SetInput(1, p16)
        f.SetGeometry(g)
        f.Update()
        assert torch.isfinite(f.GetOutput().tensor).all()


def test_cg_weights_must_match_the_projections(hip):
    """The CG weights input is read as one float per projection pixel: a weights image of
    another size (e.g. a single view for a four-view stack) or dtype is refused."""
    g = hip_geometry(hip, ROWS[:4])
    size, org, sp = _stack_img(24, 20, 4)
    p = hip.ConstantImageSource(size, sp, org, constant=1.)
    v = hip.ConstantImageSource(VSIZE, VSP, VORG)
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, v)
    f.SetInput(1, p)
    f.SetGeometry(g)
    f.SetNumberOfIterations(1)
    with pytest.raises(ValueError):
        f.SetInputWeights(p.like(p.tensor.half()))
    f.SetInputWeights(hip.ConstantImageSource((24, 20, 1), sp, org, constant=1.))
    with pytest.raises(ValueError):
        f.Update()
    f.SetInputWeights(hip.ConstantImageSource(size, sp, org, constant=1.))
    f.Update()
    assert torch.isfinite(f.GetOutput().tensor).all()
