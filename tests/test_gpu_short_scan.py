"""Parker short-scan and displaced-detector weighting on HIP against the oracle,
and the FDK / CG chains they feed.

* kernels: the reference's CUDA-vs-CPU comparison geometries
  (test/rtkshortscancompcudatest.cxx, test/rtkdisplaceddetectorcompcudatest.cxx:
  tilted views with source and projection offsets, 1e-5 there) plus the
  displaced-detector cases of test/rtkdisplaceddetectortest.cxx; relative L2 vs the
  float oracle <= 1e-6 and exact zeros in the padding;
* view subsets (stack index > 0), in-place, widths that are not multiples of 4;
* FDK end to end: test/rtkshortscantest.cxx (0.09 / 22 dB) and
  test/rtkdisplaceddetectortest.cxx (0.061 / 24 dB) on the HIP chain, and the HIP
  volume within 1e-4 of the oracle's chain;
* CG with a displaced detector: weights D(1) without padding
  (rtkConjugateGradientConeBeamReconstructionFilter.hxx:54,222-242) vs the oracle.
"""
import numpy as np
import pytest
import torch

from helpers import check_image_quality, hip_geometry, rel_l2, to_device

pytestmark = pytest.mark.gpu

ODD = np.array([[600., 700., 0., 84., 35, 23, 15, 21, 26], [500., 800., 45., 21., 12, 16, 546, 14, 41],
                [700., 900., 90., 68., 68, 54, 38, 35, 56], [900., 1000., 135., 48., 35, 84, 10, 84, 59]])


def _arc(n, arc, sid=600., sdd=1200.):
    return np.array([[sid, sdd, i * arc / n, 0, 0, 0, 0, 0, 0] for i in range(n)])


def _disp(n, px):
    return np.array([[600., 1200., i * 360. / n, px, 0., 0, 0, 0, 0] for i in range(n)])


def _stack(rng, n, nv, nu):
    return (rng.normal(size=(n, nv, nu)) + 2.).astype(np.float32)


@pytest.mark.parametrize("case", ["odd", "arc240", "arc200_width127"])
def test_parker_kernel_vs_oracle(oracle, hip, case):
    rng = np.random.default_rng(11)
    if case == "odd":
        rows, size, org, sp = ODD, (128, 4, 4), (-127., -3., 0.), (2., 2., 2.)
    elif case == "arc240":
        rows, size, org, sp = _arc(110, 240.), (128, 6, 110), (-254., -10., 0.), (4., 4., 4.)
    else:
        rows, size, org, sp = _arc(90, 200.), (127, 5, 90), (-252., -8., 0.), (4., 4., 4.)
    p = _stack(rng, size[2], size[1], size[0])
    ref, short, *_ = oracle.parker(rows, oracle.image(size, org, sp), p)
    assert short
    f = hip.HipParkerShortScanImageFilter()
    f.SetInput(to_device(hip, p, org, sp))
    f.SetGeometry(hip_geometry(hip, rows))
    if case == "arc200_width127":
        # delta (8.5 deg) is below the half fan angle (12 deg): the reference warns
        with pytest.warns(UserWarning, match="not have enough data"):
            f.Update()
    else:
        f.Update()
    assert f.IsShortScan()
    out = f.GetOutput().numpy()
    e = rel_l2(out, ref)
    print(f"Parker {case}: rel L2 {e:.3g}")
    assert e <= 1e-6
    assert np.array_equal(out == 0, ref == 0)


def test_parker_subset_inplace_and_full_scan(oracle, hip):
    rng = np.random.default_rng(12)
    rows = _arc(110, 240.)
    p = _stack(rng, 110, 4, 96)
    ref = oracle.parker(rows, oracle.image((96, 4, 110), (-190., -6., 0.), (4., 4., 4.)), p)[0]
    # views 30..69 as a stack whose region starts at index 30, weighted in place
    sub = to_device(hip, p[30:70], (-190., -6., 0.), (4., 4., 4.), index=(0, 0, 30))
    t = sub.tensor
    f = hip.HipParkerShortScanImageFilter()
    f.SetInput(sub)
    f.SetGeometry(hip_geometry(hip, rows))
    f.InPlaceOn()
    f.Update()
    assert f.GetOutput().tensor.data_ptr() == t.data_ptr()
    assert rel_l2(f.GetOutput().numpy(), ref[30:70]) <= 1e-6
    # full scan: the data pass through unchanged
    full = _arc(120, 360.)
    q = _stack(rng, 120, 4, 96)
    f = hip.HipParkerShortScanImageFilter()
    f.SetInput(to_device(hip, q, (-190., -6., 0.), (4., 4., 4.)))
    f.SetGeometry(hip_geometry(hip, full))
    f.Update()
    assert not f.IsShortScan() and np.array_equal(f.GetOutput().numpy(), q)


DISPLACED = [("odd", ODD, (128, 4, 4), (-127., -3., 0.), (2., 2., 2.), {}),
             ("pos", _disp(180, 120.), (128, 4, 180), (-254., -6., 0.), (4., 4., 4.), {}),
             ("neg", _disp(180, -120.), (128, 4, 180), (-254., -6., 0.), (4., 4., 4.), {}),
             ("origin_neg", _disp(180, 0.), (128, 4, 180), (-400., -6., 0.), (4., 4., 4.), {}),
             ("origin_pos", _disp(180, 0.), (128, 4, 180), (-100., -6., 0.), (4., 4., 4.), {}),
             ("nopad", _disp(180, 120.), (128, 4, 180), (-254., -6., 0.), (4., 4., 4.), {"pad": False}),
             ("width_odd", _disp(60, 90.), (127, 3, 60), (-252., -4., 0.), (4., 4., 4.), {}),
             ("offsets", _disp(60, 0.), (128, 3, 60), (-254., -4., 0.), (4., 4., 4.), {"offsets": (-100., 120.)})]


@pytest.mark.parametrize("name,rows,size,org,sp,kw", DISPLACED, ids=[d[0] for d in DISPLACED])
def test_displaced_kernel_vs_oracle(oracle, hip, name, rows, size, org, sp, kw):
    rng = np.random.default_rng(13)
    p = _stack(rng, size[2], size[1], size[0])
    ref, oim = oracle.displaced(rows, oracle.image(size, org, sp), p, pad=kw.get("pad", True),
                                offsets=kw.get("offsets"))
    f = hip.HipDisplacedDetectorImageFilter()
    f.SetInput(to_device(hip, p, org, sp))
    f.SetGeometry(hip_geometry(hip, rows))
    f.SetPadOnTruncatedSide(kw.get("pad", True))
    if "offsets" in kw:
        f.SetOffsets(*kw["offsets"])
    f.Update()
    o = f.GetOutput()
    assert o.size == tuple(oim.size)
    assert o.origin[0] + o.index[0] * o.spacing[0] == pytest.approx(oim.origin[0])
    out = o.numpy()
    e = rel_l2(out, ref)
    print(f"displaced {name}: out {o.size} index {o.index}, rel L2 {e:.3g}")
    assert e <= 1e-6
    assert np.array_equal(out == 0, ref == 0)


def test_displaced_subset_streaming(oracle, hip):
    """Views split over two stacks (regions starting at view 0 and 90), like the
    reference's streamed comparison, equal the whole stack."""
    rng = np.random.default_rng(14)
    rows = _disp(180, 120.)
    p = _stack(rng, 180, 4, 128)
    ref = oracle.displaced(rows, oracle.image((128, 4, 180), (-254., -6., 0.), (4., 4., 4.)), p)[0]
    parts = []
    for a, b in ((0, 90), (90, 180)):
        f = hip.HipDisplacedDetectorImageFilter()
        f.SetInput(to_device(hip, p[a:b], (-254., -6., 0.), (4., 4., 4.), index=(0, 0, a)))
        f.SetGeometry(hip_geometry(hip, rows))
        f.Update()
        parts.append(f.GetOutput().numpy())
    assert rel_l2(np.concatenate(parts), ref) <= 1e-6


def _shepp_logan(hip, oracle, rows, size, org, sp):
    from rtk_amd import phantoms
    q = phantoms.shepp_logan(116.)
    proj = oracle.project_quadrics(rows, oracle.image(size, org, sp), np.zeros(size[::-1], np.float32), q)
    vol = oracle.image((128,) * 3, (-127.,) * 3, (2.,) * 3)
    ref = oracle.draw_quadrics(vol, np.zeros((128,) * 3, np.float32), q)
    return proj, vol, ref


def _hip_fdk(hip, rows, proj_img):
    f = hip.HipFDKConeBeamReconstructionFilter()
    f.SetInput(0, hip.ConstantImageSource((128,) * 3, (2.,) * 3, (-127.,) * 3))
    f.SetInput(1, proj_img)
    f.SetGeometry(hip_geometry(hip, rows))
    f.SetProjectionSubsetSize(16)
    f.Update()
    return f.GetOutput().numpy()


def test_short_scan_fdk_reference_test(oracle, hip):
    rows, size, org, sp = _arc(110, 240.), (128, 128, 110), (-254.,) * 3, (4.,) * 3
    proj, vol, ref = _shepp_logan(hip, oracle, rows, size, org, sp)
    f = hip.HipParkerShortScanImageFilter()
    f.SetInput(to_device(hip, proj, org, sp))
    f.SetGeometry(hip_geometry(hip, rows))
    f.Update()
    rec = _hip_fdk(hip, rows, f.GetOutput())
    err, psnr, ok = check_image_quality(rec, ref, 0.09, 22., 2.)
    print(f"short-scan FDK: err {err:.4f} psnr {psnr:.2f}")
    assert ok, (err, psnr)
    oref = oracle.fdk(rows, vol, np.zeros((128,) * 3, np.float32), oracle.image(size, org, sp),
                      oracle.parker(rows, oracle.image(size, org, sp), proj)[0], gpf=13)
    assert rel_l2(rec, oref) <= 1e-4


@pytest.mark.parametrize("px,ou", [(120., -254.), (-120., -254.), (0., -254.), (0., -400.), (0., -100.)])
def test_displaced_detector_fdk_reference_test(oracle, hip, px, ou):
    rows, size, org, sp = _disp(180, px), (128, 128, 180), (ou, -254., -254.), (4.,) * 3
    proj, vol, ref = _shepp_logan(hip, oracle, rows, size, org, sp)
    f = hip.HipDisplacedDetectorImageFilter()
    f.SetInput(to_device(hip, proj, org, sp))
    f.SetGeometry(hip_geometry(hip, rows))
    f.Update()
    rec = _hip_fdk(hip, rows, f.GetOutput())
    err, psnr, ok = check_image_quality(rec, ref, 0.061, 24., 2.)
    print(f"displaced FDK px={px} origin={ou}: err {err:.4f} psnr {psnr:.2f}")
    assert ok, (err, psnr)
    wp, oim = oracle.displaced(rows, oracle.image(size, org, sp), proj)
    oref = oracle.fdk(rows, vol, np.zeros((128,) * 3, np.float32), oim, wp, gpf=13)
    assert rel_l2(rec, oref) <= 1e-4


def test_cg_displaced_detector_weights(oracle, hip):
    from rtk_amd import phantoms
    n = 60
    rows = _disp(n, 100.)
    size, org, sp = (64, 64, n), (-255., -255., 0.), (8., 8., 8.)
    q = phantoms.ellipsoid((0, 0, 0), (90, 90, 90), None, 1.0)[None]
    st = oracle.image(size, org, sp)
    proj = oracle.project_quadrics(rows, st, np.zeros((n, 64, 64), np.float32), q)
    w, _ = oracle.displaced(rows, st, np.ones_like(proj), pad=False)
    ov = oracle.image((64,) * 3, (-127.,) * 3, (4.,) * 3)
    ref = oracle.conjugate_gradient(rows, ov, np.zeros((64,) * 3), st, proj.astype(np.float64), 6,
                                    weights=w.astype(np.float64))[0]
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, hip.ConstantImageSource((64,) * 3, (4.,) * 3, (-127.,) * 3))
    f.SetInput(1, to_device(hip, proj, org, sp))
    f.SetGeometry(hip_geometry(hip, rows))
    f.SetNumberOfIterations(6)
    f.Update()
    e = rel_l2(f.GetOutput().numpy(), ref)
    print(f"CG displaced: rel L2 vs FP64 oracle {e:.3g}")
    assert e <= 1e-4
    # disabling the filter changes the result (the weights matter here)
    f.SetDisableDisplacedDetectorFilter(True)
    f.Update()
    assert rel_l2(f.GetOutput().numpy(), ref) > 1e-3
