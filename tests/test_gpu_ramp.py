"""The FFT ramp filter on HIP (rtk::FFTRampImageFilter's counterpart) at the filter
level: the reference's known answers (test/rtkrampfiltertest2.cxx:54-83, including the
2-D HannCutFrequencyY kernel), padded lengths with prime factors up to 13 (RTK's CUDA
filter and FFTW builds, include/rtkFFTProjectionsConvolutionImageFilter.hxx:35-48,280-341)
against the oracle, the 2-D kernel inside FDK against the oracle and the reference's
thresholds (test/rtkrampfiltertest.cxx:114-119), and the column filter entry point
alone against numpy.  Bars: 1e-6 absolute on the known answers (the reference's),
1e-5 relative L2 on filtered stacks, 1e-4 on FDK volumes."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import check_image_quality, hip_geometry, rel_l2, to_device

pytestmark = pytest.mark.gpu


def _ramp(hip, img, **kw):
    f = hip.HipFFTRampImageFilter()
    f.SetInput(img)
    for k, v in kw.items():
        getattr(f, k)(v)
    f.Update()
    return f.GetOutput().numpy()


@pytest.mark.parametrize("gpf", [2, 13])
def test_ramp_known_answers(hip, gpf):
    """A 64^3 image of 10s; the value at index (3, 21, 26) after ramp filtering with no
    window, with a Hann window (0.8) and with the 2-D kernel (HannY 0.1)."""
    img = hip.ConstantImageSource((64, 64, 64), constant=10.)
    for kw, expected in (({}, 0.132652), ({"SetHannCutFrequency": 0.8}, 0.149724),
                         ({"SetHannCutFrequency": 0.8, "SetHannCutFrequencyY": 0.1}, 0.150181)):
        out = _ramp(hip, img, SetGreatestPrimeFactor=gpf, **kw)
        print(gpf, kw, out[26, 21, 3])
        assert abs(out[26, 21, 3] - expected) <= 1e-6, (kw, out[26, 21, 3])


# (nu, nv, options): padded widths 2 nu made 13-smooth: 114 -> 231 = 3 7 11, 77 -> 154 =
# 2 7 11, 100 -> 200, 600 -> 1200, 65 -> 130 = 2 5 13; the 2-D kernel pads v to 2 nv
# made smooth (nv 9 -> 18, 13 -> 26, 20 -> 40)
CASES = [(114, 9, dict(hann=0.8)), (114, 9, dict(truncation=0.1)), (77, 13, dict(shepplogan=0.7)),
         (100, 9, dict(cosine=0.6)), (600, 5, dict(hamming=0.9)), (65, 13, dict(ramlak=0.5)),
         (114, 20, dict(hann=0.8, hannY=0.3)), (77, 9, dict(hannY=0.6, truncation=0.1)),
         (64, 13, dict(hannY=1.0))]
SETTERS = dict(hann="SetHannCutFrequency", hannY="SetHannCutFrequencyY", cosine="SetCosineCutFrequency",
               hamming="SetHammingFrequency", ramlak="SetRamLakCutFrequency", shepplogan="SetSheppLoganCutFrequency",
               truncation="SetTruncationCorrection")


@pytest.mark.parametrize("gpf", [13, 2])
@pytest.mark.parametrize("nu,nv,kw", CASES, ids=[f"{a}x{b}-{c}" for a, b, c in CASES])
def test_ramp_matches_oracle(oracle, hip, nu, nv, kw, gpf):
    n = 3
    org, sp = (-(nu - 1) * 0.35, -(nv - 1) * 0.4, 0.), (0.7, 0.8, 1.)
    p = (1 + 0.2 * np.random.default_rng(nu + nv).normal(size=(n, nv, nu))).astype(np.float32)
    ref = oracle.ramp(oracle.image((nu, nv, n), org, sp), p, gpf=gpf, **kw)
    out = _ramp(hip, to_device(hip, p, org, sp), SetGreatestPrimeFactor=gpf, **{SETTERS[k]: v for k, v in kw.items()})
    xpad, ypad = oracle.padded_size(nu, nv, gpf=gpf, hannY=kw.get("hannY", 0.))
    e = rel_l2(out, ref)
    print(f"{nu}x{nv} {kw} gpf {gpf} (pad {xpad} x {ypad}): rel L2 {e:.3g}")
    assert e <= 1e-5


def test_column_filter_entry_point_vs_numpy(hip):
    """rtkhip_column_filter alone: circular convolution of each zero-padded column with a
    real kernel given by its Hermitian spectrum (not an even one), FP32 and BF16 outputs."""
    from rtk_amd import _lib
    nu, nv, n, ypad = 37, 21, 2, 55  # 55 = 5 11
    rng = np.random.default_rng(3)
    h = rng.normal(size=ypad)
    K = np.fft.fft(h)
    x = rng.normal(size=(n, nv, nu)).astype(np.float32)
    pad = np.zeros((n, ypad, nu))
    pad[:, :nv] = x
    ref = np.real(np.fft.ifft(np.fft.fft(pad, axis=1) * K[None, :, None], axis=1))[:, :nv]
    kk, kp = _lib.dbl(np.stack([K.real, K.imag], -1).ravel())
    xin = torch.from_numpy(x).cuda()
    for dt, tol in ((torch.float32, 1e-5), (torch.bfloat16, 1e-2)):
        y = torch.empty(x.shape, dtype=dt, device="cuda")
        code = {torch.float32: _lib.F32, torch.bfloat16: _lib.BF16}[dt]
        _lib.check(_lib.lib().rtkhip_column_filter(_lib.int3((nu, nv, n)), ypad, kp, ctypes.c_void_p(xin.data_ptr()),
                                                   ctypes.c_void_p(y.data_ptr()), code, None))
        torch.cuda.synchronize()
        assert rel_l2(y.float().cpu().numpy(), ref) <= tol, dt
    # a prime factor above 13 is refused
    bad = _lib.lib().rtkhip_column_filter(_lib.int3((nu, nv, n)), 34, kp, ctypes.c_void_p(xin.data_ptr()),
                                          ctypes.c_void_p(xin.data_ptr()), _lib.F32, None)
    assert bad == _lib.RTKHIP_ERR_UNSUPPORTED


def _noisy_shepp_logan(hip, nu=128):
    from rtk_amd import phantoms
    n = 180
    rows = np.array([[600., 1200., i * 360. / n, 0, 0, 0, 0, 0, 0] for i in range(n)])
    q = phantoms.shepp_logan(128.)
    g = hip_geometry(hip, rows)
    clean = hip.project_quadrics(g, hip.ConstantImageSource((nu, 128, n), (4., 4., 4.), (-254.,) * 3, 0.), q)
    return rows, q, g, clean


def test_fdk_hann_y_reference_thresholds_and_oracle(oracle, hip):
    """test/rtkrampfiltertest.cxx:114-119: the noisy Shepp-Logan FDK of the Hann test
    with HannCutFrequencyY 0.8 added: error per voxel < 0.74 and PSNR > 21.8 dB; the HIP
    volume matches the float oracle on the same noisy stack (1e-4), and FP16 / BF16
    filtered stacks (the column pass writing them) stay within 1e-3 / 1e-2 of it."""
    rows, q, g, clean = _noisy_shepp_logan(hip)
    n = len(rows)
    noisy = clean.numpy() + np.random.default_rng(12).normal(0., 1., size=(n, 128, 128)).astype(np.float32)
    out = {}
    for dt in ("f32", "f16", "bf16"):
        f = hip.HipFDKConeBeamReconstructionFilter()
        f.SetInput(0, hip.ConstantImageSource((128, 128, 128), (2., 2., 2.), (-127.,) * 3, 0.))
        f.SetInput(1, to_device(hip, noisy, (-254.,) * 3, (4.,) * 3))
        f.SetGeometry(g)
        f.GetRampFilter().SetHannCutFrequency(0.8)
        f.GetRampFilter().SetHannCutFrequencyY(0.8)
        f.SetStackDtype(dt)
        f.Update()
        out[dt] = f.GetOutput().numpy()
    ov = oracle.image((128,) * 3, (-127.,) * 3, (2.,) * 3)
    os_ = oracle.image((128, 128, n), (-254.,) * 3, (4.,) * 3)
    ref = oracle.fdk(rows, ov, np.zeros((128,) * 3, np.float32), os_, noisy, hann=0.8, hannY=0.8, gpf=13)
    assert rel_l2(out["f32"], ref) <= 1e-4
    assert rel_l2(out["f16"], ref) <= 1e-3
    assert rel_l2(out["bf16"], ref) <= 1e-2
    phantom = oracle.draw_quadrics(ov, np.zeros((128,) * 3, np.float32), q)
    err, psnr, ok = check_image_quality(out["f32"], phantom, 0.74, 21.8, 2.)
    print(f"HannY FDK: err {err:.4f} psnr {psnr:.2f}")
    assert ok, (err, psnr)


def test_fdk_truncated_detector_13_smooth_padding(oracle, hip):
    """test/rtkrampfiltertest.cxx:121-134 as RTK's CUDA FDK runs it: the 114-column
    detector is padded to 231 = 3 7 11 columns (greatest prime factor 13) with a 0.1
    truncation correction: error per voxel < 0.12 and PSNR > 20.8 dB, and the volume
    matches the oracle run with the same padding (1e-4)."""
    rows, q, g, clean = _noisy_shepp_logan(hip, nu=114)
    f = hip.HipFDKConeBeamReconstructionFilter()
    f.SetInput(0, hip.ConstantImageSource((128, 128, 128), (2., 2., 2.), (-127.,) * 3, 0.))
    f.SetInput(1, clean)
    f.SetGeometry(g)
    f.GetRampFilter().SetTruncationCorrection(0.1)
    f.Update()
    rec = f.GetOutput().numpy()
    n = len(rows)
    ov = oracle.image((128,) * 3, (-127.,) * 3, (2.,) * 3)
    os_ = oracle.image((114, 128, n), clean.origin, clean.spacing)
    assert oracle.padded_size(114, 128, gpf=13)[0] == 231
    ref = oracle.fdk(rows, ov, np.zeros((128,) * 3, np.float32), os_, clean.numpy(), truncation=0.1, gpf=13)
    assert rel_l2(rec, ref) <= 1e-4
    phantom = oracle.draw_quadrics(ov, np.zeros((128,) * 3, np.float32), q)
    err, psnr, ok = check_image_quality(rec, phantom, 0.12, 20.8, 2.)
    print(f"truncated FDK (231-sample rows): err {err:.4f} psnr {psnr:.2f}")
    assert ok, (err, psnr)
