"""Pins the oracle's ramp filter and FDK weighting.

* Known answers of test/rtkrampfiltertest2.cxx:54-83 (also pin ITK's FFT
  normalisation and the HannY 2-D kernel path).
* The oracle's own FFT (radix-2 + Bluestein) against numpy.fft on a padded size
  that is not a power of two (greatest prime factor 13, the FFTW configuration).
* The Hann window law on the kernel spectrum.
* FDK weighting against the closed-form cosine of the ray angle.
"""
import numpy as np
import pytest

from helpers import circular_rows


@pytest.mark.parametrize("kw,expected", [({}, 0.132652), ({"hann": 0.8}, 0.149724),
                                         ({"hann": 0.8, "hannY": 0.1}, 0.150181)])
def test_ramp_known_answers(oracle, kw, expected):
    st = oracle.image((64, 64, 64), (0, 0, 0), (1, 1, 1))
    p = np.full((64, 64, 64), 10, np.float32)
    out = oracle.ramp(st, p, **kw)
    assert abs(out[26, 21, 3] - expected) <= 1e-6


def test_ramp_matches_numpy_fft_non_pow2(oracle):
    nu, nv = 57, 3
    xpad, ypad = oracle.padded_size(nu, nv, gpf=13)
    assert xpad == 117 and ypad == 3  # 114 -> 117 = 3^2 * 13
    st = oracle.image((nu, nv, 2), (0, 0, 0), (0.7, 1, 1))
    rng = np.random.default_rng(1)
    p = rng.normal(size=(2, nv, nu))
    out = oracle.ramp(st, p, hann=0.7, gpf=13)
    K = oracle.ramp_kernel(xpad, ypad, 0.7, hann=0.7)[0]
    # Independent route: numpy real FFT of the zero-padded row times the kernel.
    z = (xpad - nu) // 2
    pad = np.zeros((2, nv, xpad))
    pad[..., z:z + nu] = p
    ref = np.fft.irfft(np.fft.rfft(pad, axis=-1) * K, n=xpad, axis=-1)[..., z:z + nu]
    assert np.max(np.abs(out - ref)) < 1e-12 * np.max(np.abs(ref)) * 100


def test_ramp_kernel_spectrum_and_hann_window(oracle):
    xpad, sp = 256, 0.8
    K0 = oracle.ramp_kernel(xpad, 1, sp)[0]
    # Kak & Slaney spatial Ram-Lak kernel transformed by numpy.
    k = np.zeros(xpad)
    k[0] = 1 / (4 * sp)
    n = np.arange(1, xpad // 2, 2)
    k[n] = k[xpad - n] = -1 / (np.pi ** 2 * n ** 2 * sp)
    np.testing.assert_allclose(K0, np.fft.fft(k)[: xpad // 2 + 1], atol=1e-13)
    Kh = oracle.ramp_kernel(xpad, 1, sp, hann=0.5)[0]
    nb = xpad // 2 + 1
    ncut = int(np.floor(nb * 0.5 + 0.5))
    i = np.arange(nb)
    win = np.where(i < ncut, 0.5 * (1 + np.cos(np.pi * i / ncut)), 0.)
    np.testing.assert_allclose(Kh, K0 * win, atol=1e-13)


def test_fdk_weight_closed_form(oracle):
    # Centred full scan: cos(gamma) = sdd / sqrt(sdd^2 + x^2 + y^2), factor = (2 pi / N) sdd / (2 sid).
    n, sid, sdd = 8, 1000., 1536.
    rows = circular_rows(n, sid, sdd)
    st = oracle.image((9, 7, n), (-8.0, -6.0, 0), (2.0, 2.0, 1))
    p = np.ones((n, 7, 9))
    w = oracle.fdk_weight(rows, st, p)
    x = -8.0 + 2.0 * np.arange(9)
    y = -6.0 + 2.0 * np.arange(7)
    X, Y = np.meshgrid(x, y)
    ref = (2 * np.pi / n) * sdd / (2 * sid) * sdd / np.sqrt(sdd ** 2 + X ** 2 + Y ** 2)
    for k in range(n):
        np.testing.assert_allclose(w[k], ref, rtol=1e-12)


def test_fdk_weight_offsets_are_ray_cosines(oracle):
    # With detector offsets the weight is still factor * cos(angle between the ray
    # to the pixel and the central ray), the central ray being source -> isocentre
    # direction at the detector (no tilt when the source is not offset in x).
    n, sid, sdd = 6, 800., 1200.
    rows = circular_rows(n, sid, sdd, proj_off=[(3.0 * i, -2.0 * i) for i in range(n)])
    st = oracle.image((5, 4, n), (-4.0, -3.0, 0), (2.0, 2.0, 1))
    w = oracle.fdk_weight(rows, st, np.ones((n, 4, 5)))
    for k in range(n):
        x = -4.0 + 2.0 * np.arange(5) + rows[k, 3]
        y = -3.0 + 2.0 * np.arange(4) + rows[k, 4]
        X, Y = np.meshgrid(x, y)
        cos = sdd / np.sqrt(sdd ** 2 + X ** 2 + Y ** 2)
        np.testing.assert_allclose(w[k], (2 * np.pi / n) * sdd / (2 * sid) * cos, rtol=1e-12)
