"""Host logic of the short-scan and displaced-detector filters (no GPU work): the
library's output region, corners and error statuses equal the oracle's
(include/rtkDisplacedDetectorImageFilter.hxx:81-173)."""
import ctypes

import numpy as np
import pytest

from helpers import hip_geometry

# test/rtkdisplaceddetectorcompcudatest.cxx:32-36 (source offsets, tilts, odd distances)
ODD = np.array([[600., 700., 0., 84., 35, 23, 15, 21, 26], [500., 800., 45., 21., 12, 16, 546, 14, 41],
                [700., 900., 90., 68., 68, 54, 38, 35, 56], [900., 1000., 135., 48., 35, 84, 10, 84, 59]])


def _disp(n, px):
    return np.array([[600., 1200., i * 360. / n, px, 0., 0, 0, 0, 0] for i in range(n)])


def _region(rtk, rows, size, origin, spacing, index=(0, 0, 0), **kw):
    from rtk_amd import _lib
    g = hip_geometry(rtk, rows)
    d = _lib.Image()
    for i in range(3):
        d.size[i], d.index[i], d.origin[i], d.spacing[i] = size[i], index[i], origin[i], spacing[i]
    out = _lib.Image()
    c = (ctypes.c_double * 2)()
    wt = ctypes.c_int()
    mn, mx = kw.get("offsets", (0., 0.))
    rc = _lib.lib().rtkhip_displaced_detector_region(g.handle, ctypes.byref(d), int(kw.get("disable", False)),
                                                     int(kw.get("pad", True)), int("offsets" in kw), mn, mx,
                                                     ctypes.byref(out), c, ctypes.byref(wt))
    return rc, out, (c[0], c[1]), bool(wt.value)


CASES = [(_disp(180, 120.), (-254., -254., 0.), {}), (_disp(180, -120.), (-254., -254., 0.), {}),
         (_disp(180, 0.), (-400., -254., 0.), {}), (_disp(180, 0.), (-100., -254., 0.), {}),
         (_disp(180, 0.), (-254., -254., 0.), {}), (_disp(180, 120.), (-254., -254., 0.), {"pad": False}),
         (_disp(180, 120.), (-254., -254., 0.), {"disable": True}), (ODD, (-127., -3., 0.), {}),
         (_disp(180, 0.), (-254., -254., 0.), {"offsets": (-10., 60.)})]


@pytest.mark.parametrize("rows,origin,kw", CASES)
def test_displaced_region_matches_oracle(oracle, rows, origin, kw):
    import rtk_amd
    n = rows.shape[0]
    size, spacing = (128, 4, n), (4., 4., 4.) if rows is not ODD else (2., 2., 2.)
    rc, out, corners, wt = _region(rtk_amd, rows, size, origin, spacing, **kw)
    assert rc == 0
    ost = oracle.image(size, origin, spacing)
    oout, off, oc, owt = oracle.displaced_setup(rows, ost, disable=kw.get("disable", False), pad=kw.get("pad", True),
                                                offsets=kw.get("offsets"))
    assert list(out.size) == list(oout.size) and wt == owt
    if not kw.get("disable"):
        assert corners == pytest.approx(oc, abs=1e-9)
    # ITK keeps the origin and moves the region index; the oracle moves the origin
    assert out.origin[0] + out.index[0] * out.spacing[0] == pytest.approx(oout.origin[0])
    assert out.index[0] == -off


def test_displaced_region_errors():
    import rtk_amd
    from rtk_amd import _lib
    rc, *_ = _region(rtk_amd, _disp(8, 300.), (128, 4, 8), (-254., -254., 0.), (4., 4., 4.))
    assert rc == _lib.RTKHIP_ERR_UNSUPPORTED and b"50%" in _lib.lib().rtkhip_last_error()
    par = np.array([[600., 0., i * 45., 0, 0, 0, 0, 0, 0] for i in range(8)])
    rc, *_ = _region(rtk_amd, par, (128, 4, 8), (-254., -254., 0.), (4., 4., 4.))
    assert rc == _lib.RTKHIP_ERR_UNSUPPORTED and b"parallel" in _lib.lib().rtkhip_last_error()
    # disabled: no check at all
    rc, *_ = _region(rtk_amd, par, (128, 4, 8), (-254., -254., 0.), (4., 4., 4.), disable=True)
    assert rc == 0
