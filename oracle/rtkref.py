"""ctypes front end of the CPU oracle ``librtkref.so`` (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the correctness reference or the timed CPU
baseline.
Arrays are numpy, x-fastest (index order ``[z, y, x]`` / ``[view, v, u]``).
Geometry rows are the nine ``AddProjection`` arguments
(src/rtkThreeDCircularProjectionGeometry.cxx:55-76).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "librtkref.so")


class Image(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int * 3), ("origin", ctypes.c_double * 3),
                ("spacing", ctypes.c_double * 3)]


def image(size, origin, spacing):
    im = Image()
    for i in range(3):
        im.size[i] = int(size[i])
        im.origin[i] = float(origin[i])
        im.spacing[i] = float(spacing[i])
    return im


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.rtkref_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().rtkref_last_error().decode())


def _geom(g9):
    g = np.ascontiguousarray(np.asarray(g9, dtype=np.float64).reshape(-1, 9))
    return g.shape[0], g


def _ramp_opts(hann=0., hannY=0., cosine=0., hamming=0., ramlak=0., shepplogan=0., truncation=0., gpf=2):
    return np.array([hann, hannY, cosine, hamming, ramlak, shepplogan, truncation, gpf], dtype=np.float64)


def _suffix(dtype):
    return {np.float32: "f", np.float64: "d"}[np.dtype(dtype).type]


def geometry(g9):
    """Per-view (matrices[n,3,4], source_pos[n,3], source_angles, angular_gaps, tilt_angles)."""
    n, g = _geom(g9)
    m = np.zeros((n, 3, 4)); s = np.zeros((n, 3)); a = np.zeros(n); ga = np.zeros(n); t = np.zeros(n)
    _check(lib().rtkref_geometry(n, _ptr(g), _ptr(m), _ptr(s), _ptr(a), _ptr(ga), _ptr(t)))
    return m, s, a, ga, t


def bp_matrices(g9, vol, stack, transpose=True, fdk=False):
    n, g = _geom(g9)
    out = np.zeros((n, 3, 4))
    _check(lib().rtkref_bp_matrices(n, _ptr(g), ctypes.byref(vol), ctypes.byref(stack), int(transpose),
                                    int(fdk), _ptr(out)))
    return out


def padded_size(nu, nv, **ramp):
    x = ctypes.c_int(); y = ctypes.c_int()
    _check(lib().rtkref_padded_size(nu, nv, _ptr(_ramp_opts(**ramp)), ctypes.byref(x), ctypes.byref(y)))
    return x.value, y.value


def ramp_kernel(xpad, ypad, spacing, **ramp):
    n = xpad // 2 + 1
    out = np.zeros((ypad, n, 2))
    kh = ctypes.c_int()
    _check(lib().rtkref_ramp_kernel(xpad, ypad, ctypes.c_double(spacing), _ptr(_ramp_opts(**ramp)), _ptr(out),
                                    ctypes.byref(kh)))
    k = out[..., 0] + 1j * out[..., 1]
    return k[:kh.value]


def fdk_weight(g9, stack, proj, first=0, threads=8):
    n, g = _geom(g9)
    proj = np.ascontiguousarray(proj)
    out = np.empty_like(proj)
    fn = getattr(lib(), "rtkref_fdk_weight_" + _suffix(proj.dtype))
    _check(fn(n, _ptr(g), ctypes.byref(stack), first, _ptr(proj), _ptr(out), threads))
    return out


def untilted(g9, view, tilted):
    """ThreeDCircularProjectionGeometry::ToUntiltedCoordinateAtIsocenter."""
    n, g = _geom(g9)
    out = ctypes.c_double()
    _check(lib().rtkref_untilted(n, _ptr(g), int(view), ctypes.c_double(tilted), ctypes.byref(out)))
    return out.value


def parker(g9, stack, proj, threshold=np.pi / 9, threads=8):
    """ParkerShortScanImageFilter: (weighted stack, is_short_scan, delta, first_angle,
    first view lacking data or -1)."""
    n, g = _geom(g9)
    proj = np.ascontiguousarray(proj)
    out = np.empty_like(proj)
    s = ctypes.c_int(); df = np.zeros(2); w = ctypes.c_int()
    fn = getattr(lib(), "rtkref_parker_" + _suffix(proj.dtype))
    _check(fn(n, _ptr(g), ctypes.byref(stack), _ptr(proj), _ptr(out), ctypes.c_double(threshold), threads,
              ctypes.byref(s), _ptr(df), ctypes.byref(w)))
    return out, bool(s.value), df[0], df[1], w.value


def displaced_setup(g9, stack, disable=False, pad=True, offsets=None):
    """DisplacedDetectorImageFilter output information: (out image, column offset of the
    input in it, (inferior, superior) corners, weighting applies)."""
    n, g = _geom(g9)
    out = Image(); off = ctypes.c_int(); c = np.zeros(2); wt = ctypes.c_int()
    mn, mx = offsets if offsets is not None else (0., 0.)
    _check(lib().rtkref_displaced_setup(n, _ptr(g), ctypes.byref(stack), int(disable), int(pad),
                                        int(offsets is not None), ctypes.c_double(mn), ctypes.c_double(mx),
                                        ctypes.byref(out), ctypes.byref(off), _ptr(c), ctypes.byref(wt)))
    return out, off.value, (c[0], c[1]), bool(wt.value)


def displaced(g9, stack, proj, disable=False, pad=True, offsets=None, threads=8):
    """DisplacedDetectorImageFilter: (weighted, possibly zero-padded stack, out image)."""
    n, g = _geom(g9)
    out_im = displaced_setup(g9, stack, disable, pad, offsets)[0]
    proj = np.ascontiguousarray(proj)
    out = np.empty((proj.shape[0], proj.shape[1], out_im.size[0]), dtype=proj.dtype)
    mn, mx = offsets if offsets is not None else (0., 0.)
    fn = getattr(lib(), "rtkref_displaced_" + _suffix(proj.dtype))
    _check(fn(n, _ptr(g), ctypes.byref(stack), _ptr(proj), _ptr(out), int(disable), int(pad),
              int(offsets is not None), ctypes.c_double(mn), ctypes.c_double(mx), threads))
    return out, out_im


def ramp(stack, proj, threads=8, **ramp):
    proj = np.ascontiguousarray(proj)
    out = np.empty_like(proj)
    fn = getattr(lib(), "rtkref_ramp_" + _suffix(proj.dtype))
    _check(fn(ctypes.byref(stack), _ptr(_ramp_opts(**ramp)), _ptr(proj), _ptr(out), threads))
    return out


def backproject(g9, vol, volume, stack, proj, fdk=False, threads=8, geom_offset=0):
    """Returns volume + BP(proj); plain voxel BP or FDK BP."""
    n, g = _geom(g9)
    v = np.array(volume, copy=True, order="C")
    proj = np.ascontiguousarray(proj, dtype=v.dtype)
    fn = getattr(lib(), "rtkref_backproject_" + _suffix(v.dtype))
    _check(fn(n, _ptr(g), ctypes.byref(vol), _ptr(v), ctypes.byref(stack), _ptr(proj), int(fdk), threads,
              geom_offset))
    return v


def joseph(g9, stack, proj_in, vol, volume, threads=8, geom_offset=0, return_samples=False):
    n, g = _geom(g9)
    volume = np.ascontiguousarray(volume)
    pin = np.ascontiguousarray(proj_in, dtype=volume.dtype)
    out = np.empty_like(pin)
    cnt = ctypes.c_longlong()
    fn = getattr(lib(), "rtkref_joseph_" + _suffix(volume.dtype))
    _check(fn(n, _ptr(g), ctypes.byref(stack), _ptr(pin), _ptr(out), ctypes.byref(vol), _ptr(volume), threads,
              geom_offset, ctypes.byref(cnt)))
    return (out, cnt.value) if return_samples else out


def joseph_back(g9, vol, volume, stack, proj, threads=8, geom_offset=0):
    """Returns volume + JosephBP(proj) (the splat transpose of the Joseph FP)."""
    n, g = _geom(g9)
    v = np.array(volume, copy=True, order="C")
    proj = np.ascontiguousarray(proj, dtype=v.dtype)
    fn = getattr(lib(), "rtkref_joseph_back_" + _suffix(v.dtype))
    _check(fn(n, _ptr(g), ctypes.byref(vol), _ptr(v), ctypes.byref(stack), _ptr(proj), threads, geom_offset))
    return v


def fdk(g9, vol, volume, stack, proj, threads=8, subset=16, **ramp):
    n, g = _geom(g9)
    v = np.array(volume, copy=True, order="C")
    proj = np.ascontiguousarray(proj, dtype=v.dtype)
    fn = getattr(lib(), "rtkref_fdk_" + _suffix(v.dtype))
    _check(fn(n, _ptr(g), ctypes.byref(vol), _ptr(v), ctypes.byref(stack), _ptr(proj), _ptr(_ramp_opts(**ramp)),
              threads, subset))
    return v


def conjugate_gradient(g9, vol, x0, stack, proj, niter, weights=None, threads=8, joseph_bp=False):
    """Returns (x, residual_norm2[niter+1]) of plain CG (scalars in the pixel type);
    joseph_bp selects BP_JOSEPH instead of BP_VOXELBASED."""
    n, g = _geom(g9)
    x = np.array(x0, copy=True, order="C")
    proj = np.ascontiguousarray(proj, dtype=x.dtype)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=x.dtype)
    res = np.zeros(niter + 1)
    fn = getattr(lib(), "rtkref_cg_" + _suffix(x.dtype))
    _check(fn(n, _ptr(g), ctypes.byref(vol), _ptr(x), ctypes.byref(stack), _ptr(proj), _ptr(w), niter, threads,
              _ptr(res), int(joseph_bp)))
    return x, res


def project_quadrics(g9, stack, proj_in, quadrics, threads=8):
    n, g = _geom(g9)
    pin = np.ascontiguousarray(proj_in)
    out = np.empty_like(pin)
    q = np.ascontiguousarray(np.asarray(quadrics, dtype=np.float64).reshape(-1, 11))
    fn = getattr(lib(), "rtkref_project_quadrics_" + _suffix(pin.dtype))
    _check(fn(n, _ptr(g), ctypes.byref(stack), _ptr(pin), _ptr(out), q.shape[0], _ptr(q), threads))
    return out


def draw_quadrics(vol, volume_in, quadrics, threads=8):
    vin = np.ascontiguousarray(volume_in)
    out = np.empty_like(vin)
    q = np.ascontiguousarray(np.asarray(quadrics, dtype=np.float64).reshape(-1, 11))
    fn = getattr(lib(), "rtkref_draw_quadrics_" + _suffix(vin.dtype))
    _check(fn(ctypes.byref(vol), _ptr(vin), _ptr(out), q.shape[0], _ptr(q), threads))
    return out
