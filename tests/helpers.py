"""Shared test helpers: geometries and image descriptors in both the oracle's and
the HIP package's representation, and comparison metrics."""
import numpy as np


def circular_rows(n, sid=1000., sdd=1536., arc=360., start=0., proj_off=None, src_off=(0., 0.), oop=0., ip=0.):
    rows = []
    for i in range(n):
        px, py = (0., 0.) if proj_off is None else proj_off[i]
        rows.append([sid, sdd, start + i * arc / n, px, py, oop, ip, src_off[0], src_off[1]])
    return np.array(rows, dtype=np.float64)


def hip_geometry(rtk, rows, radius=0.):
    g = rtk.ThreeDCircularProjectionGeometry()
    for r in rows:
        g.AddProjection(*r)
    if radius:
        g.SetRadiusCylindricalDetector(radius)
    return g


def centered_origin(size, spacing):
    return [-(n - 1) * s / 2. for n, s in zip(size, spacing)]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.)


def check_image_quality(recon, ref, err_tol, psnr_tol, ref_value):
    """test/rtkTest.h:38-105: mean |err| per pixel and PSNR against ref_value."""
    d = np.asarray(ref, np.float64) - np.asarray(recon, np.float64)
    err = np.abs(d).mean()
    mse = (d ** 2).mean()
    psnr = 20 * np.log10(ref_value) - 10 * np.log10(mse)
    return err, psnr, (err < err_tol and psnr > psnr_tol)


def to_device(rtk, arr, origin, spacing, index=(0, 0, 0), dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float32)).cuda()
    if dtype is not None:
        t = t.to(dtype)
    return rtk.Image(t, origin, spacing, index)
