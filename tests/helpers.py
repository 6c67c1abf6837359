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


def row_band(oracle, rows, lo, hi, v_origin, v_spacing, nv, margin=2):
    """Detector rows [r0, r1) that points of the box [lo, hi] (physical x, y, z) reach
    in any view, plus `margin` rows on each side.  A central projection maps the box
    into the convex hull of its corners' images, so the corners bound the rows."""
    m, *_ = oracle.geometry(rows)
    c = np.array([[x, y, z, 1.] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    h = np.einsum("prc,nc->pnr", m, c)
    v = h[..., 1] / h[..., 2]
    r0 = int(np.floor((v.min() - v_origin) / v_spacing)) - margin
    r1 = int(np.ceil((v.max() - v_origin) / v_spacing)) + 1 + margin
    return max(r0, 0), min(r1, nv)
