"""Writes the golden fixtures in this folder from the double-precision oracle.

Each fixture is a small .npz holding the inputs (geometry rows, image descriptors,
pixel data) and the oracle<double> output of one hot-path function, at sizes where
the oracle finishes in well under a second.  The oracle itself is pinned against
the reference's known answers and closed forms (tests/test_oracle_*.py); the
fixtures freeze its double results so that the HIP kernels can be checked on a
machine where only the fixtures are present, and so that any later change of the
oracle that moves a result shows up as a failing fixture test.

Usage (from the repository root): python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import rtkref  # noqa: E402
from rtk_amd import phantoms  # noqa: E402

# geometry rows: sid, sdd, gantry, projOffsetX, projOffsetY, outOfPlane, inPlane, srcOffsetX, srcOffsetY
ROWS_OFFSETS = np.array([[600., 1000., 360. / 10 * i + 4., 2. * i - 7., 1.5 - 0.4 * i, 0., 0., 3. - 0.5 * i, 1.]
                         for i in range(10)])
ROWS_CIRC = np.array([[1000., 1536., 360. / 24 * i, 0., 0., 0., 0., 0., 0.] for i in range(24)])
VOL = dict(size=(20, 18, 16), origin=(-28.5, -25.5, -22.5), spacing=(3., 3., 3.))
STACK = dict(size=(30, 26, 10), origin=(-60.9, -52.5, 0.), spacing=(4.2, 4.2, 1.))


def _img(d):
    return rtkref.image(d["size"], d["origin"], d["spacing"])


def _save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(f"{path}: {os.path.getsize(path)} bytes")


def make():
    rng = np.random.default_rng(20260002)
    vz, vy, vx = VOL["size"][::-1]
    sz = (STACK["size"][2], STACK["size"][1], STACK["size"][0])
    # inputs are stored as float32 (what the HIP path takes); the oracle runs in double on them
    volume = rng.random((vz, vy, vx)).astype(np.float32).astype(np.float64)
    proj = rng.normal(size=sz).astype(np.float32).astype(np.float64)
    common = dict(vol_size=VOL["size"], vol_origin=VOL["origin"], vol_spacing=VOL["spacing"],
                  stack_size=STACK["size"], stack_origin=STACK["origin"], stack_spacing=STACK["spacing"],
                  rows=ROWS_OFFSETS)
    ov, os_ = _img(VOL), _img(STACK)
    # Joseph forward projection of a random volume
    fp = rtkref.joseph(ROWS_OFFSETS, os_, np.zeros(sz), ov, volume)
    _save("joseph_forward", volume=volume.astype(np.float32), expected=fp, **common)
    # Joseph backprojection and voxel-driven plain / FDK backprojection of random projections
    jb = rtkref.joseph_back(ROWS_OFFSETS, ov, np.zeros((vz, vy, vx)), os_, proj)
    vb = rtkref.backproject(ROWS_OFFSETS, ov, np.zeros((vz, vy, vx)), os_, proj, fdk=False)
    fb = rtkref.backproject(ROWS_OFFSETS, ov, np.zeros((vz, vy, vx)), os_, proj, fdk=True)
    _save("backprojections", proj=proj.astype(np.float32), expected_joseph=jb, expected_voxel=vb, expected_fdk=fb, **common)
    # FDK (Hann 0.8) of the analytic Shepp-Logan projections, 24 views, 64^2 detector
    q = phantoms.shepp_logan(24.)
    fvol = dict(size=(24, 24, 24), origin=(-23.,) * 3, spacing=(2., 2., 2.))
    fst = dict(size=(64, 64, 24), origin=(-50.4, -50.4, 0.), spacing=(1.6, 1.6, 1.))
    sl = rtkref.project_quadrics(ROWS_CIRC, _img(fst), np.zeros((24, 64, 64)), q).astype(np.float32)
    rec = rtkref.fdk(ROWS_CIRC, _img(fvol), np.zeros((24, 24, 24)), _img(fst), sl.astype(np.float64), hann=0.8)
    _save("fdk_shepp_logan", proj=sl, expected=rec, rows=ROWS_CIRC, hann=0.8,
          vol_size=fvol["size"], vol_origin=fvol["origin"], vol_spacing=fvol["spacing"],
          stack_size=fst["size"], stack_origin=fst["origin"], stack_spacing=fst["spacing"])
    # CG with the matched Joseph pair: 6 iterations from zero on the FP of the random volume
    fp32 = fp.astype(np.float32)
    x, res = rtkref.conjugate_gradient(ROWS_OFFSETS, ov, np.zeros((vz, vy, vx)), os_, fp32.astype(np.float64), 6,
                                       joseph_bp=True)
    _save("cg_joseph_pair", proj=fp32, niter=6, expected=x, expected_residuals=np.sqrt(res / res[0]), **common)


if __name__ == "__main__":
    make()
