"""Writes tests/golden/cg_c3.npz: the double-precision oracle's plain CG at full size.

Workload: a 256^3 volume at 1 mm (centred) reconstructed from 360 views of 512 x 512
pixels at 0.8 mm (SID 1000 mm, SDD 1536 mm, 360 degrees), the projections being the
FP64 analytic line integrals of the seeded 20-ellipsoid phantom (seed PHANTOM_SEED,
rounded to float32 as the HIP phantom projector stores them), with Poisson noise at
I0 = 1e4 drawn per pixel by inversion of numpy uniforms (seed NOISE_SEED,
rtk_amd.phantoms.poisson_line_integrals) and converted back to line integrals.
Thirty iterations of CG from zero with RTK's default projector pair (Joseph forward,
voxel-based back projection; include/rtkConjugateGradientConeBeamReconstructionFilter.hxx:195-305),
scalars in double.

Stored (the volume itself is 64 MB, so only what the GPU test compares):
  residual       ||R_k|| / ||R_0|| for k = 0..30 (||R_0|| = ||B|| from a zero start)
  b_norm         ||B|| = ||BP(p)||
  block_means    8^3-voxel block means of X_30, shape (32, 32, 32)
  slice_y        X_30[:, 128, :] (the central plane normal to the rotation axis)
  slice_z        X_30[128, :, :]
  x_norm, x_sum  ||X_30|| and sum(X_30)
  counts_sum     sum of the Poisson counts (checks that the GPU test drew the same input)

With --float, the same 30 iterations run through the float oracle (RTK's float
filters restated, CG scalars in float as in include/rtkConjugateGradientImageFilter.hxx)
on the same input, and its residual history, planes and block means are added to the
fixture under the same names with the suffix _f32: how far an FP32 implementation of
RTK's own algorithm lies from the double one on this problem.

Each run needs about an hour on 8 cores.  Usage (repository root):
  python tests/golden/make_cg_c3.py [--threads 8] [--float]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import rtkref  # noqa: E402
from rtk_amd import phantoms  # noqa: E402

NX, NU, NVIEWS, I0, NITER = 256, 512, 360, 1e4, 30
PHANTOM_SEED, NOISE_SEED = 20260004, 20260005


def rows():
    return np.array([[1000., 1536., i * 360. / NVIEWS, 0., 0., 0., 0., 0., 0.] for i in range(NVIEWS)])


def stack_image():
    det0 = -(NU - 1) * 0.8 / 2
    return dict(size=(NU, NU, NVIEWS), origin=(det0, det0, 0.), spacing=(0.8, 0.8, 1.))


def volume_image():
    return dict(size=(NX, NX, NX), origin=(-(NX - 1) / 2.,) * 3, spacing=(1., 1., 1.))


def quadrics():
    return phantoms.random_ellipsoids(PHANTOM_SEED, n=20, half_fov=NX / 2.)


def noisy(clean32, threads=8):
    """Poisson counts and noisy line integrals from the float32 clean stack, drawn by
    rtk_amd.phantoms.poisson_line_integrals (per-pixel inversion): the GPU test and the
    benchmark apply the same function to the HIP phantom projector's output, which
    equals this one's to FP32 rounding."""
    p, counts = phantoms.poisson_line_integrals(clean32, I0, NOISE_SEED, return_counts=True, threads=threads)
    return counts, p


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--float", action="store_true", help="add the float oracle's results to the fixture")
    args = ap.parse_args()
    s, v = stack_image(), volume_image()
    st = rtkref.image(s["size"], s["origin"], s["spacing"])
    ov = rtkref.image(v["size"], v["origin"], v["spacing"])
    t0 = time.time()
    clean = rtkref.project_quadrics(rows(), st, np.zeros((NVIEWS, NU, NU)), quadrics(), threads=args.threads)
    counts, p = noisy(clean.astype(np.float32), args.threads)
    del clean
    print(f"projections {time.time() - t0:.0f} s, counts sum {int(counts.sum())}", flush=True)
    path = os.path.join(HERE, "cg_c3.npz")
    if args.float:
        fx = dict(np.load(path))
        assert int(counts.sum()) == int(fx["counts_sum"])
        x, res2 = rtkref.conjugate_gradient(rows(), ov, np.zeros((NX,) * 3, np.float32), st, p, NITER,
                                            threads=args.threads)
        print(f"float CG {time.time() - t0:.0f} s", flush=True)
        xd = x.astype(np.float64)
        fx.update(residual_f32=np.sqrt(res2 / res2[0]),
                  block_means_f32=xd.reshape(32, 8, 32, 8, 32, 8).mean(axis=(1, 3, 5)).astype(np.float32),
                  slice_y_f32=x[:, NX // 2, :].copy(), slice_z_f32=x[NX // 2].copy())
        np.savez_compressed(path, **fx)
        print(f"{path}: {os.path.getsize(path)} bytes", flush=True)
        return
    x, res2 = rtkref.conjugate_gradient(rows(), ov, np.zeros((NX,) * 3), st, p.astype(np.float64), NITER,
                                        threads=args.threads)
    print(f"CG {time.time() - t0:.0f} s", flush=True)
    res = np.sqrt(res2 / res2[0])
    bm = x.reshape(32, 8, 32, 8, 32, 8).mean(axis=(1, 3, 5))
    np.savez_compressed(path, residual=res, b_norm=np.sqrt(res2[0]), block_means=bm.astype(np.float32),
                        slice_y=x[:, NX // 2, :].astype(np.float32), slice_z=x[NX // 2].astype(np.float32),
                        x_norm=np.linalg.norm(x), x_sum=x.sum(), counts_sum=np.int64(counts.sum()),
                        phantom_seed=PHANTOM_SEED, noise_seed=NOISE_SEED, i0=I0, niter=NITER)
    print(f"{path}: {os.path.getsize(path)} bytes; residual {res[-1]:.6g}", flush=True)


if __name__ == "__main__":
    main()
