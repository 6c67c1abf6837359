"""Writes tests/golden/cg_warmstart_oracle.npz: the double-precision oracle's CG
solution on the warm-start workload, the point the learned start's end point is
measured against.

Workload: a 256^3 volume at 1 mm (centred) reconstructed from 90 views of 512 x 512
pixels at 0.8 mm (SID 1000 mm, SDD 1536 mm, 360 degrees), the projections being the
FP64 analytic line integrals of the 20-ellipsoid phantom of held-out seed PHANTOM_SEED
(rounded to float32 as the HIP phantom projector stores them), with Poisson noise at
I0 = 1e4 drawn per pixel by inversion of numpy uniforms (seed NOISE_SEED,
rtk_amd.phantoms.poisson_line_integrals) and converted back to line integrals.
CG from zero with the matched Joseph pair (Joseph forward, Joseph back projection;
include/rtkConjugateGradientConeBeamReconstructionFilter.hxx:195-305), scalars in
double, for NITER iterations: the first iteration count at which the oracle's own
relative residual ||R_k|| / ||B|| is at or below 1e-3, the stopping rule the warm-start
evaluation applies (the fixture's residual history shows it).

Stored (the volume itself is 64 MB, so only what the GPU test compares):
  residual       ||R_k|| / ||R_0|| for k = 0..NITER
  block_means    8^3-voxel block means of X_NITER, shape (32, 32, 32)
  slice_y        X_NITER[:, 128, :]
  slice_z        X_NITER[128, :, :]
  x_norm, x_sum  ||X_NITER|| and sum(X_NITER)
  counts_sum     sum of the Poisson counts (checks that the GPU test drew the same input)

About half an hour on 8 cores.  Usage (repository root):
  python tests/golden/make_warmstart_oracle.py [--threads 8] [--niter N]
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

NX, NU, NVIEWS, I0, NITER, TOL = 256, 512, 90, 1e4, 35, 1e-3
PHANTOM_SEED, NOISE_SEED = 5000, 20260050


def rows():
    return np.array([[1000., 1536., i * 360. / NVIEWS, 0., 0., 0., 0., 0., 0.] for i in range(NVIEWS)])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--niter", type=int, default=NITER)
    args = ap.parse_args()
    det0 = -(NU - 1) * 0.8 / 2
    st = rtkref.image((NU, NU, NVIEWS), (det0, det0, 0.), (0.8, 0.8, 1.))
    ov = rtkref.image((NX,) * 3, (-(NX - 1) / 2.,) * 3, (1., 1., 1.))
    t0 = time.time()
    q = phantoms.random_ellipsoids(PHANTOM_SEED, n=20, half_fov=NX / 2.)
    clean = rtkref.project_quadrics(rows(), st, np.zeros((NVIEWS, NU, NU)), q, threads=args.threads)
    p, counts = phantoms.poisson_line_integrals(clean.astype(np.float32), I0, NOISE_SEED, return_counts=True,
                                                threads=args.threads)
    del clean
    print(f"projections {time.time() - t0:.0f} s, counts sum {int(counts.sum())}", flush=True)
    x, res2 = rtkref.conjugate_gradient(rows(), ov, np.zeros((NX,) * 3), st, p.astype(np.float64), args.niter,
                                        threads=args.threads, joseph_bp=True)
    res = np.sqrt(res2 / res2[0])
    print(f"CG {time.time() - t0:.0f} s; residual " + " ".join(f"{r:.4g}" for r in res[-4:]), flush=True)
    if not (res[-1] <= TOL < res[-2]):
        print(f"the oracle crosses {TOL} at another iteration: rerun with --niter "
              f"{int(np.argmax(res <= TOL)) if (res <= TOL).any() else 'larger'}", flush=True)
        return 1
    bm = x.reshape(32, 8, 32, 8, 32, 8).mean(axis=(1, 3, 5))
    path = os.path.join(HERE, "cg_warmstart_oracle.npz")
    np.savez_compressed(path, residual=res, block_means=bm.astype(np.float32),
                        slice_y=x[:, NX // 2, :].astype(np.float32), slice_z=x[NX // 2].astype(np.float32),
                        x_norm=np.linalg.norm(x), x_sum=x.sum(), counts_sum=np.int64(counts.sum()),
                        phantom_seed=PHANTOM_SEED, noise_seed=NOISE_SEED, i0=I0, niter=args.niter, tol=TOL)
    print(f"{path}: {os.path.getsize(path)} bytes", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
