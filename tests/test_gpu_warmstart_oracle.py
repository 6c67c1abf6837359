"""The learned start's end point against the double-precision oracle's CG solution.

Workload: the warm-start evaluation's (256^3 @1 mm from 90 views of 512^2 @0.8 mm, SID
1000 mm, SDD 1536 mm), held-out phantom seed 5000 with Poisson noise at I0 = 1e4 drawn
per pixel by numpy (the draw tests/golden/make_warmstart_oracle.py made), CG with the
matched Joseph pair.  tests/golden/cg_warmstart_oracle.npz holds the oracle's CG from
zero, in double, stopped at the first iteration whose relative residual is at or below
1e-3: its residual history, central planes and 8^3 block means.

Checks:
  * plain CG from zero on the device (FP32 projectors, FP64 scalars) follows the
    oracle's residual history, stops at the same iteration and ends at the oracle's
    solution (planes and block means);
  * CG from CNN(FDK) with the committed weights reaches the same tolerance in fewer
    iterations, and its end point lies closer to the oracle's solution than the FDK
    start's end point does, within the bar stated below.
"""
import os

import numpy as np
import pytest
import torch

from helpers import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "cg_warmstart_oracle.npz")
NX, NU, NVIEWS = 256, 512, 90


def _planes(x):
    a = x.tensor.double()
    return torch.cat([a[:, NX // 2, :].reshape(-1), a[NX // 2].reshape(-1)]).cpu().numpy()


def _block_means(x):
    return x.tensor.double().reshape(32, 8, 32, 8, 32, 8).mean(dim=(1, 3, 5)).cpu().numpy()


def _rms(a, b):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))


def test_cnn_start_end_point_vs_fp64_oracle_solution(hip):
    if not os.path.exists(FIXTURE):
        pytest.skip("tests/golden/cg_warmstart_oracle.npz has not been generated")
    from rtk_amd import phantoms
    from rtk_amd import warmstart_train as W
    from rtk_amd.warmstart import WarmStartCNN, default_weights_path
    fx = dict(np.load(FIXTURE))
    g = W.geometry()
    q = phantoms.random_ellipsoids(int(fx["phantom_seed"]), n=20, half_fov=NX / 2.)
    det0 = -(NU - 1) * 0.8 / 2
    clean = hip.project_quadrics(g, hip.ConstantImageSource((NU, NU, NVIEWS), (0.8, 0.8, 1.), (det0, det0, 0.)), q)
    p, counts = phantoms.poisson_line_integrals(clean.numpy(), float(fx["i0"]), int(fx["noise_seed"]),
                                                return_counts=True)
    assert abs(int(counts.sum()) - int(fx["counts_sum"])) <= 1e-9 * int(fx["counts_sum"])
    proj = clean.like(torch.from_numpy(p).cuda())
    del clean
    vol = W.volume_source()
    tol, niter = float(fx["tol"]), int(fx["niter"])

    x, res, it = W.cg(g, proj, vol, 3 * niter, tol)
    res = np.asarray(res, dtype=np.float64)
    dev = float(np.max(np.abs(res / fx["residual"] - 1.))) if len(res) == niter + 1 else float("nan")
    ref_planes = np.concatenate([fx["slice_y"].ravel(), fx["slice_z"].ravel()])
    zero_planes = _planes(x)
    err_zero = {"planes": rel_l2(zero_planes, ref_planes), "block_means": rel_l2(_block_means(x), fx["block_means"])}

    net = WarmStartCNN.load(default_weights_path())
    x0 = W.fdk(g, proj, vol)
    xf, resf, itf = W.cg(g, proj, x0, 3 * niter, tol)
    xc, resc, itc = W.cg(g, proj, x0.like(net(x0.tensor)), 3 * niter, tol)
    d_fdk, d_cnn, d_zero = _rms(_planes(xf), ref_planes), _rms(_planes(xc), ref_planes), _rms(zero_planes, ref_planes)
    truth = hip.draw_quadrics(vol, q)
    t_planes = _planes(truth)
    print(f"oracle stops at {niter} (residual {fx['residual'][-1]:.6g}); zero start {it} "
          f"(residual history max rel dev {dev:.3g}), vs oracle {err_zero}; "
          f"FDK start {itf}, CNN start {itc}; RMS distance to the oracle's solution on the central planes: "
          f"zero {d_zero:.3g}, FDK {d_fdk:.3g}, CNN {d_cnn:.3g} mm^-1; to the phantom: oracle "
          f"{_rms(ref_planes, t_planes):.4g}, CNN {_rms(_planes(xc), t_planes):.4g}, FDK {_rms(_planes(xf), t_planes):.4g}")
    # plain CG is the oracle's algorithm: same stopping iteration, residual history
    # within 1e-4, end point within 1e-3 on the planes (high-frequency noise fitted by
    # 35 iterations, where FP32 rounding of the projectors is amplified) and 1e-4 on the
    # smooth block means
    assert it == niter
    assert dev <= 1e-4
    assert err_zero["planes"] <= 1e-3 and err_zero["block_means"] <= 1e-4
    # the warm starts end anywhere in the tolerance's slack around the solution; the
    # learned one in fewer iterations and at most half as far from the oracle's solution
    # as the FDK start, and within 0.004 mm^-1 RMS (about a third of the mean
    # attenuation, 0.0125 mm^-1)
    assert resc[-1] <= tol and itc < itf < it
    assert d_cnn <= 0.5 * d_fdk
    assert d_cnn <= 4e-3
