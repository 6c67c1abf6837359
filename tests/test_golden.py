"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
double-precision oracle).

CPU: the oracle still reproduces every fixture (double: to rounding; float: within
the north-star bars), so a change of the restatement cannot go unnoticed.
GPU: the HIP path against the stored double results, with no oracle needed at run
time: Joseph projections and all backprojections within 1e-5 relative L2, the FDK
volume within 1e-4, CG with the matched Joseph pair within 1e-4."""
import os

import numpy as np
import pytest

from helpers import hip_geometry, rel_l2, to_device

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def images(oracle, d):
    return (oracle.image(tuple(d["vol_size"]), tuple(d["vol_origin"]), tuple(d["vol_spacing"])),
            oracle.image(tuple(d["stack_size"]), tuple(d["stack_origin"]), tuple(d["stack_spacing"])))


def zeros_vol(d, dtype=np.float64):
    return np.zeros(tuple(d["vol_size"])[::-1], dtype)


def zeros_stack(d, dtype=np.float64):
    s = d["stack_size"]
    return np.zeros((s[2], s[1], s[0]), dtype)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_oracle_reproduces_projector_fixtures(oracle, dtype, tol):
    d = load("joseph_forward")
    ov, os_ = images(oracle, d)
    fp = oracle.joseph(d["rows"], os_, zeros_stack(d, dtype), ov, d["volume"].astype(dtype))
    assert rel_l2(fp, d["expected"]) <= tol
    b = load("backprojections")
    ov, os_ = images(oracle, b)
    p = b["proj"].astype(dtype)
    assert rel_l2(oracle.joseph_back(b["rows"], ov, zeros_vol(b, dtype), os_, p), b["expected_joseph"]) <= tol
    assert rel_l2(oracle.backproject(b["rows"], ov, zeros_vol(b, dtype), os_, p), b["expected_voxel"]) <= tol
    assert rel_l2(oracle.backproject(b["rows"], ov, zeros_vol(b, dtype), os_, p, fdk=True),
                  b["expected_fdk"]) <= tol


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-4)])
def test_oracle_reproduces_fdk_and_cg_fixtures(oracle, dtype, tol):
    d = load("fdk_shepp_logan")
    ov, os_ = images(oracle, d)
    rec = oracle.fdk(d["rows"], ov, zeros_vol(d, dtype), os_, d["proj"].astype(dtype), hann=float(d["hann"]))
    assert rel_l2(rec, d["expected"]) <= tol
    c = load("cg_joseph_pair")
    ov, os_ = images(oracle, c)
    x, res = oracle.conjugate_gradient(c["rows"], ov, zeros_vol(c, dtype), os_, c["proj"].astype(dtype),
                                       int(c["niter"]), joseph_bp=True)
    assert rel_l2(x, c["expected"]) <= tol
    np.testing.assert_allclose(np.sqrt(res / res[0]), c["expected_residuals"], rtol=max(tol, 1e-10))


def _hip_images(hip, d, vol, proj):
    return (to_device(hip, vol, tuple(d["vol_origin"]), tuple(d["vol_spacing"])),
            to_device(hip, proj, tuple(d["stack_origin"]), tuple(d["stack_spacing"])))


@pytest.mark.gpu
def test_hip_projectors_match_fixtures(hip):
    d = load("joseph_forward")
    vol, stack = _hip_images(hip, d, d["volume"], zeros_stack(d, np.float32))
    f = hip.HipForwardProjectionImageFilter()
    f.SetInput(0, stack)
    f.SetInput(1, vol)
    f.SetGeometry(hip_geometry(hip, d["rows"]))
    f.Update()
    e = rel_l2(f.GetOutput().numpy(), d["expected"])
    assert e <= 1e-5, f"Joseph forward: {e:.3g}"
    b = load("backprojections")
    g = hip_geometry(hip, b["rows"])
    for cls, key in ((hip.HipJosephBackProjectionImageFilter, "expected_joseph"),
                     (hip.HipBackProjectionImageFilter, "expected_voxel"),
                     (hip.HipFDKBackProjectionImageFilter, "expected_fdk")):
        vol, stack = _hip_images(hip, b, zeros_vol(b, np.float32), b["proj"])
        f = cls()
        f.SetInput(0, vol)
        f.SetInput(1, stack)
        f.SetGeometry(g)
        f.Update()
        e = rel_l2(f.GetOutput().numpy(), b[key])
        assert e <= 1e-5, f"{cls.__name__}: {e:.3g}"


@pytest.mark.gpu
def test_hip_fdk_and_cg_match_fixtures(hip):
    d = load("fdk_shepp_logan")
    vol, stack = _hip_images(hip, d, zeros_vol(d, np.float32), d["proj"])
    f = hip.HipFDKConeBeamReconstructionFilter()
    f.SetInput(0, vol)
    f.SetInput(1, stack)
    f.SetGeometry(hip_geometry(hip, d["rows"]))
    f.GetRampFilter().SetHannCutFrequency(float(d["hann"]))
    f.Update()
    e = rel_l2(f.GetOutput().numpy(), d["expected"])
    assert e <= 1e-4, f"FDK: {e:.3g}"
    c = load("cg_joseph_pair")
    vol, stack = _hip_images(hip, c, zeros_vol(c, np.float32), c["proj"])
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, vol)
    f.SetInput(1, stack)
    f.SetGeometry(hip_geometry(hip, c["rows"]))
    f.SetBackProjectionFilter(f.BP_JOSEPH)
    f.SetNumberOfIterations(int(c["niter"]))
    f.Update()
    e = rel_l2(f.GetOutput().numpy(), c["expected"])
    assert e <= 1e-4, f"CG: {e:.3g}"
    np.testing.assert_allclose(f.GetResiduals(), c["expected_residuals"], rtol=2e-3)
