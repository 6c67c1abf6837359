"""HIP conjugate gradient (FP64 scalars) vs the oracle, and the phantom generators.

CG bar: the HIP iterate agrees with the double oracle (RTK's CG on double
pixels) within 1e-4 relative L2 after 5 iterations of the reference's CG test
case (test/rtkconjugategradientreconstructiontest.cxx:99-118), meets that test's
0.08 / 23 dB thresholds, and the residual history tracks the oracle's."""
import numpy as np
import pytest

from helpers import check_image_quality, hip_geometry, rel_l2, to_device

pytestmark = pytest.mark.gpu


def _problem(hip, n=180):
    from rtk_amd import phantoms
    rows = np.array([[600., 1200., i * 360. / n, 0, 0, 0, 0, 0, 0] for i in range(n)])
    q = phantoms.ellipsoid((0, 0, 0), (90, 90, 90), None, 1.0)[None]
    g = hip_geometry(hip, rows)
    proj = hip.project_quadrics(g, hip.ConstantImageSource((64, 64, n), (8., 8., 8.), (-255.,) * 3, 0.), q)
    return rows, q, g, proj


def test_cg_matches_double_oracle(oracle, hip):
    rows, q, g, proj = _problem(hip)
    vol = hip.ConstantImageSource((64, 64, 64), (4., 4., 4.), (-127.,) * 3, 0.)
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, vol)
    f.SetInput(1, proj)
    f.SetGeometry(g)
    f.SetNumberOfIterations(5)
    f.Update()
    x = f.GetOutput().numpy()
    ov = oracle.image((64,) * 3, (-127.,) * 3, (4.,) * 3)
    os_ = oracle.image((64, 64, 180), (-255.,) * 3, (8.,) * 3)
    p = proj.numpy()
    xd, rd = oracle.conjugate_gradient(rows, ov, np.zeros((64,) * 3), os_, p.astype(np.float64), 5)
    xf, rf = oracle.conjugate_gradient(rows, ov, np.zeros((64,) * 3, np.float32), os_, p, 5)
    e_d, e_f = rel_l2(x, xd), rel_l2(xf, xd)
    assert e_d <= 1e-4, f"HIP vs double oracle {e_d:.3g} (float oracle vs double {e_f:.3g})"
    res = f.GetResiduals()
    np.testing.assert_allclose(res, np.sqrt(rd / rd[0]), rtol=1e-3)
    phantom = oracle.draw_quadrics(ov, np.zeros((64,) * 3, np.float32), q)
    err, psnr, ok = check_image_quality(x, phantom, 0.08, 23., 2.)
    assert ok, (err, psnr)


def test_cg_weights_and_tolerance_stop(oracle, hip):
    rows, q, g, proj = _problem(hip, n=60)
    vol = hip.ConstantImageSource((64, 64, 64), (4., 4., 4.), (-127.,) * 3, 0.)
    w = proj.like(proj.tensor * 0 + 2.0)
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, vol)
    f.SetInput(1, proj)
    f.SetInputWeights(w)
    f.SetGeometry(g)
    f.SetNumberOfIterations(40)
    f.SetResidualTolerance(0.05)
    f.Update()
    res = f.GetResiduals()
    assert res[-1] <= 0.05 and f.GetNumberOfIterationsDone() < 40
    assert np.all(res[1:-1] > 0.05)
    ov = oracle.image((64,) * 3, (-127.,) * 3, (4.,) * 3)
    os_ = oracle.image((64, 64, 60), (-255.,) * 3, (8.,) * 3)
    p = proj.numpy().astype(np.float64)
    it = f.GetNumberOfIterationsDone()
    xd, _ = oracle.conjugate_gradient(rows, ov, np.zeros((64,) * 3), os_, p, it, weights=np.full_like(p, 2.0))
    assert rel_l2(f.GetOutput().numpy(), xd) <= 1e-4


def test_phantom_generators(oracle, hip):
    from rtk_amd import phantoms
    q = np.concatenate([phantoms.shepp_logan(60.), phantoms.random_ellipsoids(3, n=6, half_fov=60.)])
    rows = np.array([[500., 900., 37. * i, 1., -2., 0, 0, 3., 1.] for i in range(9)])
    g = hip_geometry(hip, rows)
    proj = hip.project_quadrics(g, hip.ConstantImageSource((50, 40, 9), (3., 3., 1.), (-75., -60., 0.), 0.), q)
    os_ = oracle.image((50, 40, 9), (-75., -60., 0.), (3., 3., 1.))
    ref = oracle.project_quadrics(rows, os_, np.zeros((9, 40, 50), np.float32), q)
    assert rel_l2(proj.numpy(), ref) <= 1e-6
    vol = hip.draw_quadrics(hip.ConstantImageSource((40, 36, 30), (3., 3., 4.), (-60., -52., -58.), 0.), q)
    ov = oracle.image((40, 36, 30), (-60., -52., -58.), (3., 3., 4.))
    vref = oracle.draw_quadrics(ov, np.zeros((30, 36, 40), np.float32), q)
    assert np.mean(vol.numpy() != vref) < 1e-4


@pytest.mark.parametrize("joseph_bp", [False, True])
def test_cg_warm_start_matches_double_oracle(oracle, hip, joseph_bp):
    """Non-zero start: R0 = B - A X0 takes the projector pair (a cold start skips it,
    A 0 = 0); the iterate and the residual history follow the double oracle from the
    same X0, with either backprojector."""
    rows, q, g, proj = _problem(hip, n=60)
    ov = oracle.image((64,) * 3, (-127.,) * 3, (4.,) * 3)
    rng = np.random.default_rng(31)
    x0 = (0.5 * oracle.draw_quadrics(ov, np.zeros((64,) * 3, np.float32), q)
          + 0.05 * rng.normal(size=(64,) * 3)).astype(np.float32)
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, to_device(hip, x0, (-127.,) * 3, (4.,) * 3))
    f.SetInput(1, proj)
    f.SetGeometry(g)
    if joseph_bp:
        f.SetBackProjectionFilter(f.BP_JOSEPH)
    f.SetNumberOfIterations(3)
    f.Update()
    os_ = oracle.image((64, 64, 60), (-255.,) * 3, (8.,) * 3)
    xd, rd = oracle.conjugate_gradient(rows, ov, x0.astype(np.float64), os_, proj.numpy().astype(np.float64), 3,
                                       joseph_bp=joseph_bp)
    assert rel_l2(f.GetOutput().numpy(), xd) <= 1e-4
    res = f.GetResiduals()
    # the HIP residuals are relative to ||B||; the oracle's are squared norms of R_k
    np.testing.assert_allclose(np.asarray(res) / res[0], np.sqrt(rd / rd[0]), rtol=1e-3)
