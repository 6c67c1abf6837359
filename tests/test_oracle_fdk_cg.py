"""End-to-end pins of the oracle's FDK and CG against the reference's own
integration tests and CG identities.

* FDK: test/rtkfdktest.cxx:30-117 (Shepp-Logan scale 116, 180 views, source
  offsets (20, 15)); the reference masks with FieldOfViewImageFilter, here the
  field of view is the set of voxels that every view sees (constant-1
  backprojection count == number of views).  Thresholds 0.03 / 26 dB (ref 2).
* CG: test/rtkconjugategradientreconstructiontest.cxx:99-118 (ellipsoid of
  axis 90, 5 iterations, Joseph FP + voxel BP), thresholds 0.08 / 23 dB.
* CG identity: X_1 = alpha_0 B with alpha_0 = B.B / B.(A B) from X_0 = 0.
* Row-band restriction (used by the full-size GPU parity tests): the FDK of a few
  y slices from only the detector rows they reach equals the same slices of the
  FDK from the whole stack, to the rounding of the pixel coordinates (the band's
  origin moves by r0 rows); narrowing the band changes them.
"""
import numpy as np

from helpers import check_image_quality, row_band


def test_fdk_shepp_logan(oracle):
    from rtk_amd import phantoms
    n = 180
    rows = np.array([[600., 1200., i * 360. / n, 0, 0, 0, 0, 20, 15] for i in range(n)])
    st = oracle.image((128, 128, n), (-254, -254, -254), (4, 4, 4))
    q = phantoms.shepp_logan(116.)
    proj = oracle.project_quadrics(rows, st, np.zeros((n, 128, 128), np.float32), q)
    vol = oracle.image((128, 128, 128), (-127,) * 3, (2,) * 3)
    ref = oracle.draw_quadrics(vol, np.zeros((128, 128, 128), np.float32), q)
    rec = oracle.fdk(rows, vol, np.zeros((128, 128, 128), np.float32), st, proj)
    cnt = oracle.backproject(rows, vol, np.zeros((128, 128, 128), np.float32), st, np.ones_like(proj))
    fov = cnt > n - 0.5
    assert fov.sum() > 0.3 * fov.size
    err, psnr, ok = check_image_quality(np.where(fov, rec, 0.), np.where(fov, ref, 0.), 0.03, 26., 2.)
    assert ok, (err, psnr)


def _cg_problem(oracle, n=180):
    from rtk_amd import phantoms
    rows = np.array([[600., 1200., i * 360. / n, 0, 0, 0, 0, 0, 0] for i in range(n)])
    st = oracle.image((64, 64, n), (-255, -255, -255), (8, 8, 8))
    vol = oracle.image((64, 64, 64), (-127,) * 3, (4,) * 3)
    q = phantoms.ellipsoid((0, 0, 0), (90, 90, 90), None, 1.0)[None]
    return rows, st, vol, q


def test_cg_ellipsoid(oracle):
    rows, st, vol, q = _cg_problem(oracle)
    proj = oracle.project_quadrics(rows, st, np.zeros((180, 64, 64), np.float32), q)
    ref = oracle.draw_quadrics(vol, np.zeros((64, 64, 64), np.float32), q)
    x, res = oracle.conjugate_gradient(rows, vol, np.zeros((64, 64, 64), np.float32), st, proj, 5,
                                       weights=np.ones_like(proj))
    err, psnr, ok = check_image_quality(x, ref, 0.08, 23., 2.)
    assert ok, (err, psnr)
    assert res[-1] < res[0]


def test_cg_first_step_identity(oracle):
    rows, st, vol, q = _cg_problem(oracle, n=30)
    proj = oracle.project_quadrics(rows, st, np.zeros((30, 64, 64)), q)
    z = np.zeros((64, 64, 64))
    b = oracle.backproject(rows, vol, z, st, proj)
    ab = oracle.backproject(rows, vol, z, st, oracle.joseph(rows, st, np.zeros_like(proj), vol, b))
    alpha = (b * b).sum() / (b * ab).sum()
    x1, res = oracle.conjugate_gradient(rows, vol, z, st, proj, 1)
    np.testing.assert_allclose(x1, alpha * b, rtol=1e-10, atol=1e-12 * np.abs(alpha * b).max())
    np.testing.assert_allclose(res[0], (b * b).sum(), rtol=1e-12)
    r1 = b - alpha * ab
    np.testing.assert_allclose(res[1], (r1 * r1).sum(), rtol=1e-9)


def test_fdk_row_band_equals_full_stack(oracle):
    from rtk_amd import phantoms
    nx, nu, n, vsp, psp = 64, 128, 45, 4., 3.2
    vorg, porg = -(nx - 1) * vsp / 2, -(nu - 1) * psp / 2
    rng = np.random.default_rng(4)
    offs = rng.uniform(-2., 2., size=(n, 2))
    rows = np.array([[1000., 1536., i * 360. / n, offs[i, 0], offs[i, 1], 0, 0, 0, 0] for i in range(n)])
    q = phantoms.random_ellipsoids(4, n=20, half_fov=nx * vsp / 2)
    st = oracle.image((nu, nu, n), (porg, porg, 0.), (psp, psp, 1.))
    proj = oracle.project_quadrics(rows, st, np.zeros((n, nu, nu), np.float32), q)
    full = oracle.fdk(rows, oracle.image((nx,) * 3, (vorg,) * 3, (vsp,) * 3), np.zeros((nx,) * 3, np.float32), st,
                      proj)
    for j in (0, 31, 50):
        y0 = vorg + j * vsp
        r0, r1 = row_band(oracle, rows, (vorg, y0, vorg), (-vorg, y0 + vsp, -vorg), porg, psp, nu)
        assert r1 - r0 < nu // 2
        outs = []
        for a, b in ((r0, r1), (r0 + 3, r1 - 3)):
            sb = oracle.image((nu, b - a, n), (porg, porg + a * psp, 0.), (psp, psp, 1.))
            outs.append(oracle.fdk(rows, oracle.image((nx, 2, nx), (vorg, y0, vorg), (vsp,) * 3),
                                   np.zeros((nx, 2, nx), np.float32), sb, np.ascontiguousarray(proj[:, a:b])))
        scale = np.abs(full).max()
        assert np.abs(outs[0] - full[:, j:j + 2]).max() <= 1e-7 * scale
        assert np.abs(outs[1] - full[:, j:j + 2]).max() > 1e-3 * scale
