"""HIP Joseph backprojector (BP_JOSEPH) vs the oracle, the adjoint identity on the
GPU (the pair the CG operator uses), and CG with the matched pair vs the double
oracle.  FP32 atomics: agreement to rounding, not bitwise."""
import numpy as np
import pytest

from helpers import circular_rows, hip_geometry, rel_l2, to_device

pytestmark = pytest.mark.gpu

SMALL = dict(vsize=(37, 33, 29), vsp=(3., 2.5, 3.5), vorig=(-53., -40., -49.), nu=61, nv=47,
             sorig=(-96., -73., 0.), ssp=(3.2, 3.2, 1.))
# Full 16 x 16 detector tiles at the clinical shape: views on the axes, on the
# diagonals (tiles split between x- and z-dominant rays) and in between.
FULL = dict(vsize=(128, 128, 128), vsp=(1., 1., 1.), vorig=(-63.5,) * 3, nu=256, nv=256,
            sorig=(-102., -102., 0.), ssp=(0.8, 0.8, 1.))
GEOMS = {
    "offsets": (np.array([[600., 1000., 22.5 * i + 1, 2.5 * i - 4, -3. + 0.5 * i, 0, 0, 5. - 0.3 * i, 2.]
                          for i in range(16)]), SMALL),
    "inner_tilted": (np.array([[30., 1000., 22.5 * i, 0, 0, 3., 2., 0, 0] for i in range(16)]), SMALL),
    "clinical_tiles": (np.array([[1000., 1536., a, 0, 0, 0, 0, 0, 0]
                                 for a in (0., 45., 90., 135., 180., 225., 270., 315., 17., 62., 113., 331.)]),
                       FULL),
}


@pytest.mark.parametrize("gname", list(GEOMS))
def test_joseph_bp_parity_and_adjoint(oracle, hip, gname):
    rows, sz = GEOMS[gname]
    vsize, vsp, vorig = sz["vsize"], sz["vsp"], sz["vorig"]
    nu, nv, n = sz["nu"], sz["nv"], len(rows)
    sorig, ssp = sz["sorig"], sz["ssp"]
    rng = np.random.default_rng(21)
    p = rng.normal(size=(n, nv, nu)).astype(np.float32)
    v0 = rng.normal(size=vsize[::-1]).astype(np.float32)
    ov, os_ = oracle.image(vsize, vorig, vsp), oracle.image((nu, nv, n), sorig, ssp)
    ref = oracle.joseph_back(rows, ov, v0, os_, p)
    g = hip_geometry(hip, rows)
    f = hip.HipJosephBackProjectionImageFilter()
    f.SetInput(0, to_device(hip, v0, vorig, vsp))
    f.SetInput(1, to_device(hip, p, sorig, ssp))
    f.SetGeometry(g)
    f.Update()
    out = f.GetOutput().numpy()
    e = rel_l2(out - v0, ref - v0)
    assert e <= 1e-5, f"{gname}: rel L2 vs oracle {e:.3g}"
    # adjoint identity with the HIP forward projector
    x = rng.random(vsize[::-1]).astype(np.float32)
    fp = hip.HipForwardProjectionImageFilter()
    fp.SetInput(0, to_device(hip, np.zeros_like(p), sorig, ssp))
    fp.SetInput(1, to_device(hip, x, vorig, vsp))
    fp.SetGeometry(g)
    fp.Update()
    bp = hip.HipJosephBackProjectionImageFilter()
    bp.SetInput(0, to_device(hip, np.zeros_like(x), vorig, vsp))
    bp.SetInput(1, to_device(hip, p, sorig, ssp))
    bp.SetGeometry(g)
    bp.Update()
    a = float((fp.GetOutput().numpy().astype(np.float64) * p).sum())
    b = float((x.astype(np.float64) * bp.GetOutput().numpy()).sum())
    assert abs(a / b - 1) < 1e-5, (a, b)


def test_cg_joseph_pair_matches_double_oracle(oracle, hip):
    from rtk_amd import phantoms
    n = 60
    rows = np.array([[600., 1200., i * 360. / n, 0, 0, 0, 0, 0, 0] for i in range(n)])
    q = phantoms.ellipsoid((0, 0, 0), (90, 90, 90), None, 1.0)[None]
    g = hip_geometry(hip, rows)
    proj = hip.project_quadrics(g, hip.ConstantImageSource((64, 64, n), (8., 8., 8.), (-255.,) * 3, 0.), q)
    vol = hip.ConstantImageSource((64, 64, 64), (4., 4., 4.), (-127.,) * 3, 0.)
    f = hip.HipConjugateGradientConeBeamReconstructionFilter()
    f.SetInput(0, vol)
    f.SetInput(1, proj)
    f.SetGeometry(g)
    f.SetBackProjectionFilter(f.BP_JOSEPH)
    f.SetNumberOfIterations(8)
    f.Update()
    ov = oracle.image((64,) * 3, (-127.,) * 3, (4.,) * 3)
    os_ = oracle.image((64, 64, n), (-255.,) * 3, (8.,) * 3)
    xd, rd = oracle.conjugate_gradient(rows, ov, np.zeros((64,) * 3), os_, proj.numpy().astype(np.float64), 8,
                                       joseph_bp=True)
    assert rel_l2(f.GetOutput().numpy(), xd) <= 1e-4
    np.testing.assert_allclose(f.GetResiduals(), np.sqrt(rd / rd[0]), rtol=2e-3)
    # symmetric operator: the residual history decreases monotonically here
    assert np.all(np.diff(f.GetResiduals()) < 0)


def test_tables_follow_the_geometry_across_calls(oracle, hip):
    """The library keeps the per-view tables of the last call on a stream and reuses them
    when the next call passes the same arguments.  Backprojecting with geometry A, then
    B (the same sizes, other angles and offsets), then A again must give each geometry's
    own result, and the second A call the first's."""
    sz = FULL
    vsize, vsp, vorig = sz["vsize"], sz["vsp"], sz["vorig"]
    nu, nv, sorig, ssp = sz["nu"], sz["nv"], sz["sorig"], sz["ssp"]
    rows_a = np.array([[1000., 1536., a, 0, 0, 0, 0, 0, 0] for a in (0., 45., 100., 200., 290.)])
    rows_b = np.array([[1000., 1536., a + 7., 1.5, -2., 0, 0, 0, 0] for a in (0., 45., 100., 200., 290.)])
    rng = np.random.default_rng(8)
    p = rng.normal(size=(len(rows_a), nv, nu)).astype(np.float32)
    ov, os_ = oracle.image(vsize, vorig, vsp), oracle.image((nu, nv, len(rows_a)), sorig, ssp)
    zero = np.zeros(vsize[::-1], np.float32)
    outs = []
    for rows in (rows_a, rows_b, rows_a):
        f = hip.HipJosephBackProjectionImageFilter()
        f.SetInput(0, to_device(hip, zero, vorig, vsp))
        f.SetInput(1, to_device(hip, p, sorig, ssp))
        f.SetGeometry(hip_geometry(hip, rows))
        f.Update()
        outs.append(f.GetOutput().numpy())
    for rows, out in zip((rows_a, rows_b, rows_a), outs):
        ref = oracle.joseph_back(rows, ov, zero, os_, p)
        assert rel_l2(out, ref) <= 1e-5
    assert rel_l2(outs[2], outs[0]) <= 1e-6
    assert rel_l2(outs[1], outs[0]) > 0.1
