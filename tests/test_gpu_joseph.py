"""HIP Joseph forward projection vs the oracle (RTK's CPU Joseph filter restated).
Bar: relative L2 <= 1e-5 against the float oracle, identical Joseph sample counts."""
import numpy as np
import pytest

from helpers import circular_rows, hip_geometry, rel_l2, to_device

pytestmark = pytest.mark.gpu

TOL = 1e-5

GEOMS = {
    # includes 45-degree views where the dominant axis switches within a view
    "circular": circular_rows(16, 600., 1000., start=0.),
    "offsets": np.array([[600., 1000., 22.5 * i + 1, 2.5 * i - 4, -3. + 0.5 * i, 0, 0, 5. - 0.3 * i, 2.]
                         for i in range(16)]),
    "tilted": np.array([[600., 1000., 22.5 * i, 1., -2., 6., 4., 0, 0] for i in range(16)]),
    # source inside the volume: rays clipped at the source
    "inner": circular_rows(16, 30., 1000., start=5.),
}


@pytest.mark.parametrize("gname", list(GEOMS))
def test_joseph_parity(oracle, hip, gname):
    rows = GEOMS[gname]
    vsize, vsp, vorig = (37, 33, 29), (3., 2.5, 3.5), (-53., -40., -49.)
    nu, nv, n = 61, 47, len(rows)
    ssp, sorig = (3.2, 3.2), (-96., -73.)
    rng = np.random.default_rng(11)
    vol = rng.random(vsize[::-1]).astype(np.float32)
    p0 = rng.normal(size=(n, nv, nu)).astype(np.float32)
    ov, os_ = oracle.image(vsize, vorig, vsp), oracle.image((nu, nv, n), sorig + (0.,), ssp + (1.,))
    ref, cnt = oracle.joseph(rows, os_, p0, ov, vol, return_samples=True)
    ref_d = oracle.joseph(rows, os_, p0.astype(np.float64), ov, vol.astype(np.float64))
    g = hip_geometry(hip, rows)
    f = hip.HipForwardProjectionImageFilter()
    f.SetInput(0, to_device(hip, p0, sorig + (0.,), ssp + (1.,)))
    f.SetInput(1, to_device(hip, vol, vorig, vsp))
    f.SetGeometry(g)
    f.Update(count_samples=True)
    out = f.GetOutput().numpy()
    e_f, e_d = rel_l2(out - p0, ref - p0), rel_l2(out - p0, ref_d - p0)
    assert e_f <= TOL, f"{gname}: rel L2 vs float oracle {e_f:.3g} (vs double {e_d:.3g})"
    assert f.GetNumberOfSamples() == cnt


def test_joseph_view_subset_and_volume_index(oracle, hip):
    rows = GEOMS["offsets"]
    vsize, vsp, vorig, vindex = (30, 28, 26), (3., 3., 3.), (-60., -50., -45.), (3, 5, 2)
    rng = np.random.default_rng(12)
    vol = rng.random(vsize[::-1]).astype(np.float32)
    oorig = tuple(o + i * s for o, i, s in zip(vorig, vindex, vsp))
    ov = oracle.image(vsize, oorig, vsp)
    # detector buffered region starts at (4, 6): fold into the origin for the oracle
    os_ = oracle.image((40, 30, 6), (-96. + 4 * 3.2, -73. + 6 * 3.2, 0.), (3.2, 3.2, 1.))
    ref = oracle.joseph(rows[7:13], os_, np.zeros((6, 30, 40), np.float32), ov, vol)
    g = hip_geometry(hip, rows)
    f = hip.HipForwardProjectionImageFilter()
    f.SetInput(0, to_device(hip, np.zeros((6, 30, 40)), (-96., -73., 0.), (3.2, 3.2, 1.), index=(4, 6, 7)))
    f.SetInput(1, to_device(hip, vol, vorig, vsp, index=vindex))
    f.SetGeometry(g)
    f.Update()
    assert rel_l2(f.GetOutput().numpy(), ref) <= TOL


def test_joseph_constant_volume_chords_large(oracle, hip):
    """Size-independent property at a large shape: a constant-1 volume projects to
    the chord of the voxel-centre box (mm), for every ray."""
    n = 24
    rows = circular_rows(n, 1000., 1536., start=1.)
    g = hip_geometry(hip, rows)
    vol = hip.ConstantImageSource((256, 256, 256), (1., 1., 1.), (-127.5,) * 3, 1.)
    proj = hip.ConstantImageSource((512, 512, n), (0.8, 0.8, 1.), (-204.4, -204.4, 0.), 0.)
    f = hip.HipForwardProjectionImageFilter()
    f.SetInput(0, proj)
    f.SetInput(1, vol)
    f.SetGeometry(g)
    f.Update()
    got = f.GetOutput().numpy()
    # analytic chords, vectorised slab test in physical space
    _, src, *_ = oracle.geometry(rows)
    from test_oracle_geometry import _rx, _ry, _rz
    u = -204.4 + 0.8 * np.arange(512)
    U, V = np.meshgrid(u, u)
    lo, hi = -127.5, 127.5
    for k in range(0, n, 5):
        r = rows[k]
        R = _rz(0) @ _rx(0) @ _ry(-np.deg2rad(r[2]))
        det = np.stack([U, V, np.full_like(U, r[0] - r[1])], -1) @ R  # R^T applied to row vectors
        d = det - src[k]
        t0 = np.zeros(U.shape)
        t1 = np.ones(U.shape)
        for a in range(3):
            with np.errstate(divide="ignore", invalid="ignore"):
                ta, tb = (lo - src[k][a]) / d[..., a], (hi - src[k][a]) / d[..., a]
            t0 = np.maximum(t0, np.minimum(ta, tb))
            t1 = np.minimum(t1, np.maximum(ta, tb))
        chord = np.maximum(0., t1 - t0) * np.linalg.norm(d, axis=-1)
        assert np.max(np.abs(got[k] - chord)) < 1e-4 * 256, k
        assert rel_l2(got[k], chord) < 1e-6


# A geometry the column gathers take (detector columns vertical in the volume, default
# clips, row step b = 0.9 mm x (distance from the source / 800 mm) between 0.51 and 0.62
# voxel over the box, inside the forward gather's (0.71) and the backprojection gather's
# (0.4 to 2/3) bounds), with detector and source offsets, views on the diagonals (columns
# whose rays switch dominant axis), and sizes that are not multiples of the workgroup
# tiles (16 columns, 64 rows).
GATHER_ROWS = np.array([[500., 800., 7. + 15. * i, 2.5 - 0.4 * i, -3. + 0.3 * i, 0., 0., 1.5, -2. + 0.2 * i]
                        for i in range(24)])
GATHER_VOL = dict(size=(64, 48, 56), spacing=(1., 1., 1.), origin=(-31., -24.5, -27.))
GATHER_DET = dict(size=(120, 100), spacing=(0.9, 0.9), origin=(-54.5, -44.))


def _gather_case(oracle, hip, seed):
    rows, v, d = GATHER_ROWS, GATHER_VOL, GATHER_DET
    n = len(rows)
    rng = np.random.default_rng(seed)
    vol = rng.random(v["size"][::-1]).astype(np.float32)
    p0 = rng.normal(size=(n, d["size"][1], d["size"][0])).astype(np.float32)
    ov = oracle.image(v["size"], v["origin"], v["spacing"])
    os_ = oracle.image(d["size"] + (n,), d["origin"] + (0.,), d["spacing"] + (1.,))
    return rows, v, d, vol, p0, ov, os_


def test_joseph_gather_forward_parity(oracle, hip):
    """The forward gather (a call that does not count samples) against the float and
    double oracles, with a non-zero input stack, and against the per-ray projector (the
    call that counts samples) on the same inputs."""
    rows, v, d, vol, p0, ov, os_ = _gather_case(oracle, hip, 31)
    ref = oracle.joseph(rows, os_, p0, ov, vol)
    ref_d = oracle.joseph(rows, os_, p0.astype(np.float64), ov, vol.astype(np.float64))
    g = hip_geometry(hip, rows)
    outs = []
    for count in (False, True):
        f = hip.HipForwardProjectionImageFilter()
        f.SetInput(0, to_device(hip, p0, d["origin"] + (0.,), d["spacing"] + (1.,)))
        f.SetInput(1, to_device(hip, vol, v["origin"], v["spacing"]))
        f.SetGeometry(g)
        f.Update(count_samples=count)
        outs.append(f.GetOutput().numpy())
    gather, per_ray = outs
    assert rel_l2(gather - p0, ref - p0) <= TOL
    assert rel_l2(gather - p0, ref_d - p0) <= TOL
    assert rel_l2(gather - p0, per_ray - p0) <= TOL
    # every view, including the diagonal ones
    for k in range(len(rows)):
        assert rel_l2(gather[k] - p0[k], ref[k] - p0[k]) <= 3 * TOL, k


def test_joseph_gather_backprojection_and_adjoint(oracle, hip):
    """The backprojection gather on the same geometry against the oracle, and the
    adjoint identity <FP x, p> = <x, BP p> between the two gathers."""
    rows, v, d, vol, p0, ov, os_ = _gather_case(oracle, hip, 32)
    zero = np.zeros(v["size"][::-1], np.float32)
    ref = oracle.joseph_back(rows, ov, zero, os_, p0)
    g = hip_geometry(hip, rows)
    bp = hip.HipJosephBackProjectionImageFilter()
    bp.SetInput(0, to_device(hip, zero, v["origin"], v["spacing"]))
    bp.SetInput(1, to_device(hip, p0, d["origin"] + (0.,), d["spacing"] + (1.,)))
    bp.SetGeometry(g)
    bp.Update()
    back = bp.GetOutput().numpy()
    assert rel_l2(back, ref) <= TOL
    fp = hip.HipForwardProjectionImageFilter()
    fp.SetInput(0, to_device(hip, np.zeros_like(p0), d["origin"] + (0.,), d["spacing"] + (1.,)))
    fp.SetInput(1, to_device(hip, vol, v["origin"], v["spacing"]))
    fp.SetGeometry(g)
    fp.Update()
    a = float((fp.GetOutput().numpy().astype(np.float64) * p0).sum())
    b = float((vol.astype(np.float64) * back).sum())
    assert abs(a / b - 1) < 1e-5, (a, b)
