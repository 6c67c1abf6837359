"""Pins the oracle's geometry against the closed-form matrices of
documentation/docs/Geometry.md:60-140 and geometric invariants, and checks the
product's host-side geometry (librtkhip, no GPU needed) against the oracle."""
import numpy as np
import pytest

from helpers import circular_rows, hip_geometry


def _rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def doc_matrix(sid, sdd, ga, px, py, oop, ip, sx, sy):
    """Final matrix of Geometry.md: M_P = T(s-p) Mag T(-s) M_R (angles in degrees)."""
    d = np.pi / 180
    R = np.eye(4)
    R[:3, :3] = _rz(-ip * d) @ _rx(-oop * d) @ _ry(-ga * d)
    T2 = np.array([[1, 0, sx - px], [0, 1, sy - py], [0, 0, 1.]])
    Mag = np.array([[-sdd, 0, 0, 0], [0, -sdd, 0, 0], [0, 0, 1, -sid]], dtype=float)
    Ts = np.eye(4)
    Ts[0, 3], Ts[1, 3] = -sx, -sy
    return T2 @ Mag @ Ts @ R


def random_rows(n, seed):
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n):
        rows.append([rng.uniform(500, 1200), rng.uniform(1000, 1600), rng.uniform(-400, 400),
                     rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(-10, 10), rng.uniform(-30, 30),
                     rng.uniform(-15, 15), rng.uniform(-15, 15)])
    return np.array(rows)


def test_matrices_match_documentation(oracle):
    rows = random_rows(12, 3)
    m, *_ = oracle.geometry(rows)
    for k, r in enumerate(rows):
        np.testing.assert_allclose(m[k], doc_matrix(*r), rtol=1e-12, atol=1e-9)


def test_source_is_centre_of_projection(oracle):
    rows = random_rows(10, 4)
    m, s, *_ = oracle.geometry(rows)
    for k in range(len(rows)):
        h = m[k] @ np.append(s[k], 1.)
        assert abs(h[2]) < 1e-9 * rows[k, 0]
        # source at distance sid from the isocentre when there is no source offset
    rows[:, 7:] = 0
    _, s, *_ = oracle.geometry(rows)
    np.testing.assert_allclose(np.linalg.norm(s, axis=1), rows[:, 0], rtol=1e-12)


def test_circular_invariants(oracle):
    n = 36
    rows = circular_rows(n, 1000., 1536.)
    m, s, ang, gaps, tilt = oracle.geometry(rows)
    np.testing.assert_allclose(gaps, 2 * np.pi / n, rtol=1e-12)
    t = np.mod(tilt + np.pi, 2 * np.pi) - np.pi
    np.testing.assert_allclose(t, 0., atol=1e-12)
    g = np.deg2rad(rows[:, 2])
    np.testing.assert_allclose(np.cos(ang), np.cos(-g), atol=1e-12)
    np.testing.assert_allclose(np.sin(ang), np.sin(-g), atol=1e-12)
    # Isocentre projects to the detector origin; a point at x=1 mm (gantry 0)
    # projects to u = sdd/sid (magnification).
    for k in range(n):
        h = m[k] @ np.array([0, 0, 0, 1.])
        assert abs(h[0] / h[2]) < 1e-9 and abs(h[1] / h[2]) < 1e-9
    h = m[0] @ np.array([1., 0, 0, 1])
    assert abs(h[0] / h[2] - 1536. / 1000.) < 1e-12


def test_angular_gaps_unsorted_and_uneven(oracle):
    ga = np.array([0., 200., 10., 90., 350.])
    rows = np.array([[1000, 1500, a, 0, 0, 0, 0, 0, 0] for a in ga], float)
    _, _, ang, gaps, _ = oracle.geometry(rows)
    # source angles are -gantry mod 360; gaps are half the distance between neighbours
    sa = np.mod(-ga, 360.)
    order = np.argsort(sa)
    s = sa[order]
    ext = np.concatenate([[s[-1] - 360.], s, [s[0] + 360.]])
    ref = np.empty(len(ga))
    ref[order] = 0.5 * (ext[2:] - ext[:-2])
    np.testing.assert_allclose(np.rad2deg(gaps), ref, rtol=1e-12)


def test_product_geometry_matches_oracle(oracle):
    import rtk_amd
    rows = random_rows(16, 5)
    g = hip_geometry(rtk_amd, rows)
    m, s, ang, _, _ = oracle.geometry(rows)
    assert len(g) == 16
    np.testing.assert_allclose(g.GetMatrices(), m, rtol=1e-12, atol=1e-9)
    for k in range(16):
        np.testing.assert_allclose(g.GetSourcePosition(k), s[k], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(g.GetSourceAngles(), ang, atol=1e-12)


def test_product_geometry_rejects_mixed_parallel(oracle):
    import rtk_amd
    g = rtk_amd.ThreeDCircularProjectionGeometry()
    g.AddProjection(1000., 1500., 0.)
    with pytest.raises(rtk_amd.RtkHipError) as e:
        g.AddProjection(1000., 0., 10.)
    assert e.value.code == 2
