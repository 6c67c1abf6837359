"""Host-side noise model of the benchmark inputs (rtk_amd.phantoms.poisson_line_integrals):
per-pixel Poisson draws by inversion.  Checks: the inversion is the exact Poisson
quantile (scipy.stats.poisson.ppf, an independent implementation) over six decades of
mean counts; moments of the draws; a pixel's count depends only on its own line
integral and variate (moving a few pixels leaves every other count unchanged, the
property the GPU test relies on to reproduce the oracle fixture's input); and the
line integrals are -ln(max(c, 1) / I0)."""
import numpy as np
from scipy import stats

from rtk_amd import phantoms


def test_inversion_is_the_exact_poisson_quantile():
    rng = np.random.default_rng(11)
    lam = 10. ** rng.uniform(-2, 4.3, size=100000)
    u = rng.random(lam.size)
    k = phantoms._poisson_inverse(lam.copy(), u)
    ref = stats.poisson.ppf(u, lam).astype(np.int64)
    assert np.array_equal(k, ref)


def test_moments():
    for lam in (0.5, 20., 5000.):
        p = np.full(200000, -np.log(lam / 1e4), np.float32)
        _, c = phantoms.poisson_line_integrals(p, 1e4, 3, return_counts=True, threads=4)
        m = float(1e4 * np.exp(-np.float64(p[0])))
        se = np.sqrt(m / c.size)
        assert abs(c.mean() - m) < 5 * se, (lam, c.mean())
        assert abs(c.var() / m - 1.) < 0.03, (lam, c.var())


def test_a_pixel_draw_depends_only_on_that_pixel():
    p = np.random.default_rng(4).uniform(0., 4., size=(6, 50, 70)).astype(np.float32)
    a, ca = phantoms.poisson_line_integrals(p, 1e4, 9, return_counts=True, threads=3)
    q = p.copy()
    moved = [(0, 3, 5), (2, 40, 1), (5, 49, 69)]
    for i in moved:
        q[i] = np.nextafter(q[i], np.float32(10.))
    b, cb = phantoms.poisson_line_integrals(q, 1e4, 9, return_counts=True, threads=5)
    mask = np.ones(p.shape, bool)
    for i in moved:
        mask[i] = False
    assert np.array_equal(ca[mask], cb[mask])
    np.testing.assert_array_equal(a, (-np.log(np.maximum(ca, 1) / 1e4)).astype(np.float32))
