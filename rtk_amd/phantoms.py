"""Analytic phantoms as quadric coefficient rows (A..J, density) for the phantom
generators of librtkhip (and the oracle).

* ``shepp_logan`` restates rtk::SheppLoganPhantom (src/rtkSheppLoganPhantom.cxx:23-35):
  ten ellipsoids, each rotated about y by its angle (QuadricShape::SetEllipsoid,
  src/rtkQuadricShape.cxx:259-317), then scaled by PhantomScale (default 128,
  include/rtkSheppLoganPhantomFilter.hxx:29) as ProjectGeometricPhantom does
  (rotate -> translate -> rescale, include/rtkProjectGeometricPhantomImageFilter.hxx:71-74).
* ``random_ellipsoids`` is the seeded random-ellipsoid phantom of the benchmark
  workloads: a 0.02 mm^-1 base sphere plus ellipsoids with random centres, semi-axes,
  orientations and density offsets (the distributions are in its docstring; they are
  this project's definition, RTK has no such generator).
"""
import numpy as np

SHEPP_LOGAN = [
    # semi-axes (x, y, z), centre (x, y, z), angle about y (deg), density
    (0.69, 0.90, 0.92, 0., 0., 0., 0., 2.),
    (0.6624, 0.880, 0.874, 0., 0., 0., 0., -0.98),
    (0.41, 0.21, 0.16, -0.22, -0.25, 0., 108., -0.02),
    (0.31, 0.22, 0.11, 0.22, -0.25, 0., 72., -0.02),
    (0.21, 0.50, 0.25, 0., -0.25, 0.35, 0., 0.02),
    (0.046, 0.046, 0.046, 0., -0.25, 0.10, 0., 0.02),
    (0.046, 0.02, 0.023, -0.08, -0.25, -0.65, 0., 0.01),
    (0.046, 0.02, 0.023, 0.06, -0.25, -0.65, 90., 0.01),
    (0.056, 0.1, 0.04, 0.06, 0.625, -0.105, 90., 0.02),
    (0.056, 0.1, 0.056, 0., 0.625, 0.1, 0., -0.02),
]


def quadric_from_form(M, g, J, density):
    """Row (A..J, density) of x^T M x + g.x + J <= 0 (RTK's A..J naming)."""
    M = np.asarray(M, dtype=np.float64)
    return np.array([M[0, 0], M[1, 1], M[2, 2], 2 * M[0, 1], 2 * M[0, 2], 2 * M[1, 2],
                     g[0], g[1], g[2], J, density], dtype=np.float64)


def ellipsoid(center, axes, rotation=None, density=1.0, scale=1.0):
    """Ellipsoid {x : |D^-1 R^T (x - c)| <= 1} (columns of R are the axes'
    directions), then the polynomial rescaled as f(x / scale) (QuadricShape::Rescale)."""
    R = np.eye(3) if rotation is None else np.asarray(rotation, dtype=np.float64)
    c = np.asarray(center, dtype=np.float64)
    a = np.asarray(axes, dtype=np.float64)
    M = R @ np.diag(1.0 / a ** 2) @ R.T
    g = -2.0 * M @ c
    J = c @ M @ c - 1.0
    s = np.broadcast_to(np.asarray(scale, dtype=np.float64), (3,))
    Si = np.diag(1.0 / s)
    return quadric_from_form(Si @ M @ Si, Si @ g, J, density)


def _rot_y(deg):
    # The rotation QuadricShape::SetEllipsoid applies for a positive angle about y
    # corresponds to axis directions R = Ry(-angle) (E = 2 cos sin (A - C)).
    t = np.deg2rad(deg)
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, 0., -s], [0., 1., 0.], [s, 0., c]])


def shepp_logan(scale=128.0):
    """Ten quadric rows of rtk::SheppLoganPhantom scaled by PhantomScale."""
    rows = []
    for ax, ay, az, cx, cy, cz, ang, rho in SHEPP_LOGAN:
        rows.append(ellipsoid((cx, cy, cz), (ax, ay, az), _rot_y(ang), rho, scale))
    return np.stack(rows)


def _random_rotation(rng):
    u1, u2, u3 = rng.random(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def random_ellipsoids(seed, n=20, half_fov=128.0, base=0.02, base_radius=0.9, centre_radius=0.6,
                      axes_range=(0.05, 0.40), density_range=(-0.005, 0.02)):
    """Seeded random-ellipsoid phantom (attenuation in mm^-1):

    * a base sphere of radius ``base_radius * half_fov`` and density ``base``;
    * ``n`` ellipsoids with centres uniform in the ball of radius
      ``centre_radius * half_fov``, semi-axes uniform in ``axes_range * half_fov``,
      uniformly random orientations and density offsets uniform in ``density_range``.
    """
    rng = np.random.default_rng(seed)
    rows = [ellipsoid((0., 0., 0.), (base_radius * half_fov,) * 3, None, base)]
    for _ in range(n):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        c = d * centre_radius * half_fov * rng.random() ** (1.0 / 3.0)
        a = rng.uniform(axes_range[0], axes_range[1], size=3) * half_fov
        rows.append(ellipsoid(c, a, _random_rotation(rng), rng.uniform(*density_range)))
    return np.stack(rows)


def _poisson_inverse(lam, u):
    """Smallest k with P(Poisson(lam) <= k) >= u, element-wise: a Cornish-Fisher start
    (lam + sqrt(lam) z + (z^2 - 1) / 6, z = Phi^-1(u)) corrected one count at a time
    against the exact CDF, P(X <= k) = Q(k + 1, lam) (scipy.special.pdtr)."""
    from scipy import special
    z = special.ndtri(u)
    k = np.nan_to_num(np.floor(lam + np.sqrt(lam) * z + (z * z - 1.) / 6.), nan=0., posinf=0., neginf=0.)
    np.maximum(k, 0., out=k)
    for _ in range(200):
        low = special.pdtr(k, lam) < u
        k[low] += 1.
        high = (k > 0.) & (special.pdtr(k - 1., lam) >= u)
        k[high] -= 1.
        if not low.any() and not high.any():
            return k.astype(np.int64)
    raise RuntimeError("Poisson inversion did not converge")


def poisson_line_integrals(p, i0, seed, return_counts=False, threads=16):
    """Noisy line integrals: counts ~ Poisson(I0 exp(-p)), p_hat = -ln(max(c, 1) / I0).

    The counts are drawn by inversion, one uniform variate per pixel from numpy's
    generator (`seed`), so a pixel's count depends only on its own line integral and
    its own variate: the same float32 p and seed give the same draw on any host, and
    two stacks that differ in a few pixels by a rounding step draw the same counts
    everywhere else (a sequential sampler would change every later draw).
    return_counts: also return the counts.  `threads` workers share the inversion."""
    from concurrent.futures import ThreadPoolExecutor
    lam = i0 * np.exp(-np.asarray(p, dtype=np.float64))
    u = np.random.default_rng(seed).random(lam.shape)
    flat_l, flat_u = lam.ravel(), u.ravel()
    c = np.empty(flat_l.shape, np.int64)
    chunk = 1 << 20

    def run(a):
        c[a:a + chunk] = _poisson_inverse(flat_l[a:a + chunk], flat_u[a:a + chunk])

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(run, range(0, flat_l.size, chunk)))
    c = c.reshape(lam.shape)
    out = (-np.log(np.maximum(c, 1) / i0)).astype(np.float32)
    return (out, c) if return_counts else out
