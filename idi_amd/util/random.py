"""Random samplers for emitter positions (drop-in for idi/util/random.py and poissondisk.py).

Every sampler takes an optional ``rng`` (numpy Generator) so that a seeded
generator gives reproducible emitters.
"""
import numpy as _np

_TWO_PI = 2 * _np.pi


def _rng(rng):
    return _np.random.default_rng() if rng is None else rng


def rndSphere(r, N, rng=None):
    """N points uniform in a ball of radius r (idi/util/random.py:39-51).

    Uniform cos(polar angle), uniform azimuth, radius r * U^(1/3).
    """
    rng = _rng(rng)
    N = int(N)
    t = rng.uniform(-1.0, 1.0, N)
    p = rng.uniform(0.0, _TWO_PI, N)
    u = r * rng.uniform(0.0, 1.0, N) ** (1.0 / 3.0)
    st = _np.sqrt(1.0 - t * t)
    out = _np.empty((N, 3))
    out[:, 0] = u * t
    out[:, 1] = u * st * _np.sin(p)
    out[:, 2] = u * st * _np.cos(p)
    return out


def rndgennorm(mu, fwhm, rho, N, rng=None):
    """Generalised normal samples with shape rho, mean mu and full width at half maximum fwhm.

    |x - mu| = fwhm/2 * (G / log 2)^(1/rho), G ~ Gamma(1/rho), random sign
    (idi/util/random.py:6-26).  rho = 2 is a normal distribution; large rho tends to
    U(-fwhm/2, fwhm/2).  mu, fwhm, rho broadcast against N.
    """
    rng = _rng(rng)
    if _np.isscalar(N):
        if not _np.isscalar(fwhm):
            N = (N, len(fwhm))
        elif not _np.isscalar(rho):
            N = (N, len(rho))
    rho = _np.asarray(rho, dtype=float)
    sign = rng.choice(_np.array([-1.0, 1.0]), N)
    g = rng.gamma(1.0 / rho, 1.0, N)
    return mu + _np.asarray(fwhm, dtype=float) / 2 * (g / _np.log(2)) ** (1.0 / rho) * sign


def random_rotation(rng=None, amount=1):
    """Uniformly random rotation matrix (Arvo, Graphics Gems III); idi/util/random.py:130-139."""
    rng = _rng(rng)
    theta = rng.uniform(0, _TWO_PI * amount)
    phi = rng.uniform(0, _TWO_PI)
    z = rng.uniform(0, amount)
    v = _np.array([_np.cos(phi) * _np.sqrt(z), _np.sin(phi) * _np.sqrt(z), _np.sqrt(1.0 - z)])
    c, s = _np.cos(theta), _np.sin(theta)
    Rz = _np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])
    return (2 * _np.outer(v, v) - _np.eye(3)) @ Rz


def rndConvexpoly(N, n, x, rotmatrix, r, rng=None):
    """N points uniform inside the convex polyhedron {p : n_i . p >= n_i . x_i} (unit scale), times r."""
    rng = _rng(rng)
    n = _np.asarray(n, dtype=float)
    x = _np.asarray(x, dtype=float)
    if rotmatrix is not None:
        n = n @ _np.asarray(rotmatrix).T
        x = x @ _np.asarray(rotmatrix).T
    crit = _np.einsum("ij,ij->i", n, x)
    found, out = 0, []
    while found < N:
        pts = rndSphere(1.0, int(_np.clip(2 * (N - found), 1024, max(N, 1024))), rng)
        inside = _np.all(pts @ n.T > crit, axis=1)
        pts = pts[inside] * r
        out.append(pts)
        found += len(pts)
    return _np.concatenate(out)[: int(N)]


def rndIcosahedron(r, N, rotmatrix=None, rng=None):
    """N points uniform inside a regular icosahedron with circumradius r."""
    phi = (1 + _np.sqrt(5)) / 2
    verts = []
    for a in (-1, 1):
        for b in (-phi, phi):
            verts += [(0, a, b), (a, b, 0), (b, 0, a)]
    verts = _np.array(verts, dtype=float) / _np.sqrt(1 + phi**2)
    edge = _np.min([_np.linalg.norm(verts[i] - verts[j]) for i in range(12) for j in range(i)])
    normals, points = [], []
    for i in range(12):
        for j in range(i):
            for k in range(j):
                if all(abs(_np.linalg.norm(verts[a] - verts[b]) - edge) < 1e-9 for a, b in ((i, j), (j, k), (i, k))):
                    c = (verts[i] + verts[j] + verts[k]) / 3
                    normals.append(-c / _np.linalg.norm(c))  # inward normal: n . p >= n . c inside
                    points.append(c)
    return rndConvexpoly(N, _np.array(normals), _np.array(points), rotmatrix, r, rng)


def rndtruncexp(lam, b, N, rng=None):
    """Exponential distribution with scale lam truncated to [0, b] (inverse CDF); idi/util/random.py:142-151."""
    rng = _rng(rng)
    u = rng.random(N)
    return -_np.log1p(u * _np.expm1(-b / lam)) * lam


def poisson_disc_sample(r, d, N=_np.inf, ndim=3, k=30, method="auto", rng=None):
    """Random points inside an ndim-ball of radius r with pairwise distance >= d (dart throwing).

    Candidates are drawn uniformly in the ball and accepted when no accepted point
    lies within d (checked on a grid of cell size d/sqrt(ndim)).  Stops at N points
    or after k * 1000 consecutive rejections.  Interface of idi/util/poissondisk.py:7
    (`method` is accepted for compatibility).
    """
    rng = _rng(rng)
    ndim = int(ndim)
    if ndim < 1:
        raise ValueError("ndim should be integer >=1")
    if r < 0 or d < 0:
        raise ValueError("r and d should be positive")
    if d == 0:
        if not _np.isfinite(N):
            raise ValueError("Zero distance would cause infinite number of spheres. Limit N or set distance >0")
        v = rng.normal(size=(int(N), ndim))
        v /= _np.linalg.norm(v, axis=1, keepdims=True)
        return v * r * rng.uniform(0, 1, (int(N), 1)) ** (1 / ndim)
    cell = d / _np.sqrt(ndim)
    grid = {}
    pts = []
    fails = 0
    reach = int(_np.ceil(_np.sqrt(ndim)))
    offsets = _np.array(_np.meshgrid(*[_np.arange(-reach, reach + 1)] * ndim, indexing="ij")).reshape(ndim, -1).T
    while len(pts) < N and fails < k * 1000:
        batch = rng.normal(size=(1024, ndim))
        batch /= _np.linalg.norm(batch, axis=1, keepdims=True)
        batch *= r * rng.uniform(0, 1, (1024, 1)) ** (1 / ndim)
        for p in batch:
            key = tuple(_np.floor(p / cell).astype(int))
            ok = True
            for off in offsets:
                q = grid.get(tuple(int(a) + int(b) for a, b in zip(key, off)))
                if q is not None and _np.sum((q - p) ** 2) < d * d:
                    ok = False
                    break
            if ok:
                grid[key] = p
                pts.append(p)
                fails = 0
                if len(pts) >= N:
                    break
            else:
                fails += 1
                if fails >= k * 1000:
                    break
    return _np.array(pts).reshape(-1, ndim)
