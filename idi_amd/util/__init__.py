"""Geometry and sampling helpers used by the emitter generators and the detector set-up.

Drop-in subset of ``idi.util``: rotation, angles, axisrotation, abs2
(idi/util/__init__.py:139-144, 337-383), gnorm, fwhm, fastlen, and the random
samplers of idi/util/random.py (rndSphere, rndgennorm, random_rotation, ...)
plus poisson_disc_sample (idi/util/poissondisk.py:7).
"""
import numpy as _np

from .random import (  # noqa: F401
    poisson_disc_sample,
    random_rotation,
    rndConvexpoly,
    rndgennorm,
    rndIcosahedron,
    rndSphere,
    rndtruncexp,
)

__all__ = [
    "rotation", "angles", "axisrotation", "abs2", "gnorm", "fwhm", "fastlen",
    "rndSphere", "rndgennorm", "random_rotation", "rndConvexpoly", "rndIcosahedron", "rndtruncexp",
    "poisson_disc_sample",
]


def rotation(alpha, beta, gamma):
    """Rotation matrix Rz(gamma) Ry(beta) Rx(alpha) (yaw-pitch-roll); idi/util/__init__.py:348-372."""
    ca, cb, cg = _np.cos(_np.array((alpha, beta, gamma), dtype=_np.float64))
    sa, sb, sg = _np.sin(_np.array((alpha, beta, gamma), dtype=_np.float64))
    return _np.array(
        [
            [cb * cg, sa * sb * cg - ca * sg, ca * sb * cg + sa * sg],
            [cb * sg, sa * sb * sg + ca * cg, ca * sb * sg - sa * cg],
            [-sb, sa * cb, ca * cb],
        ]
    )


def angles(rotmatrix):
    """Inverse of :func:`rotation`; idi/util/__init__.py:375-383."""
    return (
        _np.arctan2(rotmatrix[2, 1], rotmatrix[2, 2]),
        -_np.arcsin(rotmatrix[2, 0]),
        _np.arctan2(rotmatrix[1, 0], rotmatrix[0, 0]),
    )


def axisrotation(axis, theta):
    """Rotation by `theta` about `axis` (Rodrigues); idi/util/__init__.py:337-346."""
    u = _np.asarray(axis, dtype=_np.float64)
    u = u / _np.linalg.norm(u)
    cross = _np.array([[0.0, -u[2], u[1]], [u[2], 0.0, -u[0]], [-u[1], u[0], 0.0]])
    return _np.cos(theta) * _np.eye(3) + _np.sin(theta) * cross + (1 - _np.cos(theta)) * _np.outer(u, u)


def abs2(x):
    """|x|^2 of a complex array (numpy or torch); idi/util/__init__.py:139-144."""
    try:
        import torch

        if isinstance(x, torch.Tensor):
            return x.real * x.real + x.imag * x.imag
    except ImportError:
        pass
    x = _np.asarray(x)
    return x.real * x.real + x.imag * x.imag


def gnorm(x, fwhm, rho, axis=None):
    """Generalised normal profile exp(-log2 (2|x|/fwhm)^rho), summed over `axis` in the exponent."""
    s = _np.log(2) * (2 / _np.asarray(fwhm, dtype=float)) ** rho
    e = _np.abs(x) ** rho * s
    if axis is not None:
        e = _np.sum(e, axis=axis)
    return _np.exp(-e)


def fwhm(x, y):
    """Full width at half maximum of a sampled peak (linear interpolation at the flanks)."""
    x = _np.asarray(x, dtype=float)
    y = _np.asarray(y, dtype=float)
    if len(x) == len(y) + 1:  # histogram edges
        x = (x[1:] + x[:-1]) / 2
    half = y.max() / 2
    above = _np.nonzero(y >= half)[0]
    i0, i1 = above[0], above[-1]

    def cross(a, b):
        if y[b] == y[a]:
            return x[a]
        return x[a] + (half - y[a]) * (x[b] - x[a]) / (y[b] - y[a])

    left = cross(i0 - 1, i0) if i0 > 0 else x[i0]
    right = cross(i1, i1 + 1) if i1 + 1 < len(y) else x[i1]
    return right - left


def fastlen(x, factors=(2, 3, 5, 7, 11)):
    """Smallest integer >= x whose prime factors are all in `factors` (fast FFT length)."""
    def one(n):
        n = int(n)
        best = None
        stack = [1]
        seen = set()
        while stack:
            v = stack.pop()
            if v >= n:
                best = v if best is None else min(best, v)
                continue
            for f in factors:
                w = v * f
                if w not in seen and (best is None or w < best * max(factors)):
                    seen.add(w)
                    stack.append(w)
        return best

    if _np.ndim(x) == 0:
        return one(x)
    return _np.array([one(v) for v in _np.asarray(x).ravel()]).reshape(_np.shape(x))
