"""FP64 oracle for idi's direct-sum wavefield (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, the smoke check of the repository's entry module and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only to CHECK the
HIP library.  The
product package ``idi_amd`` never imports it and has no CPU fallback.

Contents
  * :class:`Geometry` -- detector geometry and the ``settings`` flags of
    ``idi/simulation/cpu.py:112-212`` / ``cuda.py:16-74``.
  * :func:`rotation` -- restatement of ``idi/util/__init__.py:348-372``.
  * :func:`field_numpy` -- vectorised NumPy FP64 restatement of
    ``idi/simulation/cpu.py:30-107`` (small sizes).
  * :func:`field_c` / :func:`adjoint_c` -- ctypes wrappers of the OpenMP C
    restatement in ``wavefield.c`` (also the timed CPU baseline).
  * :func:`loss_central_diff` -- long-double central differences of the
    squared-intensity loss, the independent check of the gradients.

How each function is pinned (checked by ``tests/test_oracle.py``):
  * field_numpy, field_c: mpmath 50-digit golden vectors (tests/golden), the
    closed forms single emitter, two emitters, emitters at the origin, Laue
    lattice sum; translation, phase-shift and Friedel invariants; numpy vs C
    brute-force agreement.
  * adjoint_c: long-double central differences of the loss, and the closed-form
    gradient of a single emitter.
  * loss_central_diff: an independent pure-NumPy long-double evaluation of the
    loss at r +/- h on tiny cases, and the closed-form phase gradient of two
    emitters.
  * rotation: orthogonality, round trip through angles, known matrices.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field as _field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def rotation(alpha, beta, gamma):
    """Rz(gamma) @ Ry(beta) @ Rx(alpha); idi/util/__init__.py:348-372."""
    ca, cb, cg = np.cos([alpha, beta, gamma])
    sa, sb, sg = np.sin([alpha, beta, gamma])
    return np.array(
        [
            [cb * cg, sa * sb * cg - ca * sg, ca * sb * cg + sa * sg],
            [cb * sg, sa * sb * sg + ca * cg, ca * sb * sg - sa * cg],
            [-sb, sa * cb, ca * cb],
        ],
        dtype=np.float64,
    )


@dataclass
class Geometry:
    """Detector geometry + phase-model flags (cpu.py:30-40, 189-190; cuda.py:53-74)."""

    n0: int
    n1: int
    pixelsize: float
    detz: float
    k: float
    detoffset: tuple = (0.0, 0.0)
    detangles: tuple = (0.0, 0.0)
    nf: bool = False
    scale: bool = False
    secondorder: bool = False
    M: np.ndarray = _field(default=None, repr=False)

    def __post_init__(self):
        if self.M is None:
            if np.any(self.detangles):
                self.M = rotation(self.detangles[0], self.detangles[1], 0.0)
            else:
                self.M = np.eye(3)

    @property
    def rot(self):
        return bool(np.any(self.detangles))

    def pixels(self, rows=None):
        """Pixel positions (detx, dety, detz), each (nrows, n1); cpu.py:76-80."""
        rows = np.arange(self.n0) if rows is None else np.asarray(rows)
        x = rows[:, None].astype(np.float64)
        y = np.arange(self.n1, dtype=np.float64)[None, :]
        detx = (x - self.n0 / 2) * self.pixelsize + self.detoffset[0]
        dety = (y - self.n1 / 2) * self.pixelsize + self.detoffset[1]
        detx, dety = np.broadcast_arrays(detx, dety)
        detz = np.full(detx.shape, float(self.detz))
        if self.rot:
            M = self.M
            v = np.stack([detx, dety, detz])
            detx, dety, detz = (M[i, 0] * v[0] + M[i, 1] * v[1] + M[i, 2] * v[2] for i in range(3))
        return detx.copy(), dety.copy(), detz.copy()


def field_numpy(geom: Geometry, atoms, amp=None, rows=None):
    """FP64 field at the listed rows (all rows by default) -> complex128 (nrows, n1).

    Restates idi/simulation/cpu.py:42-107: per pixel, emitters in order, near-
    or far-field phase, optional 1/r scale, Kahan summation of re and im.
    """
    atoms = np.ascontiguousarray(atoms, dtype=np.float64)
    detx, dety, detz = geom.pixels(rows)
    k = float(geom.k)
    if not geom.nf:
        rdist = 1 / np.sqrt(detx**2 + dety**2 + detz**2)
        qz = k * (detz * rdist - 1)
        qx = k * (detx * rdist)
        qy = k * (dety * rdist)
        c2 = k * rdist / 2
    accx = np.zeros(detx.shape)
    accy = np.zeros(detx.shape)
    cx = np.zeros(detx.shape)
    cy = np.zeros(detx.shape)
    for j in range(len(atoms)):
        ax, ay, az, aphi = atoms[j]
        a = 1.0 if amp is None else float(amp[j])
        if geom.nf:
            dist = np.sqrt((detx - ax) ** 2 + (dety - ay) ** 2 + (detz - az) ** 2)
            phase = (dist - detz + az) * k + aphi
            mag = a / dist if geom.scale else a
        else:
            phase = -(ax * qx + ay * qy + az * qz) + aphi
            if geom.secondorder:
                phase = phase + c2 * (ax * ax + ay * ay)
            mag = a * rdist if geom.scale else a
        px = mag * np.cos(phase)
        py = mag * np.sin(phase)
        yx = px - cx
        yy = py - cy
        tx = accx + yx
        ty = accy + yy
        cx = (tx - accx) - yx
        cy = (ty - accy) - yy
        accx = tx
        accy = ty
    return accx + 1j * accy


class _CGeom(ctypes.Structure):
    _fields_ = [
        ("n0", ctypes.c_int64),
        ("n1", ctypes.c_int64),
        ("pixelsize", ctypes.c_double),
        ("detz", ctypes.c_double),
        ("k", ctypes.c_double),
        ("off0", ctypes.c_double),
        ("off1", ctypes.c_double),
        ("M", ctypes.c_double * 9),
        ("rot", ctypes.c_int32),
        ("nf", ctypes.c_int32),
        ("scale", ctypes.c_int32),
        ("secondorder", ctypes.c_int32),
    ]


def build(force=False):
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_field_rows.argtypes = [
            ctypes.POINTER(_CGeom), dp, dp, ctypes.c_int64,
            ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, dp, ctypes.c_int,
        ]
        lib.oracle_field_rows.restype = None
        lib.oracle_adjoint.argtypes = [
            ctypes.POINTER(_CGeom), dp, dp, ctypes.c_int64, dp, dp, dp, dp, dp, ctypes.c_int,
        ]
        lib.oracle_adjoint.restype = None
        lib.oracle_time_intensity.argtypes = [
            ctypes.POINTER(_CGeom), dp, dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double, dp, ctypes.c_int,
        ]
        lib.oracle_time_intensity.restype = None
        lib.oracle_loss_central_diff.argtypes = [
            ctypes.POINTER(_CGeom), dp, dp, ctypes.c_int64, dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, dp,
            dp, ctypes.c_int,
        ]
        lib.oracle_loss_central_diff.restype = None
        _lib = lib
    return _lib


def _cgeom(geom: Geometry):
    g = _CGeom()
    g.n0, g.n1 = int(geom.n0), int(geom.n1)
    g.pixelsize, g.detz, g.k = float(geom.pixelsize), float(geom.detz), float(geom.k)
    g.off0, g.off1 = float(geom.detoffset[0]), float(geom.detoffset[1])
    for i, v in enumerate(np.asarray(geom.M, dtype=np.float64).ravel()):
        g.M[i] = v
    g.rot = int(geom.rot)
    g.nf, g.scale, g.secondorder = int(geom.nf), int(geom.scale), int(geom.secondorder)
    return g


def _dptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def field_c(geom: Geometry, atoms, amp=None, rows=None, nthreads=0):
    """FP64 field (OpenMP C restatement) at the listed rows -> complex128 (nrows, n1)."""
    lib = _load()
    atoms = np.ascontiguousarray(atoms, dtype=np.float64)
    if atoms.ndim != 2 or atoms.shape[1] != 4:
        raise ValueError("atoms must be (N, 4)")
    amp = None if amp is None else np.ascontiguousarray(amp, dtype=np.float64)
    if rows is None:
        nrows, rows_arr, rptr = geom.n0, None, None
    else:
        rows_arr = np.ascontiguousarray(rows, dtype=np.int64)
        nrows = len(rows_arr)
        rptr = rows_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    out = np.empty((nrows, geom.n1, 2), dtype=np.float64)
    g = _cgeom(geom)
    lib.oracle_field_rows(ctypes.byref(g), _dptr(atoms), _dptr(amp), len(atoms), rptr, nrows,
                          _dptr(out), int(nthreads))
    return out[..., 0] + 1j * out[..., 1]


def adjoint_c(geom: Geometry, atoms, dLdI, field=None, amp=None, nthreads=0):
    """Analytic FP64 gradient of L(I): returns (grad_pos (N,3), grad_phase (N,), grad_amp (N,))."""
    lib = _load()
    atoms = np.ascontiguousarray(atoms, dtype=np.float64)
    amp = None if amp is None else np.ascontiguousarray(amp, dtype=np.float64)
    if field is None:
        field = field_c(geom, atoms, amp, nthreads=nthreads)
    f2 = np.ascontiguousarray(np.stack([field.real, field.imag], -1), dtype=np.float64)
    dLdI = np.ascontiguousarray(dLdI, dtype=np.float64)
    n = len(atoms)
    gpos = np.empty((n, 3))
    gph = np.empty(n)
    gamp = np.empty(n)
    g = _cgeom(geom)
    lib.oracle_adjoint(ctypes.byref(g), _dptr(atoms), _dptr(amp), n, _dptr(f2), _dptr(dLdI),
                       _dptr(gpos), _dptr(gph), _dptr(gamp), int(nthreads))
    return gpos, gph, gamp


def loss_central_diff(geom: Geometry, atoms, target, coords, h, amp=None, nthreads=0):
    """Central differences (L(+h) - L(-h)) / 2h of L = sum_p (|E_p|^2 - T_p)^2 in 80-bit long
    double (oracle/wavefield.c: the base field once, then only the moved emitter's term per
    coordinate).  coords: (j, c) pairs, c = 0, 1, 2 for x, y, z and 3 for the phase; h: one
    step or one per coordinate (length unit / radians).  Returns float64 (len(coords),)."""
    lib = _load()
    atoms = np.ascontiguousarray(atoms, dtype=np.float64)
    amp = None if amp is None else np.ascontiguousarray(amp, dtype=np.float64)
    target = np.ascontiguousarray(target, dtype=np.float64).reshape(geom.n0, geom.n1)
    cs = np.ascontiguousarray(coords, dtype=np.int64).reshape(-1, 2)
    if np.any(cs[:, 0] < 0) or np.any(cs[:, 0] >= len(atoms)) or np.any(cs[:, 1] < 0) or np.any(cs[:, 1] > 3):
        raise ValueError("coords must be (emitter, component) pairs with component 0..3")
    hs = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (len(cs),)))
    out = np.empty(len(cs))
    g = _cgeom(geom)
    lib.oracle_loss_central_diff(ctypes.byref(g), _dptr(atoms), _dptr(amp), len(atoms), _dptr(target),
                                 cs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(cs), _dptr(hs),
                                 _dptr(out), int(nthreads))
    return out


def time_c(geom: Geometry, atoms, times, c, tau, nthreads=0):
    """FP64 time-integrated intensity of decaying emitters (oracle/timesim.c) -> float64 (n0, n1)."""
    lib = _load()
    atoms = np.ascontiguousarray(atoms, dtype=np.float64)
    times = np.ascontiguousarray(times, dtype=np.float64)
    out = np.empty((geom.n0, geom.n1))
    g = _cgeom(geom)
    lib.oracle_time_intensity(ctypes.byref(g), _dptr(atoms), _dptr(times), len(atoms), float(c), float(tau),
                              _dptr(out), int(nthreads))
    return out


def time_pairsum(geom: Geometry, atoms, times, c, tau):
    """The same intensity from the closed form tau/2 sum_{j,l} a_j conj(a_l) exp(-|t_j - t_l|/tau)
    (the time integral of |sum_j a_j e^{-(t - t_j)/tau} H(t - t_j)|^2), with no sort; tiny sizes."""
    atoms = np.asarray(atoms, dtype=np.float64)
    times = np.asarray(times, dtype=np.float64)
    out = np.empty((geom.n0, geom.n1))
    for x in range(geom.n0):
        for y in range(geom.n1):
            px, py, pz = geom.pixels(rows=[x])
            p = np.array([px[0, y], py[0, y], pz[0, y]])
            R = np.linalg.norm(p)
            if geom.nf:
                d = np.linalg.norm(atoms[:, :3] - p, axis=1)
                s = d - R + atoms[:, 2]
                t = s / c + times
                a = np.exp(1j * (s * geom.k + atoms[:, 3])) / (d if geom.scale else 1.0)
            else:
                s = atoms[:, :3] @ (p / R - np.array([0.0, 0.0, 1.0]))
                t = -s / c + times
                a = np.exp(1j * (atoms[:, 3] - s * geom.k)) / (R if geom.scale else 1.0)
            w = np.exp(-np.abs(t[:, None] - t[None, :]) / tau)
            out[x, y] = tau / 2 * np.real(np.einsum("j,l,jl->", a, np.conj(a), w))
    return out


def relative_l2(a, b):
    """||a - b||_2 / ||b||_2 over all elements."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
