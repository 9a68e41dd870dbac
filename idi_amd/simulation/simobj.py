"""Emitter generators ("simulation objects"): drop-in for idi/simulation/simobj.py.

A simobject supplies the emitters of one image:
  * ``N`` -- number of emitters, ``k`` -- wavenumber 2 pi / (1.24 / E) in 1/um
    for an energy E in eV (simobj.py:42-48);
  * ``get2()`` -> (pos (N, 3) float64, phase (N, 1) float64 ~ U[0, 2 pi))
    (simobj.py:50-61); ``get()`` -> (N, 4) C-contiguous (x, y, z, phase)
    (simobj.py:63-66), the layout the C ABI takes;
  * ``rndPos`` / ``rndPhase`` switch fresh positions / phases per call;
    changing a shape parameter drops cached positions (simobj.py:104-109).

Difference from the reference, on purpose: every generator draws from
``self.rng`` (a numpy Generator; the class-level default is seeded from the
OS), so ``obj.rng = numpy.random.default_rng(seed)`` makes positions AND
phases reproducible.  The reference's sphere draws positions from an
unseeded generator (simobj.py:130, util/random.py:40).

Optional per-emitter amplitudes (no reference counterpart): set
``obj.amplitudes`` to an (N,) array, or override ``get_amplitudes()``.
"""
import abc as _abc
from warnings import warn as _warn

import numpy as _np
from numpy import pi

from ..util import angles as _angles
from ..util import axisrotation as _axisrotation
from ..util import gnorm as _gnorm
from ..util import poisson_disc_sample as _pds
from ..util import random_rotation as _random_rotation
from ..util import rndgennorm as _rndgennorm
from ..util import rndIcosahedron as _rndIcosahedron
from ..util import rndSphere as _rndSphere
from ..util import rotation as _rotation

__all__ = ["simobj", "sphere", "icosahedron", "gnorm", "gauss", "multisphere", "crystal", "sc", "fcc", "hcp",
           "cuso4", "grating"]


class simobj(_abc.ABC):
    """Base class of the emitter generators (simobj.py:26-116)."""

    rng = _np.random.default_rng()
    _pos = None
    _resetproperties = ()
    _rotmatrix = None
    amplitudes = None

    def __init__(self, E, N):
        self.E = E
        self.rndPhase = True
        self.rndPos = True
        self._N = int(N)

    @property
    def N(self):
        return self._N

    @property
    def k(self):
        """Angular wavenumber in 1/um for E in eV, hc approximated as 1.24 eV um (simobj.py:46-48)."""
        return 2 * pi / (1.24 / self.E)

    def get2(self):
        if self.rndPos or self._pos is None:
            self._pos = None
            self.updatePos()
        if self.rndPhase:
            phase = self.rng.uniform(0, 2 * pi, (self.N, 1))
        else:
            phase = _np.zeros((self.N, 1), self._pos.dtype)
        return self._pos, phase

    def get(self):
        pos, phase = self.get2()
        out = _np.empty((self.N, 4), dtype=_np.float64, order="C")
        out[:, :3] = pos
        out[:, 3:] = phase
        return out

    def get_amplitudes(self):
        """Per-emitter amplitudes (N,) or None for unit amplitudes."""
        if self.amplitudes is None:
            return None
        a = _np.asarray(self.amplitudes, dtype=_np.float64).reshape(-1)
        if a.shape != (self.N,):
            raise ValueError(f"amplitudes must have shape ({self.N},)")
        return a

    @_abc.abstractmethod
    def updatePos(self):
        pass

    @property
    def rotangles(self):
        if self._rotmatrix is None:
            _warn("Rotation not used for this object")
            return [0, 0, 0]
        if self._rotmatrix is False:
            return [0, 0, 0]
        return _angles(self._rotmatrix)

    @rotangles.setter
    def rotangles(self, value):
        if self._rotmatrix is None:
            _warn("Rotation not used for this object")
        elif not _np.any(value):
            self._rotmatrix = False
        elif self._rotmatrix is False:
            self._rotmatrix = _rotation(*value)
        else:
            new = _rotation(*value)
            self._rotate(new @ self._rotmatrix.T)
            self._rotmatrix = new

    def __setattr__(self, name, value):
        if name in self._resetproperties:
            self._pos = None
        super().__setattr__(name, value)

    def _rotate(self, rotmatrix):
        pass


class sphere(simobj):
    """Emitters uniform in a ball of radius r (simobj.py:119-130)."""

    def __init__(self, E, N, r):
        self.r = r
        super().__init__(E, N)
        self._resetproperties = ("r",)

    def updatePos(self):
        self._pos = _rndSphere(self.r, self.N, self.rng)


class icosahedron(simobj):
    """Emitters uniform in a regular icosahedron of circumradius r (simobj.py:133-151)."""

    def __init__(self, E, N, r, rotangles=(0, 0, 0)):
        self.r = r
        self.rndOrientation = False
        self._rotmatrix = _rotation(*rotangles)
        super().__init__(E, N)
        self._resetproperties = ("r",)

    def updatePos(self):
        if self.rndOrientation:
            self._rotmatrix = _random_rotation(self.rng)
        self._pos = _rndIcosahedron(self.r, self._N, self._rotmatrix, self.rng)

    def _rotate(self, rotmatrix):
        if self._pos is not None:
            self._pos = self._pos @ rotmatrix.T


class gnorm(simobj):
    """Generalised-normal shaped cloud with full widths `fwhm` and shape `rho` (simobj.py:154-176)."""

    def __init__(self, E, N, fwhm, rho=2, rotangles=(0, 0, 0)):
        self.fwhm, self.rho = _np.array(fwhm), _np.array(rho)
        self._rotmatrix = _rotation(*rotangles) if _np.any(rotangles) else False
        super().__init__(E, N)
        self._resetproperties = ("fwhm", "rho")

    def updatePos(self):
        if _np.all(self.rho == 2):
            self._pos = self.rng.normal(scale=self.fwhm / (2 * _np.sqrt(2 * _np.log(2))), size=(self.N, 3))
        else:
            self._pos = _rndgennorm(0, self.fwhm, self.rho, (self.N, 3), self.rng)
        if self._rotmatrix is not False:
            self._rotate(self._rotmatrix)

    def _rotate(self, rotmatrix):
        if self._pos is not None:
            self._pos = self._pos @ rotmatrix.T


class gauss(gnorm):
    """Gaussian cloud with full widths `fwhm` (simobj.py:179-184)."""

    def __init__(self, E, N, fwhm):
        super().__init__(E, N, fwhm, 2)


class multisphere(simobj):
    """Many balls of radius `rsphere` placed by Poisson-disc sampling; the number of
    emitters per ball follows a generalised normal profile of width `fwhm` about the
    origin (simobj.py:187-262).  Natoms < 0 means -Natoms per ball on average."""

    def __init__(self, E, Natoms=1e6, rsphere=10, fwhm=200, rho=2, spacing=1, Nspheres=_np.inf):
        self._Nspheres, self.rsphere, self.spacing = Nspheres, rsphere, spacing
        self.fwhm, self.rho = _np.array(fwhm), _np.array(rho)
        self.rndPos = True
        self._debug = None
        if Natoms < 0:
            N = int(-Natoms * _np.mean([len(self._spherepos()) for _ in range(10)]))
        else:
            N = Natoms
        super().__init__(E, N)
        self._resetproperties = ("rndPos", "rsphere", "fwhm", "spacing", "rho", "_Nspheres")

    def _spherepos(self):
        return _pds(1.2 * float(_np.max(self.fwhm)), 2 * (self.rsphere + self.spacing), ndim=3, N=self._Nspheres,
                    rng=self.rng)

    def centres_counts(self):
        """Ball centres (Poisson-disc sampled) and the number of emitters of each ball, summing to N."""
        centres = self._spherepos()
        if len(centres) == 0:
            centres = _np.zeros((1, 3))
        weight = _gnorm(centres, self.fwhm, self.rho, axis=1)
        weight = weight / weight.sum()
        counts = self.rng.poisson(weight * self._N)
        missing = self._N - counts.sum()
        while missing:
            ids = self.rng.choice(len(counts), int(abs(missing)), replace=True, p=weight)
            _np.add.at(counts, ids, int(_np.sign(missing)))
            counts[counts < 0] = 0
            missing = self._N - counts.sum()
        return centres, counts

    def updatePos(self):
        centres, counts = self.centres_counts()
        if self.rsphere > 0:
            pos = _rndSphere(self.rsphere, int(self._N), self.rng)
        else:
            pos = _np.zeros((self._N, 3))
        pos += _np.repeat(centres, counts, axis=0)
        self._debug = (len(centres), counts.min(), counts.max(), counts.mean())
        self._pos = pos


class crystal(simobj):
    """N emitters chosen without replacement from a lattice (simobj.py:357-496).

    The lattice (constants `lconst`, angles `langle`, basis `unitcell`,
    `repeats` cells per axis) is centred on the midpoint of its bounding box.
    With `fwhm`, sites are drawn with generalised-normal weights about the centre.
    """

    def __init__(self, E, lconst, langle, unitcell, N, repeats=None, rotangles=(0, 0, 0), fwhm=None, rho=2):
        if fwhm is not None:
            ca, cb, cc = _np.cos(_np.array(langle))
            sa, sb, sc_ = _np.sin(_np.array(langle))
            extent = _np.array([1, sc_, _np.sqrt(sb**2 - ((ca - cb * cc) / sc_) ** 2)]) * _np.array(lconst)
            needed = _np.ceil(2 * _np.array(fwhm) / extent).astype(int)
            if repeats is not None:
                if _np.any(needed > repeats):
                    _warn("Number of repeats small for choosen fwhm")
            else:
                repeats = needed
        if repeats is None:
            repeats = 3 * [int(_np.rint((N / len(unitcell)) ** (1 / 3.0)))]
        if _np.prod(repeats) * len(unitcell) < N:
            _warn("Number of atoms high for atoms in focus")
        allpos = crystal._lattice(lconst, langle, unitcell, repeats)
        if _np.any(rotangles):
            self._rotmatrix = _rotation(*rotangles)
            allpos = allpos @ self._rotmatrix.T
        else:
            self._rotmatrix = False
        self._allpos = allpos
        self.rndOrientation = False
        self._p = None
        self.fwhm = None if fwhm is None else _np.array(fwhm)
        self.rho = _np.array(2) if rho is None else _np.array(rho)
        super().__init__(E, N)
        self._resetproperties = ("rho", "fwhm")

    @staticmethod
    def _lattice(lconst, langle, unitcell, repeats, sigma=0, rng=None):
        """Lattice sites (float32 offsets, as the reference builds them), centred on the bounding-box midpoint."""
        ca, cb, cc = _np.cos(_np.array(langle))
        sa, sb, sc_ = _np.sin(_np.array(langle))
        basis = _np.array(
            [[1, 0, 0], [cc, sc_, 0], [cb, (ca - cb * cc) / sc_, _np.sqrt(sb**2 - ((ca - cb * cc) / sc_) ** 2)]]
        ) * _np.asarray(lconst, dtype=float)[:, None]
        cells = _np.stack(_np.meshgrid(*[_np.arange(int(n)) for n in repeats], indexing="ij"), -1).reshape(-1, 3)
        sites = (_np.asarray(unitcell, dtype=float).reshape(-1, 3) @ basis)[None, :, :] + (cells @ basis)[:, None, :]
        atoms = sites.reshape(-1, 3).astype(_np.float32)
        if sigma:
            rng = _np.random.default_rng() if rng is None else rng
            atoms = atoms + rng.uniform(0, sigma, atoms.shape).astype(_np.float32)
        return atoms - (atoms.min(0) + atoms.max(0)) / 2.0

    def orientation(self):
        """Rotation of the stored sites for this draw (a new random one with rndOrientation), or None."""
        if not _np.any(self.rndOrientation):
            return None
        if _np.size(self.rndOrientation) == 3:
            rot = _axisrotation(_np.array(self.rndOrientation, dtype=_np.float64), self.rng.uniform(0, 2 * pi))
        else:
            rot = _random_rotation(self.rng)
        self._rotmatrix = rot @ self._rotmatrix if self._rotmatrix is not False else rot
        self._p = None
        return self._rotmatrix

    def site_weights(self, pos):
        """Normalised choice weights of the (rotated) sites `pos`, or None without fwhm."""
        if self.fwhm is None:
            return None
        if self._p is None:
            p = _gnorm(pos, self.fwhm, self.rho, axis=1)
            self._p = p / p.sum()
        return self._p

    def updatePos(self):
        rot = self.orientation()
        pos = self._allpos if rot is None else self._allpos @ rot.T
        p = self.site_weights(pos)
        missing, idx = self.N, []
        while missing:
            new = self.rng.choice(len(self._allpos), min(missing, len(self._allpos)), replace=False, p=p)
            idx.append(new)
            missing -= len(new)
        sel = idx[0] if len(idx) == 1 else _np.concatenate(idx)
        self._pos = _np.take(pos, sel, axis=0).astype(_np.float64, copy=False)  # take: faster than fancy indexing

    def __setattr__(self, name, value):
        if name in self._resetproperties:
            self._p = None
        super().__setattr__(name, value)

    def _rotate(self, rotmatrix):
        self._allpos = self._allpos @ rotmatrix.T


class sc(crystal):
    """Simple cubic lattice, constant a."""

    def __init__(self, E, N, a, repeats=None, rotangles=(0, 0, 0), fwhm=None, rho=None):
        super().__init__(E, [a, a, a], _np.array([90, 90, 90]) * pi / 180.0, [[0, 0, 0]], N, repeats, rotangles,
                         fwhm, rho)


class fcc(crystal):
    """Face-centred cubic lattice, constant a."""

    def __init__(self, E, N, a, repeats=None, rotangles=(0, 0, 0), fwhm=None, rho=None):
        cell = [[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]
        super().__init__(E, [a, a, a], _np.array([90, 90, 90]) * pi / 180.0, cell, N, repeats, rotangles, fwhm,
                         rho)


class hcp(crystal):
    """Hexagonal close-packed lattice, constant a, c = 1.633 a."""

    def __init__(self, E, N, a, repeats=None, rotangles=(0, 0, 0), fwhm=None, rho=None):
        cell = [[0, 0, 0], [1.0 / 3, 2.0 / 3, 1.0 / 2]]
        super().__init__(E, [a, a, 1.633 * a], _np.array([90, 90, 120]) * pi / 180.0, cell, N, repeats, rotangles,
                         fwhm, rho)


class cuso4(crystal):
    """CuSO4 lattice with the fixed constants of simobj.py:533-545 (argument order N, E as there)."""

    def __init__(self, N, E, repeats=None, rotangles=(0, 0, 0), fwhm=None, rho=None):
        cell = [[0, 0, 0], [0.5, 0.5, 0]]
        lconst = _np.array([6.141, 10.736, 5.986]) * 1e-4
        langle = _np.array([82.27, 107.43, 102.67]) * pi / 180.0
        super().__init__(E, lconst, langle, cell, N, repeats, rotangles, fwhm, rho)


class grating(simobj):
    """Line grating along x (lines of `linewidth`, gaps of `spacingwidth`, edge shape `rholine`)
    under a generalised-normal focus `fwhm`/`rho` (simobj.py:548-598)."""

    def __init__(self, E, N, linewidth, spacingwidth, fwhm, rho=2, rholine=10, rotangles=(0, 0, 0)):
        self.linewidth = linewidth
        self.spacingwidth = spacingwidth
        self.fwhm = fwhm
        self.rho = rho
        self.rholine = rholine
        self._rotmatrix = _rotation(*rotangles) if _np.any(rotangles) else False
        self._x = None
        super().__init__(E, int(N))
        self._resetproperties = ("linewidth", "spacingwidth", "fwhm", "rho", "rholine")

    def updatePos(self):
        rhox = self.rho if _np.isscalar(self.rho) else self.rho[0]
        fwhmx = self.fwhm if _np.isscalar(self.fwhm) else self.fwhm[0]
        a, b = self.linewidth, self.spacingwidth
        x = _np.arange(-(0.5 + 2 / rhox) * fwhmx, (0.5 + 2 / rhox) * fwhmx, min(a, b) / 10)
        lines = _np.exp(-_np.abs((x % (a + b)) - (a + b) / 2) ** self.rholine * _np.log(2) * (2 / a) ** self.rholine)
        cdf = _np.cumsum(lines * _gnorm(x, fwhmx, rhox))
        cdf /= cdf[-1]
        rhoyz = self.rho if _np.isscalar(self.rho) else _np.array(self.rho[1:])
        fwhmyz = self.fwhm if _np.isscalar(self.fwhm) else _np.array(self.fwhm[1:])
        pos = _np.zeros((self.N, 3))
        pos[:, 0] = _np.interp(self.rng.uniform(size=self.N), cdf, x)
        pos[:, 1:] = _rndgennorm(0, fwhmyz, rhoyz, (self.N, 2), self.rng)
        if self._rotmatrix is not False:
            pos = pos @ self._rotmatrix.T
        self._pos = pos

    def _rotate(self, rotmatrix):
        if self._pos is not None:
            self._pos = self._pos @ rotmatrix.T
