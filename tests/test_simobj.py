"""Emitter generators and helpers (CPU).  Behaviour tests of idi's tests/test_sim.py:15-96
and tests/test_util.py:11-18, 100-146, run against the drop-in."""
import warnings

import numpy as np
import numpy.testing as testing
import pytest


def test_import():
    with warnings.catch_warnings():
        warnings.filterwarnings("ignore", category=UserWarning)
        import idi_amd.simulation  # noqa: F401


def test_sphere():
    from idi_amd.simulation import simobj

    o = simobj.sphere(E=1, N=1e3, r=10)
    o.rndPos = False
    o.rndPhase = False
    assert o.k == pytest.approx(2 * np.pi / 1.24)
    assert o.N == 1000
    t = o.get()
    assert t.shape == (1000, 4) and t.flags["C_CONTIGUOUS"] and t.dtype == np.float64
    testing.assert_array_less(np.linalg.norm(t[:, :3], axis=1), 10)
    testing.assert_allclose(t[:, 3:], 0)
    o.r = 5
    t = o.get()
    r = np.linalg.norm(t[:, :3], axis=1)
    assert r.max() == pytest.approx(5, abs=0.5)
    testing.assert_array_less(r, 5)
    o.rndPhase = True
    pos2, phase2 = o.get2()
    assert not np.allclose(phase2, 0)
    assert phase2.mean() == pytest.approx(np.pi, abs=0.2)
    testing.assert_allclose(t[:, :3], pos2)
    o.rndPos = True
    pos3, phase3 = o.get2()
    assert not np.allclose(pos2, pos3)
    assert not np.allclose(phase2, phase3)
    with pytest.warns(Warning):
        o.rotangles = [1, 1, 1]


def test_sphere_seeded_reproducible():
    from idi_amd.simulation import simobj

    a = simobj.sphere(8000, 500, 0.05)
    b = simobj.sphere(8000, 500, 0.05)
    a.rng = np.random.default_rng(42)
    b.rng = np.random.default_rng(42)
    testing.assert_array_equal(a.get(), b.get())


def test_sc():
    from idi_amd.simulation import simobj

    o = simobj.sc(1, 100, 1)
    pos = o.get()
    assert len(pos) == 100
    assert all(np.allclose(np.unique(c), np.arange(-2, 3)) for c in pos[:, :3].T)
    o = simobj.sc(1, 1000, 1, repeats=[10, 10, 10])
    p0 = o.get2()[0]
    r0 = sorted(np.linalg.norm(p0, axis=1))
    o.rotangles = [0.5, 0.1, 0.1]
    p1 = o.get2()[0]
    r1 = sorted(np.linalg.norm(p1, axis=1))
    assert not np.allclose(p0, p1)
    testing.assert_allclose(r0, r1, rtol=1e-6)
    with pytest.warns(UserWarning):
        o = simobj.sc(1, 100000, 2, fwhm=[10, 10, 10], rho=[2, 2, 10])
    assert np.max(o.get2()[0]) >= 5
    assert np.max(o.get2()[0]) <= 20


def test_gauss_multisphere_grating_counts():
    from idi_amd.simulation import simobj

    assert len(simobj.gauss(1, 100, 1).get()) == 100
    assert len(simobj.multisphere(1, 100, 1, Nspheres=10).get()) == 100
    assert len(simobj.multisphere(1, 100, 1, fwhm=10).get()) == 100
    o = simobj.grating(1, 100, 1, 10, fwhm=[10, 10, 1], rho=[2, 2, 100], rotangles=[np.pi, 0, 0])
    assert len(o.get()) == 100
    assert len(simobj.icosahedron(1, 200, 1.0).get()) == 200
    for cls in (simobj.fcc, simobj.hcp):
        assert len(cls(1, 500, 1.0).get()) == 500
    assert len(simobj.cuso4(300, 1).get()) == 300


def test_multisphere_min_distance():
    from idi_amd.util import poisson_disc_sample

    pts = poisson_disc_sample(12, 4, ndim=3, rng=np.random.default_rng(1))
    d = np.linalg.norm(pts[:, None] - pts[None], axis=-1) + np.eye(len(pts)) * 1e9
    assert d.min() >= 4 and np.linalg.norm(pts, axis=1).max() <= 12 and len(pts) > 20


def test_amplitudes_attribute():
    from idi_amd.simulation import simobj

    o = simobj.sphere(1, 10, 1)
    assert o.get_amplitudes() is None
    o.amplitudes = np.ones(10) * 0.5
    assert o.get_amplitudes().shape == (10,)
    o.amplitudes = np.ones(9)
    with pytest.raises(ValueError):
        o.get_amplitudes()


# ---------------------------------------------------------------- util
def test_angles_rotation_roundtrip():
    from idi_amd.util import angles, rotation

    testing.assert_allclose(angles(np.eye(3)), (0, 0, 0), atol=1e-15)
    testing.assert_allclose(rotation(0, 0, 0), np.eye(3), atol=1e-15)
    for case in [(np.pi, 0, 0), (0, np.pi, 0), (0, 0, np.pi), (1, 1, 1)]:
        rm = rotation(*case)
        testing.assert_allclose(rotation(*angles(rm)), rm, atol=1e-15)
        testing.assert_allclose(rm @ rm.T, np.eye(3), atol=1e-15)
    # known matrix: 90 deg about z maps x to y
    testing.assert_allclose(rotation(0, 0, np.pi / 2) @ [1, 0, 0], [0, 1, 0], atol=1e-15)


def test_axisrotation():
    from idi_amd.util import axisrotation

    R = axisrotation(np.array([0, 0, 2.0]), np.pi / 2)
    testing.assert_allclose(R @ [1, 0, 0], [0, 1, 0], atol=1e-15)
    testing.assert_allclose(np.linalg.det(R), 1)


def test_rndSphere():
    from idi_amd.util import rndSphere

    R, N = 10, int(1e6)
    t = rndSphere(R, N, np.random.default_rng(0))
    r = np.linalg.norm(t, axis=1)
    assert t.shape == (N, 3)
    assert r.max() > R * 0.95 and r.max() <= R
    hist = np.histogram(r, bins=np.linspace(0, 1, int(N ** (1 / 3) * 0.2)) ** (1 / 3) * R)[0]
    assert np.ptp(hist) / np.mean(hist) < 0.05
    d, b = np.histogramdd(t, density=True, bins=11)
    br = np.linalg.norm(np.meshgrid(*((c[:-1] + c[1:]) / 2 for c in b)), axis=0)
    f = d[br < 0.9 * R]
    assert f.std() / f.mean() < 0.05


def test_rndgennorm():
    from idi_amd.util import fwhm, rndgennorm

    fwhms = (1, 5, 10)
    t = rndgennorm(0, fwhms, (2, 2, 100), int(2e6), np.random.default_rng(1))
    testing.assert_allclose(np.mean(t, axis=0), 0, atol=0.02)
    assert t.shape == (int(2e6), 3)
    for x, f in zip(t.T, fwhms):
        h, b = np.histogram(x, bins=int(1e3), range=(-10, 10))
        assert fwhm(b, h) == pytest.approx(f, abs=0.15)


def test_random_rotation():
    from idi_amd.util import random_rotation

    t = random_rotation(np.random.default_rng(3))
    assert t.shape == (3, 3)
    assert abs(np.linalg.det(t)) == pytest.approx(1)
    testing.assert_allclose(t @ t.T, np.eye(3), atol=1e-12)


def test_abs2_and_randomphotons():
    from idi_amd.simulation import randomphotons
    from idi_amd.util import abs2

    z = np.array([3 + 4j, 1j])
    testing.assert_allclose(abs2(z), [25, 1])
    p = np.ones((4, 8, 8))
    ph = randomphotons(p, 1e4, rng=np.random.default_rng(0))
    assert ph.shape == p.shape
    assert abs(ph.sum(axis=(1, 2)).mean() / 1e4 - 1) < 0.05
    assert randomphotons(p, np.inf) is p


def test_host_extent_bounds_the_emitters():
    """The planner's extent for host-drawn images: a crystal's largest site norm (float32 sites
    measured in FP64, unchanged by random orientations) and a sphere's radius bound every draw."""
    from idi_amd.simulation import hip, simobj

    cache = {}
    for seed in range(4):
        obj = simobj.sc(8000, 1000, 3e-4, repeats=[12, 12, 12])
        obj.rng = np.random.default_rng(seed)
        obj.rndOrientation = bool(seed % 2)
        pos, _ = obj.get2()
        assert hip._host_extent(obj, pos, cache) >= np.linalg.norm(pos, axis=1).max()
    ball = simobj.sphere(8000, 2000, 0.05)
    ball.rng = np.random.default_rng(1)
    pos, _ = ball.get2()
    assert hip._host_extent(ball, pos, cache) >= np.linalg.norm(pos, axis=1).max()
    gauss = simobj.gauss(8000, 500, 0.02)
    gauss.rng = np.random.default_rng(2)
    pos, _ = gauss.get2()
    assert hip._host_extent(gauss, pos, cache) == pytest.approx(np.linalg.norm(pos, axis=1).max(), rel=1e-15)
