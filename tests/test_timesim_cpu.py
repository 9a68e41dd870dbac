"""The FP64 restatement of the time-dependent simulation (oracle time_c), pinned on the CPU."""
import numpy as np
import pytest

from conftest import K_UM

C_UM_FS = 0.299792458  # speed of light in um / fs


def _atoms(rng, n, r):
    u = rng.normal(size=(n, 3))
    u *= (r * rng.uniform(0, 1, n) ** (1 / 3) / np.linalg.norm(u, axis=1))[:, None]
    return np.concatenate([u, rng.uniform(0, 2 * np.pi, (n, 1))], 1)


@pytest.mark.parametrize("nf,scale", [(False, False), (True, False), (False, True), (True, True)])
def test_sorted_scan_equals_pair_sum(oracle_mod, nf, scale):
    """The reference's sorted decaying scan equals the sort-free closed form
    tau/2 sum_jl a_j conj(a_l) exp(-|t_j - t_l|/tau) (same integral of |E(t)|^2)."""
    import oracle

    rng = np.random.default_rng(1)
    at = _atoms(rng, 60, 0.02)
    times = rng.normal(0, 5e-4, 60)  # fs: comparable to the ~1e-3 fs spread of path differences
    g = oracle.Geometry(5, 4, 300.0, 1e5, K_UM, nf=nf, scale=scale)
    a = oracle_mod.time_c(g, at, times, C_UM_FS, 1e-3)
    b = oracle_mod.time_pairsum(g, at, times, C_UM_FS, 1e-3)
    np.testing.assert_allclose(a, b, rtol=1e-10)


def test_one_and_two_emitters(oracle_mod):
    import oracle

    g = oracle.Geometry(3, 3, 300.0, 1e5, K_UM)
    at = np.array([[0.0, 0.0, 0.0, 0.3]])
    np.testing.assert_allclose(oracle_mod.time_c(g, at, [0.0], C_UM_FS, 5.0), 2.5, rtol=1e-14)
    at2 = np.array([[0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, np.pi / 3]])
    I = oracle_mod.time_c(g, at2, [0.0, 1.0], C_UM_FS, 2.0)
    want = 1.0 * (1 + 1 + 2 * np.cos(np.pi / 3) * np.exp(-1.0 / 2.0))
    np.testing.assert_allclose(I, want, rtol=1e-13)


def test_long_decay_limit_is_the_static_intensity(oracle_mod):
    """tau >> every arrival spread: 2 I / tau -> |sum a_j e^{i theta}|^2, the time-independent
    intensity (the near-field phase differs from the static one by a per-pixel constant)."""
    import oracle

    rng = np.random.default_rng(2)
    at = _atoms(rng, 200, 0.02)
    for nf in (False, True):
        g = oracle.Geometry(6, 5, 300.0, 1e5, K_UM, nf=nf)
        I = oracle_mod.time_c(g, at, np.zeros(200), C_UM_FS, 1e6)
        static = np.abs(oracle_mod.field_c(g, at)) ** 2
        np.testing.assert_allclose(2 * I / 1e6, static, rtol=1e-6)


def test_short_decay_limit_is_incoherent(oracle_mod):
    """tau << every gap between arrivals: I = tau/2 sum |a_j|^2."""
    import oracle

    g = oracle.Geometry(2, 2, 300.0, 1e5, K_UM, scale=True)
    at = _atoms(np.random.default_rng(3), 50, 0.02)
    times = np.arange(50) * 1.0
    I = oracle_mod.time_c(g, at, times, C_UM_FS, 1e-3)
    R = np.linalg.norm(np.stack(g.pixels(), -1), axis=-1)
    np.testing.assert_allclose(I, 1e-3 / 2 * 50 / R**2, rtol=1e-9)
