"""The NumPy restatement of the device generator (oracle/philox.py), pinned on the CPU."""
import numpy as np
import pytest


def test_philox_known_answers():
    """Known-answer vectors of Philox4x32-10 published with Random123 (kat_vectors)."""
    from oracle.philox import philox4x32_10

    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        got = philox4x32_10(np.array([ctr], dtype=np.uint64), key)[0]
        assert tuple(int(v) for v in got) == want, (ctr, key, [hex(int(v)) for v in got])


def test_u53_range_and_resolution():
    from oracle.philox import u53

    assert u53(0, 0) == 0.0
    assert u53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0**-53
    assert u53(0, 1 << 6) == 2.0**-53


def test_ball_distribution_and_independence():
    from oracle.philox import generate_ball

    R = 10.0
    a = generate_ball(200_000, seed=123, rng_stream=0, radius=R, centre=(1.0, -2.0, 3.0))
    r = np.linalg.norm(a[:, :3] - np.array([1.0, -2.0, 3.0]), axis=1)
    assert r.max() <= R and r.max() > 0.99 * R
    # uniform in the ball: r^3 / R^3 ~ U[0, 1)
    h = np.histogram((r / R) ** 3, bins=20, range=(0, 1))[0]
    assert np.ptp(h) / h.mean() < 0.1
    assert abs(a[:, 3].mean() / np.pi - 1) < 0.01 and a[:, 3].min() >= 0 and a[:, 3].max() < 2 * np.pi
    b = generate_ball(1000, seed=123, rng_stream=1, radius=R)
    c = generate_ball(1000, seed=124, rng_stream=0, radius=R)
    a0 = generate_ball(1000, seed=123, rng_stream=0, radius=R)
    assert not np.allclose(a0, b) and not np.allclose(a0, c)
    np.testing.assert_array_equal(a0, generate_ball(1000, seed=123, rng_stream=0, radius=R))
    # emitter i does not depend on n
    np.testing.assert_array_equal(a0[:10], generate_ball(10, seed=123, rng_stream=0, radius=R))


@pytest.mark.parametrize("bad", [dict(radius=-1.0), dict(radius=float("nan")), dict(n=-1)])
def test_generate_ball_argument_errors(bad):
    import ctypes

    from idi_amd import _lib

    lib = _lib.load()
    n = bad.get("n", 4)
    rc = lib.idisim_generate_ball(ctypes.c_uint64(1), ctypes.c_uint64(0), n, bad.get("radius", 1.0), None, 1, 1,
                                  None if n <= 0 else ctypes.c_void_p(1), None)
    assert rc == -1


def test_gauss_restatement_distribution():
    from oracle.philox import generate_gauss

    a = generate_gauss(200_000, seed=77, sigma=(1.0, 2.0, 0.5))
    np.testing.assert_allclose(a[:, :3].std(0), (1.0, 2.0, 0.5), rtol=0.01)
    assert np.abs(a[:, :3].mean(0)).max() < 0.01
    k = np.mean((a[:, 0] / a[:, 0].std()) ** 4)
    assert abs(k - 3) < 0.05  # Gaussian kurtosis
    assert abs(np.corrcoef(a[:, 0], a[:, 1])[0, 1]) < 0.01
    assert a[:, 3].min() >= 0 and a[:, 3].max() < 2 * np.pi
    # rotation acts on the positions only; |p| is invariant
    from idi_amd.util import rotation

    M = rotation(0.3, -0.7, 1.1)
    b = generate_gauss(1000, seed=77, sigma=1.0, rot=M)
    c = generate_gauss(1000, seed=77, sigma=1.0)
    np.testing.assert_allclose(b[:, :3], c[:, :3] @ M.T, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(np.linalg.norm(b[:, :3], axis=1), np.linalg.norm(c[:, :3], axis=1), rtol=1e-12)


def test_balls_restatement_membership():
    from oracle.philox import generate_balls

    centres = np.array([[-40.0, 0, 0], [0, 0, 0], [50.0, 0, 0]])
    counts = [100, 0, 250]
    a = generate_balls(5, centres, counts, radius=15.0)
    assert a.shape == (350, 4)
    d0 = np.linalg.norm(a[:100, :3] - centres[0], axis=1)
    d2 = np.linalg.norm(a[100:, :3] - centres[2], axis=1)
    assert d0.max() <= 15.0 and d2.max() <= 15.0


def test_choice_restatement_inclusion_probabilities():
    """Exponential keys give numpy's choice(replace=False, p) distribution: compare the inclusion
    frequency of every site with numpy's own sampler (uniform and weighted)."""
    from oracle.philox import generate_choice

    m, n, trials = 12, 4, 6000
    sites = np.stack([np.arange(m, dtype=float), np.zeros(m), np.zeros(m)], 1)
    w = np.linspace(0.2, 2.0, m)
    rng = np.random.default_rng(0)
    for weights in (None, w):
        ours = np.zeros(m)
        ref = np.zeros(m)
        p = None if weights is None else weights / weights.sum()
        for t in range(trials):
            got = generate_choice(1000 + t, n, sites, weights)
            idx = got[:, 0].astype(int)
            assert len(set(idx)) == n  # without replacement
            ours[idx] += 1
            ref[rng.choice(m, n, replace=False, p=p)] += 1
        ours /= trials
        ref /= trials
        if weights is None:
            np.testing.assert_allclose(ours, n / m, atol=0.025)
        np.testing.assert_allclose(ours, ref, atol=0.03)


def test_choice_restatement_rounds_and_rotation():
    from idi_amd.util import rotation
    from oracle.philox import generate_choice

    sites = np.random.default_rng(1).normal(size=(5, 3))
    a = generate_choice(7, 12, sites)  # 5 + 5 + 2: each round without replacement
    for k in range(0, 10, 5):
        assert sorted(map(tuple, a[k:k + 5, :3])) == sorted(map(tuple, sites))
    M = rotation(0.1, 0.2, 0.3)
    b = generate_choice(7, 12, sites, rot=M)
    np.testing.assert_allclose(b[:, :3], a[:, :3] @ M.T, rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(b[:, 3], a[:, 3])
