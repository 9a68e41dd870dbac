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
