"""The NumPy restatement of detector charge sharing (oracle/photons.py), pinned on the CPU.

The reference draws its hit positions with numpy's unseeded global generator
(common.py:51,54), so its outputs cannot be reproduced; the restatement is
pinned instead by the closed forms of the Gaussian pixel integrals.
"""
import math

import numpy as np
import pytest


def _gauss_cdf(x):
    return 0.5 * (1 + np.vectorize(math.erf)(np.asarray(x) / math.sqrt(2)))


def test_hit_order_and_truncation():
    from oracle.photons import hit_list

    counts = np.array([[0, 2.7], [-1, 1], [np.nan, 3]])
    pix, h = hit_list(counts)
    assert pix.tolist() == [1, 1, 3, 5, 5, 5]
    assert h.tolist() == [0, 1, 0, 0, 1, 2]


def test_single_hit_matches_pixel_integrals():
    """Deposit of one hit = product of the Gaussian's integrals over the pixels (via the normal CDF),
    with mu_x along the second image axis and mu_y along the first (common.py:57-67)."""
    from oracle.photons import putcharges

    s, w = 0.7, 2
    counts = np.zeros((9, 11))
    counts[4, 5] = 1
    mux, muy = 0.31, -0.17
    out = putcharges(counts, s, 0, np.array([[mux, muy]]))
    d = np.arange(-w, w + 1)
    px = _gauss_cdf((d + 0.5 - mux) / s) - _gauss_cdf((d - 0.5 - mux) / s)  # second axis
    py = _gauss_cdf((d + 0.5 - muy) / s) - _gauss_cdf((d - 0.5 - muy) / s)  # first axis
    want = np.zeros_like(out)
    want[4 - w:4 + w + 1, 5 - w:5 + w + 1] = py[:, None] * px[None, :]
    np.testing.assert_allclose(out, want, rtol=1e-13, atol=1e-15)


def test_charge_conservation_closed_form():
    """Sum over the window telescopes: each interior hit deposits
    1/4 (erf((w+1/2-mu_x)/(sqrt2 s)) + erf((w+1/2+mu_x)/(sqrt2 s))) (same in mu_y)."""
    from oracle.photons import hit_randoms, putcharges

    rng = np.random.default_rng(3)
    s = 1.2
    w = math.ceil(2 * s)
    counts = np.zeros((24, 20))
    counts[w:-w, w:-w] = rng.poisson(0.6, (24 - 2 * w, 20 - 2 * w))
    mu, _ = hit_randoms(counts, seed=11)
    out = putcharges(counts, s, 0, mu)
    t = lambda m: math.erf((w + 0.5 - m) / (math.sqrt(2) * s)) + math.erf((w + 0.5 + m) / (math.sqrt(2) * s))
    want = sum(0.25 * t(a) * t(b) for a, b in mu)
    assert out.sum() == pytest.approx(want, rel=1e-12)
    assert want < counts.sum()  # the window misses the Gaussian's tails


def test_edge_hits_lose_charge():
    """A centred hit in a corner pixel keeps only the offsets 0..w of each axis (common.py:42-45)."""
    from oracle.photons import putcharges

    s, w = 0.8, 2
    counts = np.zeros((6, 6))
    counts[0, 0] = 1
    edge = putcharges(counts, s, 0, np.array([[0.0, 0.0]])).sum()
    full = putcharges(np.pad(counts, 3), s, 0, np.array([[0.0, 0.0]])).sum()
    kept = (_gauss_cdf((w + 0.5) / s) - _gauss_cdf(-0.5 / s)) ** 2
    window = (_gauss_cdf((w + 0.5) / s) - _gauss_cdf(-(w + 0.5) / s)) ** 2
    assert edge == pytest.approx(kept, rel=1e-12)
    assert full == pytest.approx(window, rel=1e-12)


def test_normal_draws_and_gain():
    from oracle.photons import hit_randoms, pixel_gain

    counts = np.full((300, 300), 1.0)
    _, z = hit_randoms(counts, seed=5, gain=True)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert abs(np.mean(z ** 4) - 3) < 0.05  # Gaussian kurtosis
    g = pixel_gain(counts * 4, 0.1, seed=5)
    zz = (g / 4 - 1) / 0.1
    assert abs(zz.mean()) < 0.01 and abs(zz.std() - 1) < 0.01
    np.testing.assert_array_equal(pixel_gain(counts, 0.0, seed=5), counts)


def test_uniform_positions():
    from oracle.photons import hit_randoms

    mu, _ = hit_randoms(np.full((200, 200), 2.0), seed=9)
    assert mu.min() >= -0.5 and mu.max() < 0.5
    h = np.histogram(mu[:, 0], bins=10, range=(-0.5, 0.5))[0]
    assert np.ptp(h) / h.mean() < 0.05
    assert abs(np.corrcoef(mu[:, 0], mu[:, 1])[0, 1]) < 0.01
