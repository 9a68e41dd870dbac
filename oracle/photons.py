"""NumPy restatement of the detector charge sharing (TEST INFRASTRUCTURE ONLY).

chargesharing / __putcharges of idi/simulation/common.py:26-67, restated as
the reference computes it: a scatter over hits in hit order (input pixels in
row-major order, then hit number), each hit depositing the (2w+1)^2 kernel
k[x, y] = (1/4 E(mu_x, x - w)) E(mu_y, y - w) at ret[hit_x + y - w, hit_y + x - w]
(common.py:57-67), with E(mu, d) = erf(c + a) + erf(c - a), a = (mu - d)/(sqrt(2) s),
c = 1/(2 sqrt(2) s) (common.py:53,59-64), into an image padded by w and cropped
afterwards (common.py:42-45).

The reference draws mu with numpy's global generator; this restatement takes
the hit randoms as arguments (``hit_randoms`` makes the Philox4x32-10 numbers
the device kernel uses), so the GPU result can be checked value for value.
"""
import math

import numpy as np

from .philox import philox4x32_10, u53

TAG_HIT, TAG_HIT_GAIN, TAG_PIXEL_GAIN = 1, 2, 3


def _key(seed):
    return (int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF)


def hit_list(counts):
    """(pixel flat index, hit number) of every hit, in the reference's hit order (common.py:43)."""
    c = np.asarray(counts, dtype=np.float64)
    n = np.where(c >= 1.0, np.floor(np.minimum(c, 4294967295.0)), 0).astype(np.int64).ravel()
    pix = np.repeat(np.arange(n.size, dtype=np.int64), n)
    start = np.repeat(np.cumsum(n) - n, n)
    h = np.arange(pix.size, dtype=np.int64) - start
    return pix, h


def _block(pix, h, tag, seed):
    ctr = np.zeros((pix.size, 4), dtype=np.uint64)
    ctr[:, 0] = pix.astype(np.uint64) & np.uint64(0xFFFFFFFF)
    ctr[:, 1] = pix.astype(np.uint64) >> np.uint64(32)
    ctr[:, 2] = h.astype(np.uint64)
    ctr[:, 3] = tag
    return philox4x32_10(ctr, _key(seed))


def normal(r):
    """Box-Muller normal from one Philox block: sqrt(-2 ln u1) cos(2 pi u2), u1 in (0, 1]."""
    u1 = 1.0 - u53(r[:, 0], r[:, 1])
    u2 = u53(r[:, 2], r[:, 3])
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)


def hit_randoms(counts, seed, gain=False):
    """mu (nhits, 2) in [-1/2, 1/2) and, with gain, the N(0,1) draws (nhits,) of the device kernel."""
    pix, h = hit_list(counts)
    r = _block(pix, h, TAG_HIT, seed)
    mu = np.stack([u53(r[:, 0], r[:, 1]) - 0.5, u53(r[:, 2], r[:, 3]) - 0.5], 1)
    z = normal(_block(pix, h, TAG_HIT_GAIN, seed)) if gain else None
    return mu, z


def putcharges(counts, s, ns, mu, z=None):
    """Charge-shared image (float64) for the given hit randoms (common.py:26-67 with s > 0)."""
    counts = np.asarray(counts, dtype=np.float64)
    w = int(np.ceil(2 * s))
    pix, _ = hit_list(counts)
    n0, n1 = counts.shape
    ret = np.zeros((n0 + 2 * w, n1 + 2 * w))
    c = 1 / (2 * math.sqrt(2) * s)
    sq2s = math.sqrt(2) * s
    for i in range(pix.size):
        hx, hy = divmod(int(pix[i]), n1)
        hx += w
        hy += w
        mux, muy = mu[i]
        ex = np.array([math.erf(c + (mux - x) / sq2s) + math.erf(c - (mux - x) / sq2s) for x in range(-w, w + 1)])
        ey = np.array([math.erf(c + (muy - y) / sq2s) + math.erf(c - (muy - y) / sq2s) for y in range(-w, w + 1)])
        k = (0.25 * ex)[:, None] * ey[None, :]
        if ns > 0:
            k = (1 + z[i] * ns) * k
        for x in range(2 * w + 1):
            for y in range(2 * w + 1):
                ret[hx + y - w, hy + x - w] += k[x, y]
    return ret[w:n0 + w, w:n1 + w]


def pixel_gain(counts, ns, seed):
    """s <= 0 branch (common.py:33-37): counts (1 + ns N(0,1)) per pixel, or counts."""
    counts = np.asarray(counts, dtype=np.float64)
    if not ns > 0:
        return counts.copy()
    p = np.arange(counts.size, dtype=np.int64)
    z = normal(_block(p, np.zeros_like(p), TAG_PIXEL_GAIN, seed)).reshape(counts.shape)
    return counts * (1.0 + ns * z)


def chargesharing(counts, s, ns, seed):
    """The device kernel's result, restated: the reference algorithm with the device's random numbers."""
    if not s > 0:
        return pixel_gain(counts, ns, seed)
    mu, z = hit_randoms(counts, seed, gain=ns > 0)
    return putcharges(counts, s, ns, mu, z)
