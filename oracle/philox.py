"""NumPy restatement of the device emitter generator (TEST INFRASTRUCTURE ONLY).

Philox4x32-10 from Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
easy as 1, 2, 3" (SC'11): 10 rounds of
    (c0, c1, c2, c3) -> (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0))
with M0 = 0xD2511F53, M1 = 0xCD9E8D57 and the key bumped by (0x9E3779B9,
0xBB67AE85) between rounds.  Pinned by the published known-answer vectors
(tests/test_generate_cpu.py).  Used to check idisim_generate_ball.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32 counters, key: (2,) ints -> (..., 4) uint32."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for r in range(10):
        if r > 0:
            k0 = (k0 + W0) & 0xFFFFFFFF
            k1 = (k1 + W1) & 0xFFFFFFFF
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
    return np.stack([x.astype(np.uint32) for x in c], -1)


def u53(a, b):
    """53-bit double in [0, 1) from two 32-bit words (27 + 26 bits)."""
    a = np.asarray(a, dtype=np.uint64) >> np.uint64(5)
    b = np.asarray(b, dtype=np.uint64) >> np.uint64(6)
    return ((a << np.uint64(26)) | b).astype(np.float64) * (1.0 / 9007199254740992.0)


def generate_ball(n, seed, rng_stream=0, radius=1.0, centre=(0.0, 0.0, 0.0)):
    """(n, 4) emitters (x, y, z, phase) exactly as idisim_generate_ball draws them."""
    i = np.arange(n, dtype=np.uint64)
    base = np.zeros((n, 4), dtype=np.uint64)
    base[:, 0] = i & MASK
    base[:, 1] = i >> np.uint64(32)
    base[:, 2] = np.uint64(int(rng_stream) & 0xFFFFFFFF)
    key = (int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF)
    r0 = philox4x32_10(base, key)
    c1 = base.copy()
    c1[:, 3] = 1
    r1 = philox4x32_10(c1, key)
    t = 2.0 * u53(r0[:, 0], r0[:, 1]) - 1.0
    az = 2 * np.pi * u53(r0[:, 2], r0[:, 3])
    rr = radius * np.cbrt(u53(r1[:, 0], r1[:, 1]))
    st = np.sqrt(1.0 - t * t)
    out = np.empty((n, 4))
    out[:, 0] = centre[0] + rr * t
    out[:, 1] = centre[1] + rr * st * np.sin(az)
    out[:, 2] = centre[2] + rr * st * np.cos(az)
    out[:, 3] = 2 * np.pi * u53(r1[:, 2], r1[:, 3])
    return out


def _blocks(n, seed, rng_stream, word3):
    i = np.arange(n, dtype=np.uint64)
    ctr = np.zeros((n, 4), dtype=np.uint64)
    ctr[:, 0] = i & MASK
    ctr[:, 1] = i >> np.uint64(32)
    ctr[:, 2] = np.uint64(int(rng_stream) & 0xFFFFFFFF)
    ctr[:, 3] = word3
    return philox4x32_10(ctr, (int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF))


def generate_balls(seed, centres, counts, rng_stream=0, radius=1.0):
    """(N, 4) emitters as idisim_generate_balls draws them: ball draws about the origin
    shifted by the centre of each emitter's ball (counts[b] consecutive emitters per ball)."""
    counts = np.asarray(counts, dtype=np.int64)
    out = generate_ball(int(counts.sum()), seed, rng_stream, radius)
    out[:, :3] += np.repeat(np.asarray(centres, dtype=np.float64).reshape(-1, 3), counts, axis=0)
    return out


def _normal_pair(r):
    rad = np.sqrt(-2.0 * np.log(1.0 - u53(r[:, 0], r[:, 1])))
    ang = 2 * np.pi * u53(r[:, 2], r[:, 3])
    return rad * np.cos(ang), rad * np.sin(ang)


def generate_gauss(n, seed, sigma, rot=None, rng_stream=0):
    """(n, 4) emitters as idisim_generate_gauss draws them (Box-Muller on blocks 0 and 1, phase from block 2)."""
    z0, z1 = _normal_pair(_blocks(n, seed, rng_stream, 0))
    z2, _ = _normal_pair(_blocks(n, seed, rng_stream, 1))
    p = np.stack([z0, z1, z2], 1) * np.broadcast_to(np.asarray(sigma, dtype=np.float64), (3,))
    if rot is not None:
        p = p @ np.asarray(rot, dtype=np.float64).T
    out = np.empty((n, 4))
    out[:, :3] = p
    r2 = _blocks(n, seed, rng_stream, 2)
    out[:, 3] = 2 * np.pi * u53(r2[:, 0], r2[:, 1])
    return out


def generate_choice(seed, n, sites, weights=None, rot=None, rng_stream=0):
    """(n, 4) emitters as idisim_generate_choice draws them: the n smallest keys -ln(u_i)/w_i
    (rounds of fresh keys beyond len(sites)), sites rotated by rot, phases from block 0x200."""
    sites = np.asarray(sites, dtype=np.float64).reshape(-1, 3)
    m = len(sites)
    w = np.ones(m) if weights is None else np.asarray(weights, dtype=np.float64)
    out = np.empty((n, 4))
    done, rnd = 0, 0
    while done < n:
        r = _blocks(m, seed, rng_stream, 0x100 + rnd)
        u = 1.0 - u53(r[:, 0], r[:, 1])
        with np.errstate(divide="ignore"):
            keys = np.where(w > 0, -np.log(u) / np.where(w > 0, w, 1.0), np.inf)
        take = min(n - done, m)
        idx = np.argsort(keys, kind="stable")[:take]
        p = sites[idx]
        out[done:done + take, :3] = p if rot is None else p @ np.asarray(rot, dtype=np.float64).T
        done += take
        rnd += 1
    r = _blocks(n, seed, rng_stream, 0x200)
    out[:, 3] = 2 * np.pi * u53(r[:, 0], r[:, 1])
    return out
