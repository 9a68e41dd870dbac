"""NumPy restatement of the q-corrected autocorrelation (TEST INFRASTRUCTURE ONLY).

cucor.corrfunction (idi/reconstruction/cucor.py:54-96): every reference pixel
(rr, rc) is paired with the target (rr + dr, rc + dc) for 0 <= dr < qmax and
|dc| < qmax; the pair adds img_r img_t to the bin
    dqx = qmax + round((tr - xc) zd_t - (rr - xc) zd_r)
    dqy = qmax + round((tc - yc) zd_t - (rc - yc) zd_r)
    dqz = maxdqz + round(z (zd_r - zd_t))
computed in float32 without contraction (round half to even); bins outside the
output are dropped.  Then every bin with dqx index < qmax and its mirror
(-i mod shape) both take their sum; a bin that is its own mirror is doubled.
Sums here are float64 in pair order.  Pinned by the flat-detector limit (zd = 1:
the plain autocorrelation, checked against scipy.signal.correlate) and by the
reference's own all-ones case (tests/test_recon.py:84-93).
"""
import numpy as np


def qcorr(zd, img, xc, yc, z, qmax, maxdqz):
    zd = np.asarray(zd, dtype=np.float32)
    img = np.asarray(img, dtype=np.float32)
    nr, nc = zd.shape
    sz, sq = 2 * maxdqz, 2 * qmax
    out = np.zeros((sz, sq, sq))
    xc, yc, cz = np.float32(xc), np.float32(yc), np.float32(z)
    cols = np.arange(nc)
    for rr in range(nr):
        for dr in range(qmax):
            tr = rr + dr
            if tr >= nr:
                break
            tr2 = np.float32(tr) - xc
            rr2 = np.float32(rr) - xc
            for dc in range(-(qmax - 1), qmax):
                c0, c1 = max(0, -dc), min(nc, nc - dc)
                if c0 >= c1:
                    continue
                rc = cols[c0:c1]
                tc = rc + dc
                zt, zr = zd[tr, tc], zd[rr, rc]
                dqx = qmax + np.rint(tr2 * zt - rr2 * zr).astype(np.int64)
                dqy = qmax + np.rint((tc.astype(np.float32) - yc) * zt - (rc.astype(np.float32) - yc) * zr).astype(
                    np.int64)
                dqz = maxdqz + np.rint(cz * (zr - zt)).astype(np.int64)
                ok = (dqx >= 0) & (dqx < sq) & (dqy >= 0) & (dqy < sq) & (dqz >= 0) & (dqz < sz)
                v = img[rr, rc].astype(np.float64) * img[tr, tc].astype(np.float64)
                np.add.at(out, (dqz[ok], dqy[ok], dqx[ok]), v[ok])
    return assemble(out, qmax)


def assemble(out, qmax):
    out = out.copy()
    sz, sq, _ = out.shape
    done = np.zeros(out.shape, bool)
    for i0 in range(sz):
        for i1 in range(sq):
            for i2 in range(qmax):
                m = ((sz - i0) % sz, (sq - i1) % sq, (sq - i2) % sq)
                if (i0, i1, i2) == m:
                    out[i0, i1, i2] *= 2
                    continue
                if done[i0, i1, i2]:
                    continue
                s = out[i0, i1, i2] + out[m]
                out[i0, i1, i2] = out[m] = s
                done[i0, i1, i2] = done[m] = True
    return out
