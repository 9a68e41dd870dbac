"""The q-corrected correlator restatement (oracle/correlate.py), pinned on the CPU."""
import numpy as np


def test_flat_detector_limit_is_plain_autocorrelation():
    """zd = 1 (detector at infinity): bin (dqy, dqx) = pixel offset (dc, dr), the plain
    autocorrelation, for every offset after the symmetrisation."""
    import scipy.signal as ss

    from idi_amd.reconstruction.qcor import geometry
    from oracle.correlate import qcorr

    rng = np.random.default_rng(0)
    img = rng.uniform(0, 1, (12, 10)).astype(np.float32)
    zd, xc, yc, qmax, maxdqz = geometry(img.shape, 1e9, 32)
    assert np.all(zd == 1) and maxdqz == 1 and qmax == 32
    out = qcorr(zd, img, xc, yc, 1e9, qmax, maxdqz)
    g = ss.correlate(img.astype(np.float64), img.astype(np.float64), "full", "direct")  # lag (dr, dc) at (dr+11, dc+9)
    got = out[maxdqz, qmax - 9:qmax + 10, qmax - 11:qmax + 12].T
    np.testing.assert_allclose(got, g, rtol=1e-12)
    assert out.sum() == np.sum(out[maxdqz])


def test_all_ones_reference_case():
    """The reference's own check (tests/test_recon.py:84-93): ones 32 x 32, z = 1000, maxq = 32."""
    from idi_amd.reconstruction.qcor import geometry
    from oracle.correlate import qcorr

    t = np.ones((32, 32))
    zd, xc, yc, qmax, maxdqz = geometry(t.shape, 1000, 32)
    c = qcorr(zd, t, xc, yc, 1000, qmax, maxdqz)
    assert c.shape[1:] == (64, 64) and len(c) > 1
    assert c.max() == t.sum()
    assert np.unravel_index(np.argmax(c), c.shape) == tuple(i // 2 for i in c.shape)
