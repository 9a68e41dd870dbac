"""Correlators on the GPU: FFT autocorrelation vs scipy.signal (the reference's own checks,
tests/test_recon.py:106-172), correlator2d statistics, q-corrected correlator vs the restatement."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_cusimple_reference_vectors():
    import scipy.signal as ss

    from idi_amd.reconstruction import cusimple

    t = np.arange(32 * 16).reshape(32, 16).astype(float)
    c = cusimple.corr(t)
    g = ss.correlate(t, t, "full", "fft")
    np.testing.assert_allclose(c[1:, 1:], g, atol=1e-7)

    t = np.arange(15 * 33).reshape(15, 33).astype(float)
    c = cusimple.corr(t, norm=True)
    gc = ss.correlate(t / t.mean(), t / t.mean(), "full", "fft")
    gn = ss.correlate(np.ones_like(t), np.ones_like(t), "full", "fft")
    g = gc / gn
    g[gn < 0.9] = np.nan
    np.testing.assert_allclose(c[1:, 1:], g, atol=1e-7)


def test_corr_stacks_and_axes():
    import scipy.signal as ss

    from idi_amd.reconstruction import cusimple

    t = np.arange(15 * 33).reshape(15, 33).astype(float)
    st = np.stack((t, 2 * t))
    c = cusimple.corr(st, axes=(-2, -1))
    for i in range(2):
        np.testing.assert_allclose(c[i, 1:, 1:], ss.correlate(st[i], st[i], "full", "fft"), atol=1e-6)
    sw = st.swapaxes(0, -1)  # correlate along axes 0 and 1
    c = cusimple.corr(sw, axes=(0, 1))
    np.testing.assert_allclose(c[1:, 1:, 0], ss.correlate(sw[..., 0], sw[..., 0], "full", "fft"), atol=1e-6)
    d = cusimple.corr(st, returngpu=True)
    assert d.is_cuda


def test_correlator2d_statistics():
    """mean / std over images of (image autocorrelation / mask autocorrelation), float32."""
    import scipy.signal as ss

    from idi_amd.reconstruction import cusimple

    rng = np.random.default_rng(5)
    mask = np.ones((20, 24), np.float32)
    mask[3:6, 10:14] = 0
    imgs = rng.uniform(0.5, 2.0, (6, 20, 24)).astype(np.float32)
    cor = cusimple.correlator2d(mask)
    for im in imgs:
        cor.add(im)
    mc = ss.correlate(mask.astype(np.float64), mask.astype(np.float64), "full", "fft")
    per = []
    for im in imgs:
        x = (im / im.mean()).astype(np.float64) * mask
        cc = ss.correlate(x, x, "full", "fft")
        per.append(np.where(mc > 0.5, cc / np.where(mc > 0.5, mc, 1), 0))
    per = np.array(per)
    assert cor.n == 6
    np.testing.assert_allclose(cor.mean[1:, 1:], per.mean(0), rtol=2e-4, atol=2e-5)
    np.testing.assert_allclose(cor.std[1:, 1:], per.std(0), rtol=5e-3, atol=5e-5)
    assert np.all(cor.mean[0] == 0) and np.all(cor.mean[:, 0] == 0)


def test_qcorr_matches_restatement():
    import torch

    from idi_amd import _lib
    from idi_amd.reconstruction.qcor import geometry
    from oracle.correlate import qcorr

    rng = np.random.default_rng(8)
    img = rng.uniform(0, 3, (20, 27)).astype(np.float32)
    for z, maxq in ((60.0, 32), (1e4, 40)):
        zd, xc, yc, qmax, maxdqz = geometry(img.shape, z, maxq)
        out = torch.empty((2 * maxdqz, 2 * qmax, 2 * qmax), dtype=torch.float64, device="cuda")
        _lib.qcorr(torch.from_numpy(zd).cuda(), torch.from_numpy(img).cuda(), xc, yc, z, qmax, maxdqz, out)
        got = out.cpu().numpy()
        want = qcorr(zd, img, xc, yc, z, qmax, maxdqz)
        print("z", z, "maxdqz", maxdqz, "max abs diff", np.abs(got - want).max(), "sum", want.sum())
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_qcorr_workgroup_histogram_geometry():
    """A 64 x 64 image at z = 500 px with maxq 32 (the workgroup-private LDS histogram with its dqx
    window, and the global atomics for pairs outside it) against the restatement; the same image
    off-centre (beam centre at a corner), where the windows are widest."""
    import torch

    from idi_amd import _lib
    from idi_amd.reconstruction.qcor import geometry
    from oracle.correlate import qcorr

    rng = np.random.default_rng(9)
    img = rng.poisson(2.0, (64, 64)).astype(np.float32)
    for xcen, ycen in ((None, None), (0.0, 0.0)):
        zd, xc, yc, qmax, maxdqz = geometry(img.shape, 500.0, 32, xcen, ycen)
        out = torch.empty((2 * maxdqz, 2 * qmax, 2 * qmax), dtype=torch.float64, device="cuda")
        _lib.qcorr(torch.from_numpy(zd).cuda(), torch.from_numpy(img).cuda(), xc, yc, 500.0, qmax, maxdqz, out)
        want = qcorr(zd, img, xc, yc, 500.0, qmax, maxdqz)
        np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-12, atol=1e-12)


def test_cucor_reference_case():
    """tests/test_recon.py:84-93 of the reference."""
    from idi_amd.reconstruction import cucor

    t = np.ones((32, 32))
    with cucor.corrfunction(t.shape, 1000, 32) as f:
        c = f(t)
        assert c.shape[1:] == (64, 64)
        assert len(c) > 1
        assert c.max() == t.sum()
        assert np.unravel_index(np.argmax(c), c.shape) == tuple(i // 2 for i in c.shape)
    with pytest.raises(ValueError):
        f(t)


def test_qcorr_wide_image_unstaged_rows():
    """Images wider than 1024 columns read their rows from global memory instead of LDS."""
    import torch

    from idi_amd import _lib
    from idi_amd.reconstruction.qcor import geometry
    from oracle.correlate import qcorr

    rng = np.random.default_rng(12)
    img = rng.uniform(0, 1, (3, 1100)).astype(np.float32)
    zd, xc, yc, qmax, maxdqz = geometry(img.shape, 500.0, 32)
    out = torch.empty((2 * maxdqz, 2 * qmax, 2 * qmax), dtype=torch.float64, device="cuda")
    _lib.qcorr(torch.from_numpy(zd).cuda(), torch.from_numpy(img).cuda(), xc, yc, 500.0, qmax, maxdqz, out)
    np.testing.assert_allclose(out.cpu().numpy(), qcorr(zd, img, xc, yc, 500.0, qmax, maxdqz), rtol=1e-12,
                               atol=1e-11)
