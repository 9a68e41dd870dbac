"""Device charge sharing (idisim_chargesharing) vs the NumPy restatement with the same Philox draws."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _counts(rng, shape, lam):
    c = rng.poisson(lam, shape).astype(np.float64)
    c[1, 2] = 30  # a bright pixel near the edge
    c[-1, -1] = 2.9  # fractional: two hits
    return c


@pytest.mark.parametrize("s,ns", [(0.3, 0.0), (0.75, 0.0), (1.6, 0.0), (0.75, 0.1)])
def test_chargesharing_matches_restatement(s, ns):
    import torch

    from idi_amd import _lib
    from oracle import photons

    rng = np.random.default_rng(17)
    counts = _counts(rng, (40, 33), 0.8)
    d = torch.from_numpy(counts).cuda()
    out = torch.empty_like(d)
    _lib.chargesharing(d, out, s, ns, seed=2**40 + 3)
    got = out.cpu().numpy()
    want = photons.chargesharing(counts, s, ns, seed=2**40 + 3)
    print("s", s, "ns", ns, "max abs diff", np.abs(got - want).max(), "sum", got.sum(), counts.sum())
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_pixel_gain_branch():
    import torch

    from idi_amd import _lib
    from oracle import photons

    counts = _counts(np.random.default_rng(2), (17, 64), 3.0)
    d = torch.from_numpy(counts).cuda()
    out = torch.empty_like(d)
    _lib.chargesharing(d, out, 0.0, 0.25, seed=99)
    np.testing.assert_allclose(out.cpu().numpy(), photons.pixel_gain(counts, 0.25, seed=99), rtol=1e-13)
    _lib.chargesharing(d, out, float("nan"), 0.0, seed=99)
    np.testing.assert_array_equal(out.cpu().numpy(), counts)


def test_argument_errors():
    import torch

    from idi_amd import _lib

    d = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.IdisimError):
        _lib.chargesharing(d, d, 0.5, 0.0, seed=1)  # out overlaps counts
    with pytest.raises(_lib.IdisimError):
        _lib.chargesharing(d, torch.empty_like(d), 40.0, 0.0, seed=1)  # sigma beyond 32 px
    with pytest.raises(_lib.IdisimError):
        _lib.chargesharing(d, torch.empty_like(d), 0.5, -1.0, seed=1)


def test_dropin_chargesharing_after_randomphotons():
    """simulate -> randomphotons -> chargesharing stays on the device; numpy in, numpy out works too."""
    import torch

    from idi_amd.simulation import common, hip, simobj

    obj = simobj.sphere(8048.0, 2000, 0.05)
    obj.rng = np.random.default_rng(4)
    field = hip.simulate(1, obj, 64, 200.0, 1e5, settings="single nf", returngpu=True, verbose=False)[0]
    I = (field.real ** 2 + field.imag ** 2).double()
    photons_img = common.randomphotons(I, 5000, rng=np.random.default_rng(1))
    assert photons_img.is_cuda and float(photons_img.sum()) > 4000
    shared = common.chargesharing(photons_img, 0.5, rng=7)
    assert shared.is_cuda and shared.dtype == torch.float64
    assert 0.8 * float(photons_img.sum()) < float(shared.sum()) <= float(photons_img.sum())
    again = common.chargesharing(photons_img.cpu().numpy(), 0.5, rng=7)
    np.testing.assert_array_equal(again, shared.cpu().numpy())
    assert common.chargesharing(photons_img, 0.0) is photons_img
    with pytest.raises(ValueError):
        common.chargesharing(-photons_img - 1, 0.5, rng=1)
    with pytest.raises(ValueError):
        common.chargesharing(photons_img * 1e9, 0.5, rng=1)  # intensities, not counts


def test_chargesharing_empty_image():
    import torch

    from idi_amd import _lib

    d = torch.zeros((13, 9), dtype=torch.float64, device="cuda")
    out = torch.full_like(d, 7.0)
    _lib.chargesharing(d, out, 1.2, 0.1, seed=3)
    assert float(out.abs().max()) == 0.0
