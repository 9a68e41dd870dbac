"""idisim_generate_ball on the GPU vs the NumPy restatement (oracle/philox.py)."""
import numpy as np
import pytest

from conftest import K_UM
from gpu_util import rel_l2

pytestmark = pytest.mark.gpu


def test_device_generator_matches_restatement():
    import torch

    from idi_amd import _lib
    from oracle.philox import generate_ball

    n = 100_003
    out = torch.empty((n, 4), dtype=torch.float64, device="cuda")
    _lib.generate_ball(out, seed=2**40 + 17, rng_stream=3, radius=0.03, centre=(0.01, 0.0, -0.02))
    got = out.cpu().numpy()
    ref = generate_ball(n, seed=2**40 + 17, rng_stream=3, radius=0.03, centre=(0.01, 0.0, -0.02))
    np.testing.assert_array_equal(got[:, 3], ref[:, 3])  # phases: same bits
    assert np.abs(got[:, :3] - ref[:, :3]).max() <= 1e-16 * 4
    # phases only: positions untouched
    out[:, :3] = 7.0
    _lib.generate_ball(out, seed=5, positions=False)
    o = out.cpu().numpy()
    assert np.all(o[:, :3] == 7.0)
    np.testing.assert_array_equal(o[:, 3], generate_ball(n, seed=5)[:, 3])


def test_simulate_gen_device_rng(oracle_mod):
    """simulate_gen with device_rng draws the sphere on the GPU; the image matches the oracle
    evaluated on the same (device-drawn) emitters."""
    import oracle

    from idi_amd.simulation import hip, simobj

    obj = simobj.sphere(8048.0, 3000, 0.05)
    obj.rng = np.random.default_rng(9)
    gen = hip.simulate_gen(obj, 24, 100.0, 1e5, K_UM, settings="single nf", maximg=3, device_rng=True,
                           returngpu=True)
    imgs = [g.cpu().numpy() for g in gen]
    assert not np.allclose(imgs[0], imgs[1])
    # the emitters of image i are generate_ball(N, seed_i) with seeds drawn from obj.rng
    obj.rng = np.random.default_rng(9)
    g = oracle.Geometry(24, 24, 100.0, 1e5, K_UM, nf=True)
    from oracle.philox import generate_ball

    for i in range(3):
        seed = int(obj.rng.integers(2**63))
        at = generate_ball(3000, seed=seed, rng_stream=0, radius=0.05)
        assert rel_l2(imgs[i], oracle_mod.field_c(g, at)) <= 1e-4


def test_device_balls_and_gauss_match_restatement():
    import torch

    from idi_amd import _lib
    from idi_amd.util import rotation
    from oracle.philox import generate_balls, generate_gauss

    centres = np.array([[-0.04, 0.0, 0.0], [0.0, 0.01, 0.0], [0.05, 0.0, -0.02]])
    counts = [1000, 0, 2501]
    out = torch.empty((3501, 4), dtype=torch.float64, device="cuda")
    _lib.generate_balls(out, seed=123, centres=centres, counts=counts, rng_stream=2, radius=0.015)
    ref = generate_balls(123, centres, counts, rng_stream=2, radius=0.015)
    got = out.cpu().numpy()
    np.testing.assert_array_equal(got[:, 3], ref[:, 3])
    assert np.abs(got[:, :3] - ref[:, :3]).max() <= 1e-17 * 8
    M = rotation(0.2, 0.4, -0.3)
    g = torch.empty((50_001, 4), dtype=torch.float64, device="cuda")
    _lib.generate_gauss(g, seed=2**50 + 1, sigma=(0.01, 0.02, 0.005), rot=M, rng_stream=1)
    ref = generate_gauss(50_001, seed=2**50 + 1, sigma=(0.01, 0.02, 0.005), rot=M, rng_stream=1)
    got = g.cpu().numpy()
    np.testing.assert_array_equal(got[:, 3], ref[:, 3])
    np.testing.assert_allclose(got[:, :3], ref[:, :3], rtol=1e-13, atol=1e-16)  # 1 ulp at 1e-2 after cancellation


@pytest.mark.parametrize("kind", ["multisphere", "gauss"])
def test_simulate_gen_device_rng_shapes(oracle_mod, kind):
    """simulate_gen(device_rng=True) for multisphere and gauss objects matches the oracle on the
    emitters the device drew (regenerated with the restatement from the same seeds)."""
    import oracle
    from oracle.philox import generate_balls, generate_gauss

    from idi_amd.simulation import hip, simobj

    def make():
        if kind == "multisphere":
            o = simobj.multisphere(8048.0, Natoms=3000, rsphere=0.004, fwhm=0.03, spacing=0.001)
        else:
            o = simobj.gauss(8048.0, 3000, (0.02, 0.03, 0.01))
        o.rng = np.random.default_rng(21)
        return o

    obj = make()
    gen = hip.simulate_gen(obj, 24, 100.0, 1e5, K_UM, settings="single nf", maximg=2, device_rng=True,
                           returngpu=True)
    imgs = [g.cpu().numpy() for g in gen]
    ref_obj = make()
    g = oracle.Geometry(24, 24, 100.0, 1e5, K_UM, nf=True)
    for i in range(2):
        seed = int(ref_obj.rng.integers(2**63))
        if kind == "multisphere":
            centres, counts = ref_obj.centres_counts()
            at = generate_balls(seed, centres, counts, radius=ref_obj.rsphere)
        else:
            at = generate_gauss(3000, seed, ref_obj.fwhm / (2 * np.sqrt(2 * np.log(2))))
        assert rel_l2(imgs[i], oracle_mod.field_c(g, at)) <= 1e-4


def test_device_choice_matches_restatement():
    """Same sites chosen in the same order as the restatement (uniform, weighted, several rounds, rotation)."""
    import torch

    from idi_amd import _lib
    from idi_amd.util import rotation
    from oracle.philox import generate_choice

    rng = np.random.default_rng(3)
    sites = rng.normal(size=(50_000, 3)) * 0.01
    w = np.exp(-np.sum(sites**2, 1) / (2 * 0.008**2))
    M = rotation(0.4, -0.2, 0.9)
    for weights, n, rot in ((None, 30_000, None), (w, 20_000, M), (None, 120_001, None)):
        smp = _lib.ChoiceSampler(sites, weights)
        out = torch.empty((n, 4), dtype=torch.float64, device="cuda")
        smp.draw(out, seed=99, rng_stream=5, rot=rot)
        got = out.cpu().numpy()
        ref = generate_choice(99, n, sites, weights, rot, rng_stream=5)
        np.testing.assert_array_equal(got[:, 3], ref[:, 3])
        np.testing.assert_allclose(got[:, :3], ref[:, :3], rtol=1e-13, atol=1e-17)


def test_simulate_gen_device_rng_crystal(oracle_mod):
    """A simple-cubic crystal with fresh site choice and phases per image, drawn on the device."""
    import oracle
    from oracle.philox import generate_choice

    from idi_amd.simulation import hip, simobj

    def make():
        o = simobj.sc(8048.0, 3000, 0.0003, repeats=(20, 20, 20))
        o.rng = np.random.default_rng(31)
        return o

    obj = make()
    gen = hip.simulate_gen(obj, 24, 100.0, 1e5, K_UM, settings="single nf", maximg=2, device_rng=True,
                           returngpu=True)
    imgs = [g.cpu().numpy() for g in gen]
    assert not np.allclose(imgs[0], imgs[1])
    ref_obj = make()
    g = oracle.Geometry(24, 24, 100.0, 1e5, K_UM, nf=True)
    for i in range(2):
        seed = int(ref_obj.rng.integers(2**63))
        at = generate_choice(seed, 3000, ref_obj._allpos)
        assert rel_l2(imgs[i], oracle_mod.field_c(g, at)) <= 1e-4


def test_device_choice_skips_zero_weights():
    import torch

    from idi_amd import _lib

    sites = np.stack([np.arange(100.0), np.zeros(100), np.zeros(100)], 1)
    w = np.where(np.arange(100) % 2 == 0, 1.0, 0.0)
    smp = _lib.ChoiceSampler(sites, w)
    out = torch.empty((50, 4), dtype=torch.float64, device="cuda")
    smp.draw(out, seed=4)
    chosen = out.cpu().numpy()[:, 0].astype(int)
    assert sorted(chosen) == list(range(0, 100, 2))  # all 50 positive-weight sites, none of the others
