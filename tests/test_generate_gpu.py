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
