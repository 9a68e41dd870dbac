"""The idi.simulation drop-in (simulate / simulate_gen / simulate_batched) on the GPU.

Mirrors idi's tests/test_sim.py:99-141 (shapes of cpu/cuda.simulate) and adds
value checks against the oracle for the same emitters.
"""
import numpy as np
import pytest

from conftest import K_UM
from gpu_util import fp64_tol, rel_l2

pytestmark = pytest.mark.gpu


def test_reference_shape_test():
    """tests/test_sim.py:103-109 and 129-135: sphere(100, 100, 10), 4 px of 10 um at 1000 um."""
    from idi_amd.simulation import hip, simobj

    obj = simobj.sphere(100, 100, 10)
    res = hip.simulate(2, obj, 4, 10, 1000, obj.k, verbose=False)
    assert res.shape == (2, 4, 4) and res.dtype == np.complex128
    res = hip.simulate(2, obj, 4, 10, 1000, obj.k, settings="single nf", verbose=False)
    assert res.shape == (2, 4, 4) and np.all(np.isfinite(res))


@pytest.mark.parametrize("settings", ["double", "single", "single nf", "nf", "single nf scale", "secondorder single"])
def test_simulate_matches_oracle(oracle_mod, settings):
    import oracle

    from idi_amd.simulation import hip, simobj

    obj = simobj.sphere(8048.0, 500, 0.05)
    obj.rng = np.random.default_rng(7)
    n0, n1 = 33, 40
    res = hip.simulate(3, obj, (n0, n1), 75.0, 1e5, K_UM, settings=settings, verbose=False, detoffset=(100.0, -50.0))
    obj.rng = np.random.default_rng(7)
    g = oracle.Geometry(n0, n1, 75.0, 1e5, K_UM, detoffset=(100.0, -50.0), nf="nf" in settings,
                        scale="scale" in settings, secondorder="secondorder" in settings)
    tol = 1e-4 if "single" in settings else fp64_tol(g)
    for i in range(3):
        ref = oracle_mod.field_c(g, obj.get())
        assert rel_l2(res[i], ref) <= tol, (settings, i)


def test_simulate_gen_types_and_pipeline():
    import torch

    from idi_amd.simulation import hip, simobj

    obj = simobj.sphere(8048.0, 300, 0.05)
    obj.rng = np.random.default_rng(3)
    gen = hip.simulate_gen(obj, 16, 75.0, 1e5, K_UM, settings="single nf", maximg=3)
    imgs = list(gen)
    assert len(imgs) == 3 and all(i.dtype == np.complex64 and i.shape == (16, 16) for i in imgs)
    assert not np.allclose(imgs[0], imgs[1])
    gen = hip.simulate_gen(obj, 16, 75.0, 1e5, K_UM, settings="double", maximg=2, returngpu=True)
    imgs = list(gen)
    assert all(isinstance(i, torch.Tensor) and i.is_cuda and i.dtype == torch.complex128 for i in imgs)
    res = hip.simulate(2, obj, 16, 75.0, 1e5, K_UM, settings="single", returngpu=True, verbose=False)
    assert isinstance(res, torch.Tensor) and res.shape == (2, 16, 16)
    inten = hip.simulate(2, obj, 16, 75.0, 1e5, K_UM, settings="single nf", intensity=True, verbose=False)
    assert inten.dtype == np.float64 and np.all(inten >= 0)


def test_fixed_positions_random_phases_statistics():
    """rndPos=False, rndPhase=True: the mean intensity over shots tends to sum a^2 = N."""
    from idi_amd.simulation import hip, simobj

    obj = simobj.sphere(8048.0, 200, 0.05)
    obj.rng = np.random.default_rng(11)
    obj.rndPos = False
    res = hip.simulate(300, obj, 8, 2000.0, 1e5, K_UM, settings="single nf", intensity=True, verbose=False)
    assert abs(res.mean() / 200 - 1) < 0.05


def test_amplitudes_extension(oracle_mod):
    import oracle

    from idi_amd.simulation import hip, simobj

    obj = simobj.gauss(8048.0, 400, [0.02, 0.02, 0.02])
    obj.rng = np.random.default_rng(12)
    obj.amplitudes = np.linspace(0.5, 1.0, 400)
    res = hip.simulate(1, obj, 20, 100.0, 1e5, K_UM, settings="single nf", verbose=False)
    obj.rng = np.random.default_rng(12)
    g = oracle.Geometry(20, 20, 100.0, 1e5, K_UM, nf=True)
    ref = oracle_mod.field_c(g, obj.get(), obj.amplitudes)
    assert rel_l2(res[0], ref) <= 1e-4


def test_auto_is_hip():
    from idi_amd import simulation

    assert simulation.auto is simulation.hip
