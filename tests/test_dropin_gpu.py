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


def test_reference_cli_calls_match_oracle(oracle_mod):
    """The reference CLI's calls (scripts/idi_sim.py:83,85): sim.cuda.simulate(Nimg, obj, Ndet,
    pixelsize, detz, k) with cuda.py's default settings ("double", far field) and
    sim.cpu.simulate(...) with cpu.py's ("", far field, FP64); cpu.simulate with "nf scale";
    cpu.simulate_gen yields one reused complex128 buffer (cpu.py:195,212)."""
    import oracle

    from idi_amd import simulation as sim

    obj = sim.simobj.sphere(8048.0, 400, 0.05)
    g = oracle.Geometry(24, 24, 75.0, 1e5, obj.k)
    for backend in (sim.cuda, sim.cpu):
        obj.rng = np.random.default_rng(21)
        res = backend.simulate(2, obj, 24, 75.0, 1e5, obj.k)
        assert res.shape == (2, 24, 24) and res.dtype == np.complex128
        obj.rng = np.random.default_rng(21)
        for i in range(2):
            assert rel_l2(res[i], oracle_mod.field_c(g, obj.get())) <= fp64_tol(g), (backend.__name__, i)
    gnf = oracle.Geometry(24, 24, 75.0, 1e5, obj.k, nf=True, scale=True)
    obj.rng = np.random.default_rng(22)
    res = sim.cpu.simulate(2, obj, 24, 75.0, 1e5, settings="nf scale single")  # "single" is ignored: FP64
    obj.rng = np.random.default_rng(22)
    for i in range(2):
        assert rel_l2(res[i], oracle_mod.field_c(gnf, obj.get())) <= fp64_tol(gnf)
    gen = sim.cpu.simulate_gen(obj, (24, 20), 75.0, 1e5, maximg=3)
    imgs = [next(gen) for _ in range(3)]
    assert imgs[0] is imgs[1] is imgs[2] and imgs[0].dtype == np.complex128 and imgs[0].shape == (24, 20)
    with pytest.raises(StopIteration):
        next(gen)


def test_handle_cache_eviction_keeps_live_handles():
    """Nine geometries in use at once (the cache holds eight): a running simulate_gen and a
    pending backward() keep working after their cached handles were evicted."""
    import torch

    from idi_amd.simulation import autograd, hip, simobj

    hip._HANDLES.clear()
    obj = simobj.sphere(8048.0, 200, 0.03)
    obj.rng = np.random.default_rng(31)
    gen = hip.simulate_gen(obj, 12, 75.0, 1e5, K_UM, settings="single nf", maximg=3)
    first = next(gen)
    pos = torch.tensor(obj.get()[:, :3], device="cuda", requires_grad=True)
    ph = torch.zeros(200, dtype=torch.float64, device="cuda", requires_grad=True)
    I = autograd.intensity(pos, ph, Ndet=10, pixelsize=80.0, detz=1e5, k=K_UM)
    for j in range(9):  # nine more geometries evict both handles from the cache
        hip.simulate(1, obj, 8 + j, 70.0, 1e5, K_UM, settings="single nf", verbose=False)
    assert len(hip._HANDLES) == 8
    rest = list(gen)
    assert len(rest) == 2 and all(np.all(np.isfinite(r)) for r in [first] + rest)
    I.sum().backward()
    assert torch.isfinite(pos.grad).all() and torch.isfinite(ph.grad).all()
