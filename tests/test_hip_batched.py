"""Persistent batched launch (uneven shots, hipGraph replay) vs the single-shot path and the oracle."""
import numpy as np
import pytest

from conftest import K_UM
from gpu_util import ball, fp64_tol, make_handle, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch

    from idi_amd import _lib

    assert torch.cuda.is_available()
    _lib.load()
    return _lib


def _shots(rng, sizes, r):
    return [np.concatenate([ball(rng, n, r), rng.uniform(0, 2 * np.pi, (n, 1))], 1) for n in sizes]


@pytest.mark.parametrize("precision", [1, 0])
def test_batched_equals_single_shot(lib, oracle_mod, precision):
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(77)
    g = Geometry(48, 40, 100.0 * 256 / 48, 1e5, K_UM, nf=True)
    sizes = [1000, 0, 3, 2500, 257, 4100, 1]
    shots = _shots(rng, sizes, 0.03)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    cat = torch.from_numpy(np.concatenate(shots)).cuda()
    hb = make_handle(lib, g, precision)
    hs = make_handle(lib, g, precision)
    for h in (hb, hs):
        h.set_option(lib.OPT_TILE, 1)
        h.set_option(lib.OPT_SPLIT, 1)
    hs.set_option(lib.OPT_SPLIT, 1)
    out = hb.forward_batched(cat, offsets)
    torch.cuda.synchronize()
    assert hb.plan()["graph"] == 1
    out = out.cpu().numpy()
    for s, at in enumerate(shots):
        if sizes[s] == 0:
            assert np.count_nonzero(out[s]) == 0
            continue
        single = hs.forward(torch.from_numpy(at).cuda()).cpu().numpy()
        # same tile, same emitter order, no split: identical arithmetic
        np.testing.assert_array_equal(out[s], single)
        ref = oracle_mod.field_c(g, at)
        assert rel_l2(out[s], ref) <= (1e-4 if precision == 1 else fp64_tol(g))


@pytest.mark.parametrize("tile", [5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15])
def test_batched_pair_tiles_equal_single_shot(lib, oracle_mod, tile):
    """Mirror-pair tiles in the persistent batched kernel: bit-identical to the single-shot
    kernel of the same tile (no split), within 1e-4 of the oracle (75 um at 0.2 m, 10 nm balls,
    amplitudes)."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(78)
    g = Geometry(45, 70, 75.0, 2e5, K_UM, nf=True)
    sizes = [700, 0, 300, 1500]
    shots = _shots(rng, sizes, 0.01)
    amps = [rng.uniform(0.5, 1.0, n) for n in sizes]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    cat = torch.from_numpy(np.concatenate(shots)).cuda()
    acat = torch.from_numpy(np.concatenate(amps)).cuda()
    hb = make_handle(lib, g, 1)
    hs = make_handle(lib, g, 1)
    for h in (hb, hs):
        h.set_option(lib.OPT_TILE, tile)
        h.set_option(lib.OPT_SPLIT, 1)
    out = hb.forward_batched(cat, offsets, acat).cpu().numpy()
    assert hb.plan()["tile"] == tile
    for s, at in enumerate(shots):
        if sizes[s] == 0:
            assert np.count_nonzero(out[s]) == 0
            continue
        single = hs.forward(torch.from_numpy(at).cuda(), torch.from_numpy(amps[s]).cuda()).cpu().numpy()
        np.testing.assert_array_equal(out[s], single)
        assert rel_l2(out[s], oracle_mod.field_c(g, at, amps[s])) <= 1e-4


def test_graph_replay_with_new_phases(lib):
    """Replays of the captured graph, with eager work and synchronisations in between,
    each match a fresh eager (non-graph) evaluation."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(78)
    g = Geometry(64, 64, 100.0, 1e5, K_UM, nf=True)
    sizes = rng.integers(1000, 6000, 12)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    pos = torch.from_numpy(ball(rng, int(offsets[-1]), 0.03)).cuda()
    d = torch.zeros((int(offsets[-1]), 4), dtype=torch.float64, device="cuda")
    d[:, :3] = pos
    hg = make_handle(lib, g, lib.PREC_FP32SPLIT)
    he = make_handle(lib, g, lib.PREC_FP32SPLIT)
    he.set_option(lib.OPT_GRAPH, 0)
    out = torch.empty((len(sizes), 64, 64), dtype=torch.float32, device="cuda")
    for rep in range(4):
        d[:, 3] = torch.rand(d.shape[0], dtype=torch.float64, device="cuda") * 2 * np.pi
        hg.forward_batched(d, offsets, None, out, lib.OUT_INTENSITY, 0.03)
        assert hg.plan()["graph"] == 1
        eager = he.forward_batched(d, offsets, None, None, lib.OUT_INTENSITY, 0.03)
        assert he.plan()["graph"] == 0
        torch.cuda.synchronize()
        x = out.clone()
        torch.cuda.synchronize()
        assert torch.equal(x, eager), rep
        _ = (eager * 2).sum().item()  # eager work between replays


def test_batched_fallback_coarse_geometry(lib, oracle_mod):
    """Pixels of 2 mm at 0.1 m: no tile meets the FP32 bound, the batch runs the FP64 kernel
    shot by shot and still writes complex64."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(79)
    g = Geometry(24, 20, 2000.0, 1e5, K_UM, nf=True)
    sizes = [300, 0, 1000]
    shots = _shots(rng, sizes, 0.03)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    h = make_handle(lib, g, lib.PREC_FP32SPLIT)
    out = h.forward_batched(torch.from_numpy(np.concatenate(shots)).cuda(), offsets).cpu().numpy()
    assert h.plan()["fallback"] == 1 and out.dtype == np.complex64
    assert np.count_nonzero(out[1]) == 0
    for s in (0, 2):
        assert rel_l2(out[s], oracle_mod.field_c(g, shots[s])) <= max(fp64_tol(g), 1e-6)


@pytest.mark.parametrize("tile", [None, 7, 8, 12, 14, 15])
def test_batched_config_shape_rows(lib, oracle_mod, tile):
    """256 x 256 px of 100 um at 0.1 m, shots of N_s ~ U{1000..50000} in a 30 nm ball (a subset of the 256);
    the planner's choice and the mirror-pair tiles with a centre per pair (7, 8, 12, 14, 15)."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(3000)
    sizes = np.array([1000, 50000, 17000, 2500, 33000, 8000])
    shots = _shots(rng, sizes, 0.03)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    g = Geometry(256, 256, 100.0, 1e5, K_UM, nf=True)
    h = make_handle(lib, g, lib.PREC_FP32SPLIT)
    if tile is not None:
        h.set_option(lib.OPT_TILE, tile)
    out = h.forward_batched(torch.from_numpy(np.concatenate(shots)).cuda(), offsets, None, None,
                            lib.OUT_INTENSITY).cpu().numpy()
    rows = [0, 128, 255]
    errs = []
    for s in (0, 1, 3):
        ref = np.abs(oracle_mod.field_c(g, shots[s], rows=rows)) ** 2
        errs.append(rel_l2(out[s][rows], ref))
    print("batched relL2 per shot", errs, h.plan())
    assert max(errs) <= 1e-4
    if tile is not None:
        assert h.plan()["tile"] == tile and h.plan()["phase_bound"] <= 5e-6
