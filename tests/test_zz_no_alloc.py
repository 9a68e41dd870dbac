"""idisim_forward_noalloc through the C ABI, against idisim_forward (collected last: a failure here
hides no other test).

The no-alloc entry point keeps its emitter-split partial sums in the library's static
device scratch.  Checked here:
  * two calls in a row (100000 simple-cubic emitters on the 512^2 detector at 0.2 m, then
    3000 emitters in a 10 nm ball) give the same bits as idisim_forward with the same plan,
    for an FP32-split and an FP64 handle, field and intensity outputs; and the same on a
    100 x 77 detector (partly filled edge tiles) for a mirror-pair and a per-pixel tile;
  * without a tuned plan (heuristic plan, measured extent) the result is within 1e-4 of the
    FP64 oracle rows;
  * the calls leave the device's free memory unchanged;
  * a call that would need more scratch than the static buffer holds (the FP64 fallback of a
    2048^2 detector: 128 MiB of partials) fails with IDISIM_E_INVALID.
"""
import numpy as np
import pytest

from conftest import K_UM
from gpu_util import atoms_ball, fp64_tol, rel_l2, sc_ball

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch

    from idi_amd import _lib

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    _lib.load()
    return _lib


def _inputs():
    import torch

    rng = np.random.default_rng(7700)
    big, _ = sc_ball(rng, 100_000, 3e-4, 0.01)
    small = atoms_ball(rng, 3000, 0.01)
    return [(at, torch.from_numpy(at).cuda(), float(np.linalg.norm(at[:, :3], axis=1).max())) for at in (big, small)]


@pytest.mark.parametrize("precision", ["split", "fp64"])
def test_noalloc_equals_forward_two_calls(lib, precision):
    import torch

    prec = lib.PREC_FP32SPLIT if precision == "split" else lib.PREC_FP64
    # splits whose partials (16 B or 32 B per pixel and split) fit the 64 MiB static buffer
    cap = (64 << 20) // (512 * 512 * (16 if precision == "split" else 32))
    ha = lib.Handle(512, 512, 75.0, 2e5, K_UM, nf=True, precision=prec)
    hb = lib.Handle(512, 512, 75.0, 2e5, K_UM, nf=True, precision=prec)
    for kind in (lib.OUT_FIELD, lib.OUT_INTENSITY):
        for at, d, ext in _inputs():
            print(f"noalloc {precision} kind={kind} n={len(at)}", flush=True)
            ha.set_option(lib.OPT_TILE, -1)
            ha.set_option(lib.OPT_SPLIT, 0)
            ha.forward(d, None, None, kind, ext)  # the planner's (autotuned) plan
            pa = ha.plan()
            split = min(pa["split"], cap)
            ha.set_option(lib.OPT_TILE, pa["tile"])
            ha.set_option(lib.OPT_SPLIT, split)
            ref = ha.forward(d, None, None, kind, ext)
            pa = ha.plan()
            hb.set_option(lib.OPT_TILE, pa["tile"])
            hb.set_option(lib.OPT_SPLIT, split)
            out = torch.full_like(ref, float("nan"))
            hb.forward_noalloc(d, out, None, kind, ext)
            torch.cuda.synchronize()
            pb = hb.plan()
            assert (pb["tile"], pb["split"], pb["fallback"]) == (pa["tile"], pa["split"], pa["fallback"]), (pa, pb)
            assert torch.equal(out, ref), (precision, kind, len(at))


@pytest.mark.parametrize("tile", [11, 0])
def test_noalloc_ragged_detector(lib, tile):
    """A detector of 100 x 77 pixels (no multiple of 16 or 32): the partial sums of the partly
    filled edge tiles in the static buffer, for a mirror-pair tile (11) and a per-pixel tile (0) at
    4 splits, give idisim_forward's bits on two calls in a row (1024, then 300 emitters)."""
    import torch

    rng = np.random.default_rng(7701)
    ha = lib.Handle(100, 77, 75.0, 1e5, K_UM, nf=True, precision=lib.PREC_FP32SPLIT)
    hb = lib.Handle(100, 77, 75.0, 1e5, K_UM, nf=True, precision=lib.PREC_FP32SPLIT)
    for h in (ha, hb):
        h.set_option(lib.OPT_TILE, tile)
        h.set_option(lib.OPT_SPLIT, 4)
    for n in (1024, 300):
        at = atoms_ball(rng, n, 0.02)
        d = torch.from_numpy(at).cuda()
        ext = float(np.linalg.norm(at[:, :3], axis=1).max())
        for kind in (lib.OUT_FIELD, lib.OUT_INTENSITY):
            print(f"noalloc ragged tile={tile} n={n} kind={kind}", flush=True)
            ref = ha.forward(d, None, None, kind, ext)
            out = torch.full_like(ref, float("nan"))
            hb.forward_noalloc(d, out, None, kind, ext)
            torch.cuda.synchronize()
            p = hb.plan()
            assert (p["tile"], p["split"], p["fallback"]) == (tile, 4, 0) and p["phase_bound"] <= 5e-6, p
            assert torch.equal(out, ref), (tile, n, kind)


def test_noalloc_heuristic_plan_matches_oracle(lib, oracle_mod):
    import torch

    from oracle import Geometry

    g = Geometry(512, 512, 75.0, 2e5, K_UM, nf=True)
    rows = [0, 256, 511]
    h = lib.Handle(512, 512, 75.0, 2e5, K_UM, nf=True, precision=lib.PREC_FP32SPLIT)
    out = torch.empty((512, 512), dtype=torch.complex64, device="cuda")
    for at, d, ext in _inputs():
        for e in (ext, 0.0):  # given, and measured through the static buffer
            h.forward_noalloc(d, out, None, lib.OUT_FIELD, e)
            torch.cuda.synchronize()
            p = h.plan()
            assert p["fallback"] == 0 and p["split"] * 512 * 512 * 16 <= 64 << 20, p
            ref = oracle_mod.field_c(g, at, rows=rows)
            f = out.cpu().numpy()[rows]
            assert rel_l2(np.abs(f) ** 2, np.abs(ref) ** 2) <= 1e-4, (len(at), e, p)
    h64 = lib.Handle(512, 512, 75.0, 2e5, K_UM, nf=True, precision=lib.PREC_FP64)
    out64 = torch.empty((512, 512), dtype=torch.complex128, device="cuda")
    at, d, ext = _inputs()[1]
    h64.forward_noalloc(d, out64, None, lib.OUT_FIELD, ext)
    assert rel_l2(out64.cpu().numpy()[rows], oracle_mod.field_c(g, at, rows=rows)) <= fp64_tol(g)


def test_noalloc_allocates_nothing(lib):
    import torch

    inputs = _inputs()
    warm = lib.Handle(512, 512, 75.0, 2e5, K_UM, nf=True, precision=lib.PREC_FP32SPLIT)
    out = torch.empty((512, 512), dtype=torch.complex64, device="cuda")
    for _, d, ext in inputs:  # load the kernels this plan uses
        warm.forward_noalloc(d, out, None, lib.OUT_FIELD, ext)
    fresh = lib.Handle(512, 512, 75.0, 2e5, K_UM, nf=True, precision=lib.PREC_FP32SPLIT)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _, d, ext in inputs:
        fresh.forward_noalloc(d, out, None, lib.OUT_FIELD, ext)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    assert fresh.plan()["split"] >= 1


def test_noalloc_rejects_a_call_beyond_the_static_buffer(lib):
    import torch

    h = lib.Handle(2048, 2048, 75.0, 2e5, K_UM, nf=True, precision=lib.PREC_FP32SPLIT)
    h.set_option(lib.OPT_TOL_PPB, 1)  # no tile meets 1e-9 rev: the FP64 kernel with complex64 output
    d = torch.from_numpy(atoms_ball(np.random.default_rng(3), 500, 0.01)).cuda()
    out = torch.empty((2048, 2048), dtype=torch.complex64, device="cuda")
    with pytest.raises(lib.IdisimError) as ei:
        h.forward_noalloc(d, out, None, lib.OUT_FIELD, 0.01)
    assert ei.value.code == -1 and "static buffer" in str(ei.value)
