"""Parity of the HIP forward kernels (through the C ABI) with the FP64 oracle.

Bars: the literal FP64 kernel matches the oracle to the FP64 formula's own
rounding (per-term phase error of a few ulps of k|r_p|); the FP32-split kernel
matches to <= 1e-4 relative L2 on the intensity (and on the field), the
tolerance of the metric; typically it reaches ~1e-6.
"""
import numpy as np
import pytest

from conftest import K_UM, fp64_phase_tolerance, golden_field, golden_geometry, golden_term_scale
from gpu_util import atoms_ball, fp64_tol, intensity, make_handle, rel_l2, run, sc_ball

pytestmark = pytest.mark.gpu

SPLIT_RTOL = 1e-4


@pytest.fixture(scope="module")
def lib():
    import torch

    from idi_amd import _lib

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    _lib.load()
    return _lib


def test_golden_fp64(lib, golden):
    for case in golden:
        print("golden fp64", case["name"], flush=True)
        g = golden_geometry(case)
        f, _ = run(lib, g, np.array(case["atoms"]), None if case["amp"] is None else np.array(case["amp"]),
                   precision=lib.PREC_FP64)
        tol = 2 * fp64_phase_tolerance(case) * golden_term_scale(case)
        err = np.abs(f - golden_field(case)).max()
        assert err <= tol, (case["name"], err, tol)


def test_golden_split(lib, golden):
    for case in golden:
        print("golden split", case["name"], flush=True)
        g = golden_geometry(case)
        f, h = run(lib, g, np.array(case["atoms"]), None if case["amp"] is None else np.array(case["amp"]))
        plan = h.plan()
        ref = golden_field(case)
        scale = golden_term_scale(case)
        err = np.abs(f - ref).max() / scale
        if "centre" in case["name"]:
            assert plan["fallback"] == 0  # realistic pixel pitch: the split path qualifies
        if plan["fallback"]:
            # no tile met the bound (extreme near field, or a few huge pixels):
            # FP64 kernel, complex64 output
            assert err <= 2e-7 + 2 * fp64_phase_tolerance(case), (case["name"], err)
        else:
            # per-term phase error <= plan bound (rev) -> field error <= 2 pi bound * sum|a| (+ FP32 output)
            assert err <= 2 * np.pi * plan["phase_bound"] + 2e-7, (case["name"], err, plan)
            assert plan["phase_bound"] <= 5e-6


def test_golden_split_every_tile_within_its_bound(lib, golden):
    """Every FP32-split tile variant forced on every golden case: the field error stays within
    2 pi x the variant's own a-priori bound, also where that bound is above the tolerance (a variant
    the planner would not pick); variants a mode does not allow (mirror pairs without the W fold)
    refuse with an error."""
    checked = 0
    for case in golden:
        g = golden_geometry(case)
        atoms = np.array(case["atoms"])
        amp = None if case["amp"] is None else np.array(case["amp"])
        ref = golden_field(case)
        scale = golden_term_scale(case)
        for tile in range(16):
            try:
                f, h = run(lib, g, atoms, amp, tile=tile, split=1)
            except lib.IdisimError:
                continue
            plan = h.plan()
            assert plan["tile"] == tile and plan["fallback"] == 0
            err = np.abs(f - ref).max() / scale
            assert err <= 2 * np.pi * plan["phase_bound"] + 2e-7, (case["name"], tile, err, plan)
            checked += 1
    assert checked >= 8 * len(golden)


def _c1(rng):
    from oracle import Geometry

    return Geometry(128, 128, 75.0, 1e5, K_UM, nf=True), atoms_ball(rng, 1000, 0.05)


def test_ball_1000_128px_fp64_full_image(lib, oracle_mod):
    rng = np.random.default_rng(1000)
    g, at = _c1(rng)
    ref = oracle_mod.field_c(g, at)
    f, _ = run(lib, g, at, precision=lib.PREC_FP64)
    assert rel_l2(f, ref) <= fp64_tol(g)
    assert rel_l2(intensity(f), intensity(ref)) <= fp64_tol(g)


def test_ball_1000_128px_split_full_image(lib, oracle_mod):
    rng = np.random.default_rng(1000)
    g, at = _c1(rng)
    ref = oracle_mod.field_c(g, at)
    f, h = run(lib, g, at)
    assert h.plan()["fallback"] == 0
    e_i = rel_l2(intensity(f), intensity(ref))
    e_f = rel_l2(f, ref)
    print("split relL2 intensity", e_i, "field", e_f, h.plan())
    assert e_i <= SPLIT_RTOL and e_f <= SPLIT_RTOL


def test_sc_100k_512px_rows(lib, oracle_mod):
    """100k of the 155,331 simple-cubic sites (a = 0.3 nm) within 10 nm, 512^2 px of 75 um at 0.2 m."""
    from oracle import Geometry

    rng = np.random.default_rng(2000)
    at, nsites = sc_ball(rng, 100_000, 3e-4, 0.01)
    assert nsites == 155_331
    g = Geometry(512, 512, 75.0, 2e5, K_UM, nf=True)
    rows = [0, 131, 256, 511]
    ref = oracle_mod.field_c(g, at, rows=rows)
    fs, hs = run(lib, g, at)
    f64, _ = run(lib, g, at, precision=lib.PREC_FP64)
    e_split = rel_l2(intensity(fs[rows]), intensity(ref))
    e_64 = rel_l2(intensity(f64[rows]), intensity(ref))
    e_full = rel_l2(intensity(fs), intensity(f64))
    print("sc100k: split vs oracle", e_split, "fp64 vs oracle", e_64, "split vs fp64 full", e_full, hs.plan())
    assert e_64 <= fp64_tol(g)
    assert e_split <= SPLIT_RTOL
    assert e_full <= SPLIT_RTOL
    assert rel_l2(fs, f64) <= SPLIT_RTOL


MODES = [
    dict(nf=True),
    dict(nf=False),
    dict(nf=True, scale=True),
    dict(nf=False, scale=True),
    dict(nf=False, secondorder=True),
    dict(nf=False, secondorder=True, scale=True),
    dict(nf=True, detangles=(0.3, -0.2), detoffset=(1500.0, -700.0)),
    dict(nf=False, detangles=(-0.1, 0.25), detoffset=(-300.0, 900.0)),
]


@pytest.mark.parametrize("mode", range(len(MODES)))
def test_modes_ragged(lib, oracle_mod, mode):
    """Every phase model on ragged sizes: odd detector, N not a multiple of the 256-emitter chunk, amplitudes."""
    from oracle import Geometry

    rng = np.random.default_rng(300 + mode)
    m = dict(MODES[mode])
    g = Geometry(37, 53, 150.0, 1e5, K_UM, detangles=m.pop("detangles", (0.0, 0.0)),
                 detoffset=m.pop("detoffset", (0.0, 0.0)), **m)
    for n, use_amp in ((1, False), (257, True), (700, False)):
        at = atoms_ball(rng, n, 0.04)
        amp = rng.uniform(0.5, 1.0, n) if use_amp else None
        ref = oracle_mod.field_c(g, at, amp)
        f64, _ = run(lib, g, at, amp, precision=lib.PREC_FP64)
        fs, h = run(lib, g, at, amp)
        assert rel_l2(f64, ref) <= fp64_tol(g), (mode, n)
        assert rel_l2(fs, ref) <= SPLIT_RTOL, (mode, n, rel_l2(fs, ref), h.plan())
        Is, _ = run(lib, g, at, amp, kind=lib.OUT_INTENSITY)
        assert rel_l2(Is, intensity(ref)) <= SPLIT_RTOL
        I64, _ = run(lib, g, at, amp, precision=lib.PREC_FP64, kind=lib.OUT_INTENSITY)
        assert rel_l2(I64, intensity(ref)) <= fp64_tol(g)


def test_tile_split_summode_variants_agree(lib, oracle_mod):
    from oracle import Geometry

    rng = np.random.default_rng(44)
    g = Geometry(70, 90, 100.0, 1e5, K_UM, nf=True)
    at = atoms_ball(rng, 3000, 0.03)
    ref = oracle_mod.field_c(g, at)
    for tile in range(4):
        for split in (1, 3, 8):
            for summode in (0, 1):
                f, h = run(lib, g, at, tile=tile, split=split, summode=summode)
                e = rel_l2(f, ref)
                assert e <= SPLIT_RTOL, (tile, split, summode, e)
                p = h.plan()
                assert (p["tile"], p["split"], p["summode"]) == (tile, split, summode)
    for tile in range(4):
        for split in (1, 5):
            f, _ = run(lib, g, at, precision=lib.PREC_FP64, tile=tile, split=split)
            assert rel_l2(f, ref) <= fp64_tol(g), (tile, split)


def test_empty_and_errors(lib):
    import torch

    from oracle import Geometry

    g = Geometry(8, 8, 100.0, 1e5, K_UM, nf=True)
    h = make_handle(lib, g, lib.PREC_FP32SPLIT)
    out = torch.full((8, 8), 3.0, dtype=torch.complex64, device="cuda")
    h.forward(torch.empty((0, 4), dtype=torch.float64, device="cuda"), None, out)
    torch.cuda.synchronize()
    assert torch.count_nonzero(out).item() == 0
    with pytest.raises(ValueError):
        h.forward(torch.zeros((4, 3), dtype=torch.float64, device="cuda"))
    with pytest.raises(lib.IdisimError):
        h.set_option(lib.OPT_TILE, 16)  # past the last variant
    with pytest.raises(lib.IdisimError):
        h.set_option(lib.OPT_TILE, 4)  # reserved, no kernel
    with pytest.raises(lib.IdisimError):
        h.set_option(999, 1)
    with pytest.raises(lib.IdisimError):
        lib.Handle(0, 8, 100.0, 1e5, K_UM)
    # adjoint is FP32-split only
    h64 = make_handle(lib, g, lib.PREC_FP64)
    e = torch.zeros((4, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(lib.IdisimError):
        h64.adjoint(e, torch.zeros((8, 8), dtype=torch.complex64, device="cuda"),
                    torch.zeros((8, 8), dtype=torch.float32, device="cuda"))


def test_extent_measured_on_device(lib):
    import ctypes

    import torch

    rng = np.random.default_rng(5)
    at = atoms_ball(rng, 10_000, 0.05)
    d = torch.from_numpy(at).cuda()
    ext = ctypes.c_double()
    lib.check(lib.load().idisim_extent(ctypes.c_void_p(d.data_ptr()), len(at), ctypes.byref(ext), None))
    assert ext.value == pytest.approx(np.linalg.norm(at[:, :3], axis=1).max(), rel=1e-15)


def test_extent_underestimate_is_reported(lib):
    """An extent of half max|r_j| passed to the FP32-split kernels: the record stage flags it,
    idisim_get_plan reports it once the kernel has finished, and the handle's next forward,
    batched or adjoint call fails with IDISIM_E_PRECISION (then the flag is clear again).  A
    correct extent, and the measured one (extent <= 0), raise nothing."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(9)
    at = atoms_ball(rng, 3000, 0.01)
    rmax = float(np.linalg.norm(at[:, :3], axis=1).max())
    g = Geometry(64, 64, 75.0, 2e5, K_UM, nf=True)
    h = make_handle(lib, g, lib.PREC_FP32SPLIT)
    d = torch.from_numpy(at).cuda()
    for ext in (rmax, 0.0, 2 * rmax):
        h.forward(d, extent=ext)
        torch.cuda.synchronize()
        assert h.plan()["extent_exceeded"] == 0, ext
    good = h.forward(d, extent=rmax).cpu().numpy()

    def next_call_fails(call):
        torch.cuda.synchronize()
        assert h.plan()["extent_exceeded"] == 1
        with pytest.raises(lib.IdisimError) as ei:
            call()
        assert ei.value.code == lib.E_PRECISION and "extent" in str(ei.value)
        assert h.plan()["extent_exceeded"] == 1  # the call that found the flag reports it
        out = h.forward(d, extent=rmax)  # the flag was cleared: the next call runs
        torch.cuda.synchronize()
        assert h.plan()["extent_exceeded"] == 0
        np.testing.assert_array_equal(out.cpu().numpy(), good)

    # single shot, with a forced pair tile and the planner's choice
    h.forward(d, extent=0.5 * rmax)
    next_call_fails(lambda: h.forward(d, extent=rmax))
    # batched: two shots, the flag found by the next batched call
    offs = np.array([0, 1000, 3000])
    h.forward_batched(d, offs, extent=0.5 * rmax)
    next_call_fails(lambda: h.forward_batched(d, offs, extent=rmax))
    # adjoint
    f = h.forward(d, extent=rmax)
    dl = torch.ones(g.n0, g.n1, dtype=torch.float32, device="cuda")
    h.adjoint(d, f, dl, extent=0.5 * rmax)
    next_call_fails(lambda: h.adjoint(d, f, dl, extent=rmax))
    # an FP64 handle does not depend on the extent and checks nothing
    h64 = make_handle(lib, g, lib.PREC_FP64)
    h64.forward(d, extent=0.5 * rmax)
    h64.forward(d, extent=0.5 * rmax)
    torch.cuda.synchronize()
    assert h64.plan()["extent_exceeded"] == 0


def test_fallback_and_strict_precision(lib, oracle_mod):
    """Extreme near field (10 um sphere at 1 mm, the geometry of the reference's shape test)."""
    from oracle import Geometry

    rng = np.random.default_rng(6)
    k = 2 * np.pi / (1.24 / 100)
    g = Geometry(4, 4, 10.0, 1000.0, k, nf=True)
    at = atoms_ball(rng, 100, 10.0)
    ref = oracle_mod.field_c(g, at)
    f, h = run(lib, g, at)
    assert h.plan()["fallback"] == 1
    assert rel_l2(f, ref) <= max(fp64_tol(g), 1e-6)
    h.set_option(lib.OPT_FALLBACK, 0)
    import torch

    with pytest.raises(lib.IdisimError) as ei:
        h.forward(torch.from_numpy(at).cuda())
    assert ei.value.code == lib.E_PRECISION


def test_autotune_cache_across_extents(lib, oracle_mod):
    """One autotuning handle, the same emitter count at three extents: a 1 nm ball (FP32-split
    path, tuned and cached), a 10 um ball (no tile meets the bound: FP64 fallback) and the 1 nm
    ball again.  A cached choice is reused only for the kernel family and tiles it was tuned for."""
    from oracle import Geometry

    rng = np.random.default_rng(8)
    k = 2 * np.pi / (1.24 / 100)
    g = Geometry(4, 4, 10.0, 1000.0, k, nf=True)
    small = atoms_ball(rng, 100, 0.001)
    large = atoms_ball(rng, 100, 10.0)
    h = make_handle(lib, g, lib.PREC_FP32SPLIT)
    h.set_option(lib.OPT_AUTOTUNE, 1)
    for at, fb in ((small, 0), (large, 1), (small, 0)):
        f, _ = run(lib, g, at, handle=h)
        assert h.plan()["fallback"] == fb
        tol = SPLIT_RTOL if fb == 0 else max(fp64_tol(g), 1e-6)
        assert rel_l2(f, oracle_mod.field_c(g, at)) <= tol


@pytest.mark.parametrize("mode", ["nf", "ff", "ff_secondorder", "nf_scale_amp", "nf_rot", "ff_rot"])
def test_pair_tiles_match_oracle(lib, oracle_mod, mode):
    """Mirror-pair tiles (5: one 2 x 2 quad per lane, 6: two, 7: one with a centre per pair, 8: four
    pairs about a 2 x 4 block, 9..12: 5 and 7 with the pair phasors rotated from a per-wave reference
    to first or second order, 13: as 9 with the rotation angle linear in the quad offset) on the
    pixel pitch and distance of
    the 100k-site image (75 um at 0.2 m, sites of a 10 nm ball), on an odd detector so that quads
    reach past the last row and column; every emitter split and output kind; a rotated and
    offset detector."""
    from oracle import Geometry

    rng = np.random.default_rng(45)
    rot = mode.endswith("_rot")
    g = Geometry(37, 53, 75.0, 2e5, K_UM, detoffset=(900.0, -1500.0) if rot else (0.0, 0.0),
                 detangles=(0.05, -0.08) if rot else (0.0, 0.0), nf=mode.startswith("nf"),
                 scale="scale" in mode, secondorder="secondorder" in mode)
    at, _ = sc_ball(rng, 3000, 3e-4, 0.01)
    amp = rng.uniform(0.5, 1.0, len(at)) if "amp" in mode else None
    ref = oracle_mod.field_c(g, at, amp)
    for tile in (5, 6, 7, 8, 9, 10, 11, 12, 13):
        for split in (1, 3):
            f, h = run(lib, g, at, amp, tile=tile, split=split)
            p = h.plan()
            assert (p["tile"], p["split"]) == (tile, split)
            assert p["phase_bound"] <= 5e-6
            assert rel_l2(f, ref) <= SPLIT_RTOL, (tile, split, rel_l2(f, ref))
            assert rel_l2(intensity(f), intensity(ref)) <= SPLIT_RTOL
        I, _ = run(lib, g, at, amp, kind=lib.OUT_INTENSITY, tile=tile, split=2)
        assert rel_l2(I, intensity(ref)) <= SPLIT_RTOL


def test_pair_tiles_eligibility(lib, oracle_mod):
    """Pixels of 100 um at 0.1 m and a 30 nm ball: the pair remainder (~1.2e-5 rev) exceeds the
    5e-6 rev tolerance, so the planner never picks tiles 5, 6, 9 or 10 and a forced one reports its
    bound; tile 7 (a centre per pair) drops no remainder and qualifies; so does tile 12, whose
    second-order rotation from the wave reference stays ~1e-7 rev, while the first-order tile 11
    (~1.5e-5 rev) does not; the per-term Kahan mode has no pair kernel."""
    from oracle import Geometry

    rng = np.random.default_rng(46)
    # the corner region of a 256^2 detector: offset by 96 pixels
    g = Geometry(64, 64, 100.0, 1e5, K_UM, detoffset=(9600.0, 9600.0), nf=True)
    at = atoms_ball(rng, 600, 0.03)
    ref = oracle_mod.field_c(g, at)
    _, h = run(lib, g, at)
    assert h.plan()["tile"] not in (5, 6, 9, 10, 11, 13)
    for tile in (5, 9, 10, 11, 13):
        f, h = run(lib, g, at, tile=tile, split=1)
        assert h.plan()["phase_bound"] > 5e-6
    for tile in (7, 8, 12):
        f, h = run(lib, g, at, tile=tile, split=2)
        assert h.plan()["phase_bound"] <= 5e-6
        assert rel_l2(f, ref) <= SPLIT_RTOL and rel_l2(intensity(f), intensity(ref)) <= SPLIT_RTOL
    for tile in (5, 7, 8, 9, 12):
        with pytest.raises(lib.IdisimError):
            run(lib, g, at, tile=tile, split=1, summode=1)
    # the same detector at 0.5 m: eligible, and within the bound of the oracle
    g2 = Geometry(64, 64, 100.0, 5e5, K_UM, detoffset=(9600.0, 9600.0), nf=True)
    f, h = run(lib, g2, at, tile=6, split=1)
    assert h.plan()["phase_bound"] <= 5e-6
    assert rel_l2(f, oracle_mod.field_c(g2, at)) <= SPLIT_RTOL


def test_pair_tiles_tiny_detector_and_shots(lib, oracle_mod):
    """Pair tiles where every quad reaches past the detector (5 x 3 and 1 x 1 pixels) and with one or
    two emitters, single shot and batched."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(47)
    for n0, n1 in ((5, 3), (1, 1)):
        g = Geometry(n0, n1, 75.0, 2e5, K_UM, nf=True)
        for n in (1, 2, 300):
            at, _ = sc_ball(rng, n, 3e-4, 0.01)
            ref = oracle_mod.field_c(g, at)
            for tile in (5, 6, 7, 8, 9, 10, 11, 12, 13):
                f, h = run(lib, g, at, tile=tile, split=1)
                assert f.shape == (n0, n1)
                assert rel_l2(f, ref) <= SPLIT_RTOL, (n0, n1, n, tile)
            h = make_handle(lib, g, lib.PREC_FP32SPLIT)
            h.set_option(lib.OPT_TILE, {1: 9, 2: 12, 300: 8}[n])
            offsets = np.array([0, n, n, 2 * n], dtype=np.int64)
            out = h.forward_batched(torch.from_numpy(np.concatenate([at, at])).cuda(), offsets).cpu().numpy()
            assert np.count_nonzero(out[1]) == 0
            assert rel_l2(out[0], ref) <= SPLIT_RTOL and rel_l2(out[2], ref) <= SPLIT_RTOL


def test_issue_rates(lib):
    """The v_fma_f32 and v_sin_f32 / v_cos_f32 issue rates the benchmark states the forward kernels'
    ceiling with: positive, transcendentals slower than FMA, and of the order of one instruction per
    SIMD lane per clock (between 1e12 and 1e15 lane operations per second on any CDNA device)."""
    fma, trans = lib.issue_rates()
    print("issue rates", fma, trans)
    assert 1e12 < trans < fma < 1e15
    with pytest.raises(lib.IdisimError):
        lib.issue_rates(999)


def test_fp64_phases_on_both_sides_of_the_sincos_reduction_limit(lib, oracle_mod):
    """The FP64 kernel's sincos reduces by pi/2 below |x| = 2^26 C1 (1.05e8 rad) and calls the library
    above.  A 32 x 32 detector of 100 um pixels offset by 16 mm at 0.1 m spans near-field phases
    k (|r_p| - z_p) from about 8.5e7 to 1.3e8 rad, so pixels of both sides are summed; the image
    against the FP64 oracle within the two evaluations' rounding of k |r_p|, and the batched FP64
    kernel's image bit for bit."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(48)
    g = Geometry(32, 32, 100.0, 1e5, K_UM, detoffset=(16000.0, 16000.0), nf=True)
    xs = (np.arange(32) - 16) * 100.0 + 16000.0
    rho2 = xs[:, None] ** 2 + xs[None, :] ** 2
    phase = K_UM * (np.sqrt(rho2 + 1e10) - 1e5)
    lim = float.fromhex("0x1.921fb54p+26")
    assert phase.min() < 0.9 * lim and phase.max() > 1.1 * lim, (phase.min(), phase.max())
    at = atoms_ball(rng, 400, 0.03)
    ref = oracle_mod.field_c(g, at)
    f, h = run(lib, g, at, precision=lib.PREC_FP64, split=1)  # no split: the batched kernel's sums
    e = rel_l2(f, ref)
    print("fp64 across the sincos limit: relL2", e, "bound", fp64_tol(g))
    assert e <= fp64_tol(g)
    out = h.forward_batched(torch.from_numpy(at).cuda(), [0, len(at)]).cpu().numpy()
    np.testing.assert_array_equal(out[0], f)
