"""Hand-written adjoint kernel vs the FP64 analytic gradient of the oracle and FP64 central differences.

Loss L = sum_p (I_p - T_p)^2 with T from a reference configuration; gradients with
respect to positions (3N), phases (N) and amplitudes (N).
"""
import numpy as np
import pytest

from conftest import K_UM
from gpu_util import ball, make_handle, rel_l2

pytestmark = pytest.mark.gpu

GRAD_RTOL = 2e-5  # measured: 2e-7 .. 2.1e-6 (FP32 per-tile sums folded into FP64)


@pytest.fixture(scope="module")
def lib():
    import torch

    from idi_amd import _lib

    assert torch.cuda.is_available()
    _lib.load()
    return _lib


def _problem(rng, n, r, geom, oracle_mod, sigma=5e-5):
    at0 = np.concatenate([ball(rng, n, r), rng.uniform(0, 2 * np.pi, (n, 1))], 1)
    target = np.abs(oracle_mod.field_c(geom, at0)) ** 2
    at = at0.copy()
    at[:, :3] += rng.normal(0, sigma, (n, 3))
    return at, target


def _gpu_grad(lib, geom, at, target, amp=None):
    import torch

    h = make_handle(lib, geom, lib.PREC_FP32SPLIT)
    d = torch.from_numpy(at).cuda()
    da = None if amp is None else torch.from_numpy(amp).cuda()
    field = h.forward(d, da)
    I = (field.real ** 2 + field.imag ** 2).double()
    dLdI = (2 * (I - torch.from_numpy(target).cuda())).float().contiguous()
    gpos, gph, gamp = h.adjoint(d, field, dLdI, da)
    torch.cuda.synchronize()
    return gpos.cpu().numpy().astype(np.float64), gph.cpu().numpy().astype(np.float64), gamp.cpu().numpy(), h


@pytest.mark.parametrize("mode", ["nf", "ff", "ff_secondorder", "nf_scale", "ff_scale"])
def test_adjoint_vs_fp64_analytic(lib, oracle_mod, mode):
    from oracle import Geometry

    rng = np.random.default_rng(91)
    g = Geometry(64, 48, 300.0, 1e5, K_UM, nf=mode.startswith("nf"), secondorder="secondorder" in mode,
                 scale="scale" in mode)
    at, target = _problem(rng, 600, 0.02, g, oracle_mod)
    amp = rng.uniform(0.5, 1, 600)
    field = oracle_mod.field_c(g, at, amp)
    dLdI = 2 * (np.abs(field) ** 2 - target)
    rpos, rph, ramp = oracle_mod.adjoint_c(g, at, dLdI, field=field, amp=amp)
    gpos, gph, gamp, h = _gpu_grad(lib, g, at, target, amp)
    e = (rel_l2(gpos, rpos), rel_l2(gph, rph), rel_l2(gamp, ramp))
    print(mode, "adjoint relL2 pos/phase/amp", e, h.plan())
    assert max(e) <= GRAD_RTOL


def test_adjoint_4096_emitters_256px(lib, oracle_mod):
    """4096 emitters in a 20 nm ball perturbed by N(0, 0.05 nm), 256^2 px of 75 um at 0.1 m (near field)."""
    from oracle import Geometry

    rng = np.random.default_rng(4000)
    g = Geometry(256, 256, 75.0, 1e5, K_UM, nf=True)
    at, target = _problem(rng, 4096, 0.02, g, oracle_mod)
    field = oracle_mod.field_c(g, at)
    dLdI = 2 * (np.abs(field) ** 2 - target)
    rpos, rph, _ = oracle_mod.adjoint_c(g, at, dLdI, field=field)
    gpos, gph, _, h = _gpu_grad(lib, g, at, target)
    e = (rel_l2(gpos, rpos), rel_l2(gph, rph))
    print("4096 adjoint relL2 pos/phase", e, h.plan())
    assert max(e) <= GRAD_RTOL
    # central differences of the loss in long double (h = 1e-4 nm for positions, 1e-4 rad for
    # phases) on 24 random position and 8 random phase coordinates
    sel = [(int(j), int(c)) for j, c in zip(rng.choice(4096, 32, replace=False), rng.integers(0, 3, 32))]
    sel = sel[:24] + [(j, 3) for j, _ in sel[24:]]
    fd = oracle_mod.loss_central_diff(g, at, target, sel, [1e-7 if c < 3 else 1e-4 for _, c in sel])
    gpu = np.array([gpos[j, c] if c < 3 else gph[j] for j, c in sel])
    ana = np.array([rpos[j, c] if c < 3 else rph[j] for j, c in sel])
    rms = np.array([np.sqrt(np.mean(gpos**2)) if c < 3 else np.sqrt(np.mean(gph**2)) for _, c in sel])
    err_gpu = np.abs(gpu - fd) / (np.abs(fd) + rms)
    err_ana = np.abs(ana - fd) / (np.abs(fd) + rms)
    print("finite differences: max |gpu - fd| / (|fd| + rms)", err_gpu.max(), "analytic FP64", err_ana.max(),
          "min |fd| / rms", (np.abs(fd) / rms).min())
    # each coordinate within 3e-4 of its own value (plus 3e-5 of the rms gradient of its kind)
    tol = 3e-4 * np.abs(fd) + 3e-5 * rms
    bad = np.flatnonzero(np.abs(gpu - fd) > tol)
    assert len(bad) == 0, [(sel[i], gpu[i], fd[i], ana[i]) for i in bad]
    assert np.all(np.abs(ana - fd) <= tol), "the analytic FP64 gradient disagrees with the differences"


def test_autograd_function(lib, oracle_mod):
    import torch

    from idi_amd.simulation.autograd import intensity
    from oracle import Geometry

    rng = np.random.default_rng(92)
    g = Geometry(32, 32, 150.0, 1e5, K_UM, nf=True)
    at, target = _problem(rng, 300, 0.02, g, oracle_mod)
    pos = torch.from_numpy(at[:, :3].copy()).cuda().requires_grad_(True)
    ph = torch.from_numpy(at[:, 3].copy()).cuda().requires_grad_(True)
    amp = torch.ones(300, dtype=torch.float64, device="cuda", requires_grad=True)
    I = intensity(pos, ph, amp, Ndet=32, pixelsize=150.0, detz=1e5, k=K_UM, settings="single nf")
    loss = ((I.double() - torch.from_numpy(target).cuda()) ** 2).sum()
    loss.backward()
    field = oracle_mod.field_c(g, at)
    rpos, rph, ramp = oracle_mod.adjoint_c(g, at, 2 * (np.abs(field) ** 2 - target), field=field,
                                           amp=np.ones(300))
    assert rel_l2(pos.grad.cpu().numpy(), rpos) <= GRAD_RTOL
    assert rel_l2(ph.grad.cpu().numpy(), rph) <= GRAD_RTOL
    assert rel_l2(amp.grad.cpu().numpy(), ramp) <= GRAD_RTOL


@pytest.mark.parametrize("mode", ["nf", "ff_secondorder", "nf_scale"])
def test_adjoint_l2_fused_loss(lib, oracle_mod, mode):
    """idisim_adjoint_l2 (dL/dI formed inside the adjoint) against idisim_adjoint with dL/dI from
    torch, the oracle's FP64 loss and gradient; the loss is the same bit for bit on every call."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(93)
    g = Geometry(64, 48, 300.0, 1e5, K_UM, nf=mode.startswith("nf"), secondorder="secondorder" in mode,
                 scale="scale" in mode)
    at, target = _problem(rng, 600, 0.02, g, oracle_mod)
    amp = rng.uniform(0.5, 1, 600)
    h = make_handle(lib, g, lib.PREC_FP32SPLIT)
    d, da = torch.from_numpy(at).cuda(), torch.from_numpy(amp).cuda()
    T = torch.from_numpy(target).float().cuda()
    field = h.forward(d, da)
    loss, gpos, gph, gamp = h.adjoint_l2(d, field, T, da)
    loss2, gpos2, gph2, gamp2 = h.adjoint_l2(d, field, T, da)
    torch.cuda.synchronize()
    assert loss.item() == loss2.item() and torch.equal(gpos, gpos2) and torch.equal(gph, gph2)
    I = field.real.double() ** 2 + field.imag.double() ** 2
    assert loss.item() == pytest.approx(float(((I - T.double()) ** 2).sum()), rel=1e-12)
    spos, sph, samp = h.adjoint(d, field, (2 * (I - T.double())).float().contiguous(), da)
    for a_, b_ in ((gpos, spos), (gph, sph), (gamp, samp)):
        assert rel_l2(a_.cpu().numpy(), b_.cpu().numpy()) <= 1e-6
    fref = oracle_mod.field_c(g, at, amp)
    lref = np.sum((np.abs(fref) ** 2 - target.astype(np.float32).astype(np.float64)) ** 2)
    assert loss.item() == pytest.approx(lref, rel=1e-4)
    rpos, rph, ramp = oracle_mod.adjoint_c(g, at, 2 * (np.abs(fref) ** 2 - target.astype(np.float32)), field=fref,
                                           amp=amp)
    e = (rel_l2(gpos.cpu().numpy(), rpos), rel_l2(gph.cpu().numpy(), rph), rel_l2(gamp.cpu().numpy(), ramp))
    assert max(e) <= GRAD_RTOL, e


@pytest.mark.parametrize("case", ["nf_4096_256px", "nf_amp", "ff_secondorder", "nf_scale", "per_pixel_tile",
                                  "one_split"])
def test_loss_grad_l2_equals_forward_then_adjoint(lib, oracle_mod, case):
    """idisim_loss_grad_l2 against idisim_forward + idisim_adjoint_l2 on the same handle: the loss,
    the gradients and the field bit for bit, both where the adjoint's pixel pass sums the forward's
    split partials itself (mirror-pair plans with splits) and where the forward finalizes on its
    own (a per-pixel tile, or one split)."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(95)
    if case == "nf_4096_256px":  # the adjoint workload's shape
        g, n, r = Geometry(256, 256, 75.0, 1e5, K_UM, nf=True), 4096, 0.02
    else:
        g = Geometry(64, 48, 300.0, 1e5, K_UM, nf=not case.startswith("ff"), secondorder="secondorder" in case,
                     scale="scale" in case)
        n, r = 600, 0.02
    at, target = _problem(rng, n, r, g, oracle_mod)
    amp = rng.uniform(0.5, 1, n) if case in ("nf_amp", "nf_scale") else None
    h = make_handle(lib, g, lib.PREC_FP32SPLIT)
    if case == "per_pixel_tile":
        h.set_option(lib.OPT_TILE, 2)
    if case == "one_split":
        h.set_option(lib.OPT_SPLIT, 1)
    d = torch.from_numpy(at).cuda()
    da = None if amp is None else torch.from_numpy(amp).cuda()
    T = torch.from_numpy(target).float().cuda()
    ext = float(np.linalg.norm(at[:, :3], axis=1).max())
    field = h.forward(d, da, None, lib.OUT_FIELD, ext)
    ref = h.adjoint_l2(d, field, T, da, ext, want_amp=amp is not None)
    plan_adj = h.plan()
    f2 = torch.empty_like(field)
    for out in (f2, None):  # with and without the field output
        got = h.loss_grad_l2(d, T, da, ext, field=out, want_amp=amp is not None)
        torch.cuda.synchronize()
        assert got[0].item() == ref[0].item(), (case, got[0].item(), ref[0].item())
        for a_, b_ in zip(got[1:], ref[1:]):
            if b_ is not None:
                assert torch.equal(a_, b_), case
    assert torch.equal(f2, field), case
    assert h.plan()["split"] == plan_adj["split"]  # the adjoint's plan, as after adjoint_l2


@pytest.mark.parametrize("mode", ["nf", "ff_secondorder", "nf_scale"])
def test_adjoint_ragged_detector(lib, oracle_mod, mode):
    """Detectors that are no multiple of 16 or 32 pixels (odd sizes, so the centre falls between
    pixels): the last row and column of 16 x 16 adjoint tiles and of 32 x 32 forward tiles are
    partly outside the detector.  The gradients against the FP64 analytic gradient, the loss against
    the oracle's, and the one-call loss and gradient against forward + idisim_adjoint_l2 bit for bit.
    On the 300 um detector the planner's forward is a per-pixel tile (its own finalize); the 75 um one
    pins the centred mirror-pair tile 11 with 4 splits (within the 5e-6 rev bound there), so the
    adjoint's pixel pass sums the forward's split partials over the ragged tiles itself."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(97)
    for n0, n1, pix, n in ((50, 37, 300.0, 500), (100, 77, 75.0, 1024)):
        g = Geometry(n0, n1, pix, 1e5, K_UM, nf=mode.startswith("nf"), secondorder="secondorder" in mode,
                     scale="scale" in mode)
        at, target = _problem(rng, n, 0.02, g, oracle_mod)
        amp = rng.uniform(0.5, 1, n) if "scale" in mode else None
        fref = oracle_mod.field_c(g, at, amp)
        t32 = target.astype(np.float32).astype(np.float64)
        rpos, rph, ramp = oracle_mod.adjoint_c(g, at, 2 * (np.abs(fref) ** 2 - t32), field=fref, amp=amp)
        lref = np.sum((np.abs(fref) ** 2 - t32) ** 2)

        h = make_handle(lib, g, lib.PREC_FP32SPLIT)
        pinned = pix == 75.0
        if pinned:
            h.set_option(lib.OPT_TILE, 11)
            h.set_option(lib.OPT_SPLIT, 4)
        d = torch.from_numpy(at).cuda()
        da = None if amp is None else torch.from_numpy(amp).cuda()
        T = torch.from_numpy(target).float().cuda()
        ext = float(np.linalg.norm(at[:, :3], axis=1).max())
        field = h.forward(d, da, None, lib.OUT_FIELD, ext)
        fplan = h.plan()  # the plan that the forward inside loss_grad_l2 runs too
        print(mode, n0, n1, "forward plan", fplan)
        if pinned:
            assert (fplan["tile"], fplan["split"]) == (11, 4) and fplan["phase_bound"] <= 5e-6, fplan
        ref = h.adjoint_l2(d, field, T, da, ext, want_amp=amp is not None)
        f2 = torch.empty((n0, n1), dtype=torch.complex64, device="cuda")
        got = h.loss_grad_l2(d, T, da, ext, field=f2, want_amp=amp is not None)
        torch.cuda.synchronize()
        assert torch.equal(f2, field), (mode, n0, n1)
        assert got[0].item() == ref[0].item(), (mode, n0, n1)
        for a_, b_ in zip(got[1:], ref[1:]):
            if b_ is not None:
                assert torch.equal(a_, b_), (mode, n0, n1)
        assert got[0].item() == pytest.approx(lref, rel=1e-4), (mode, n0, n1)
        e = [rel_l2(got[1].cpu().numpy().astype(np.float64), rpos),
             rel_l2(got[2].cpu().numpy().astype(np.float64), rph)]
        if amp is not None:
            e.append(rel_l2(got[3].cpu().numpy().astype(np.float64), ramp))
        print(mode, n0, n1, "relL2 vs FP64 analytic", e)
        assert max(e) <= GRAD_RTOL, (mode, n0, n1, e)


def test_loss_grad_l2_fp64_fallback_geometry(lib, oracle_mod):
    """Pixels of 2 mm at 0.1 m: the forward falls back to the FP64 kernel (complex64 field through
    its finalize) and the adjoint, which has no FP64 form, refuses the geometry; the one call does
    what the two calls do."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(96)
    g = Geometry(24, 20, 2000.0, 1e5, K_UM, nf=True)
    at, target = _problem(rng, 300, 0.03, g, oracle_mod)
    h = make_handle(lib, g, lib.PREC_FP32SPLIT)
    d = torch.from_numpy(at).cuda()
    T = torch.from_numpy(target).float().cuda()
    field = h.forward(d)
    assert h.plan()["fallback"] == 1 and field.dtype == torch.complex64
    with pytest.raises(lib.IdisimError) as e2:
        h.adjoint_l2(d, field, T)
    f2 = torch.zeros_like(field)
    with pytest.raises(lib.IdisimError) as e1:
        h.loss_grad_l2(d, T, field=f2)
    assert e1.value.code == e2.value.code == lib.E_PRECISION
    torch.cuda.synchronize()
    assert torch.equal(f2, field)  # the forward ran, into the caller's complex64 buffer


def test_autograd_l2_loss(lib, oracle_mod):
    import torch

    from idi_amd.simulation.autograd import intensity, intensity_l2_loss
    from oracle import Geometry

    rng = np.random.default_rng(94)
    g = Geometry(32, 32, 150.0, 1e5, K_UM, nf=True)
    at, target = _problem(rng, 300, 0.02, g, oracle_mod)
    T = torch.from_numpy(target).float().cuda()
    kw = dict(Ndet=32, pixelsize=150.0, detz=1e5, k=K_UM, settings="single nf")
    grads = []
    for fused in (True, False):
        pos = torch.from_numpy(at[:, :3].copy()).cuda().requires_grad_(True)
        ph = torch.from_numpy(at[:, 3].copy()).cuda().requires_grad_(True)
        amp = torch.ones(300, dtype=torch.float64, device="cuda", requires_grad=True)
        if fused:
            loss = intensity_l2_loss(pos, ph, amp, target=T, **kw)
        else:
            I = intensity(pos, ph, amp, **kw)
            loss = ((I.double() - T.double()) ** 2).sum()
        (3.0 * loss).backward()  # the upstream gradient scales the stored ones
        grads.append((loss.item(), pos.grad.cpu().numpy(), ph.grad.cpu().numpy(), amp.grad.cpu().numpy()))
    (l1, *g1), (l2, *g2) = grads
    assert l1 == pytest.approx(l2, rel=1e-6)
    for a_, b_ in zip(g1, g2):
        assert rel_l2(a_, b_) <= 1e-5
