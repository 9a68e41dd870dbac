"""Coherent emitters (every phase zero) through the HIP kernels.

With phi_j = 0 the terms of a lattice add in phase at the Bragg peaks and at q = 0, so
the per-term phase errors of the FP32-split tiles (which are systematic: smooth in the
pixel and linear in r_j) and the rounding of the chunked FP32 sums add up coherently
instead of at random.  This is the regime of the reference CLI's ``--coherent``
(rndPhase = False; scripts/idi_sim.py:18,62-66) and of every crystal.

Bars (relative L2 of the intensity over every pixel, peaks included):
  * far field, against the Laue closed form of a box lattice (exact for the far-field
    formula, cpu.py:57-66): <= 1e-4 for every tile variant whose own a-priori bound
    meets the planner's tolerance, i.e. every tile the planner can pick, and for the
    planner's own choice;
  * near field, against FP64 oracle rows through the peaks: the same.
Geometries: the 512^2 / 75 um / 0.2 m detector and the 1024^2 / 50 um / 0.5 m detector
of the benchmark workloads, at k = 2 pi / 0.154 nm.
"""
import numpy as np
import pytest

from conftest import K_UM
from gpu_util import ball, fp64_tol, rel_l2

pytestmark = pytest.mark.gpu

COHERENT_RTOL = 1e-4
TOL_REV = 5e-6  # the planner's default per-term phase tolerance (IDISIM_OPT_TOL_PPB 5000)
TILES = [t for t in range(16) if t != 4]  # 4 is a reserved index without a kernel


@pytest.fixture(scope="module")
def lib():
    import torch

    from idi_amd import _lib

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    _lib.load()
    return _lib


def box(n, a):
    """Centred n[0] x n[1] x n[2] lattice of spacing a (um), phases zero -> (N, 4)."""
    g = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3) * a
    g = g - g.mean(0)
    return np.concatenate([g, np.zeros((len(g), 1))], 1)


def laue(geom, n, a):
    """prod_i sin^2(n_i q_i a / 2) / sin^2(q_i a / 2) with q = k (r_p / |r_p| - z_hat), FP64."""
    detx, dety, detz = geom.pixels()
    R = np.sqrt(detx**2 + dety**2 + detz**2)
    q = (geom.k * detx / R, geom.k * dety / R, -geom.k * (detx**2 + dety**2) / (R * (R + detz)))
    out = np.ones(detx.shape)
    for ni, qi in zip(n, q):
        s = np.sin(qi * a / 2)
        with np.errstate(invalid="ignore", divide="ignore"):
            v = np.sin(ni * qi * a / 2) ** 2 / s**2
        out *= np.where(np.abs(s) < 1e-12, float(ni * ni), v)
    return out


def forced(lib, h, d, tile, amp=None, kind=None):
    """Output of `h` with tile variant `tile` forced (planner: -1), and its plan; None if the
    variant refuses this phase model."""
    import torch

    kind = lib.OUT_INTENSITY if kind is None else kind
    try:
        h.set_option(lib.OPT_TILE, tile)
        out = h.forward(d, amp, None, kind)
    except lib.IdisimError:
        return None, None
    torch.cuda.synchronize()
    return out.cpu().numpy(), h.plan()


# (detector n, pixel um, z um) of the two benchmark detectors, and two box lattices each:
# Bragg peaks on the detector (spacing 2.1 / 3.9 nm, q_B = 2.99 / 1.61 nm^-1), and a denser
# lattice whose Bragg peaks lie off the detector but whose q = 0 peak holds N^2.
LAUE_CASES = {
    "512px_bragg": ((512, 75.0, 2e5), (10, 10, 10), 2.1e-3),
    "512px_dense": ((512, 75.0, 2e5), (31, 31, 31), 0.45e-3),
    "1024px_bragg": ((1024, 50.0, 5e5), (10, 10, 10), 3.9e-3),
    "1024px_dense": ((1024, 50.0, 5e5), (41, 41, 41), 1.2e-3),
}


@pytest.mark.parametrize("case", sorted(LAUE_CASES))
def test_laue_far_field_every_eligible_tile(lib, case):
    import torch

    from oracle import Geometry

    (n0, pix, z), n, a = LAUE_CASES[case]
    g = Geometry(n0, n0, pix, z, K_UM)
    at = box(n, a)
    expect = laue(g, n, a)
    assert expect.max() == pytest.approx(len(at) ** 2, rel=1e-12)  # the q = 0 pixel
    h = lib.Handle(n0, n0, pix, z, K_UM, precision=lib.PREC_FP32SPLIT)
    d = torch.from_numpy(at).cuda()
    checked = {}
    for tile in [-1] + TILES:
        I, plan = forced(lib, h, d, tile)
        if I is None:
            continue
        assert plan["fallback"] == 0 and plan["extent_exceeded"] == 0
        err = rel_l2(I, expect)
        if plan["phase_bound"] <= TOL_REV:
            assert err <= COHERENT_RTOL, (case, tile, err, plan)
            checked[tile] = err
        else:
            # a tile the planner would not pick still stays within its own bound:
            # |dE| <= 2 pi bound N, |dI| <= 2 |E| |dE| + |dE|^2 at every pixel
            dE = 2 * np.pi * plan["phase_bound"] * len(at)
            assert np.all(np.abs(I - expect) <= 2 * np.sqrt(expect) * dE + dE**2 + 1e-6 * expect.max()), (case, tile)
    print(case, "N", len(at), {t: f"{e:.2e}" for t, e in checked.items()})
    assert -1 in checked and len(checked) >= 6, checked


def _sc_sphere(a, r):
    m = int(np.ceil(r / a))
    gr = np.arange(-m, m + 1) * a
    pts = np.stack(np.meshgrid(gr, gr, gr, indexing="ij"), -1).reshape(-1, 3)
    return pts[np.linalg.norm(pts, axis=1) <= r * (1 + 1e-12)]


def _three_balls(rng, n):
    radii = np.array([0.015, 0.020, 0.025])
    centres = np.array([[-0.040, 0, 0], [0.0, 0, 0], [0.050, 0, 0]])
    counts = np.rint(n * radii**3 / np.sum(radii**3)).astype(int)
    counts[-1] = n - counts[:-1].sum()
    return np.concatenate([c + ball(rng, k, r) for c, k, r in zip(centres, counts, radii)])


def test_coherent_near_field_rows_every_eligible_tile(lib, oracle_mod):
    """Near field, phases zero, FP64 oracle rows through the q = 0 peak and the Bragg rows:
    100000 of the simple-cubic sites (0.3 nm) in a 10 nm ball on the 512^2 detector at 0.2 m, a
    10 x 10 x 10 lattice of 2.1 nm there (Bragg rows 61 and 451), 250000 emitters in the three
    balls of the large workload with amplitudes U[0.5, 1] on the 1024^2 detector at 0.5 m, and
    20000 simple-cubic sites (0.6 nm) in a 30 nm ball on the batched shots' 256^2 detector of
    100 um pixels at 0.1 m (where the centred tiles 7, 8, 12, 14 and 15 qualify)."""
    import torch

    from oracle import Geometry

    rng = np.random.default_rng(7100)
    sites = _sc_sphere(3e-4, 0.01)
    sc = sites[rng.choice(len(sites), 100_000, replace=False)]
    cases = [
        ("sc100k_512px", Geometry(512, 512, 75.0, 2e5, K_UM, nf=True), np.c_[sc, np.zeros(len(sc))], None,
         [0, 200, 255, 256, 257, 400]),
        ("bragg_512px", Geometry(512, 512, 75.0, 2e5, K_UM, nf=True), box((10, 10, 10), 2.1e-3), None,
         [60, 61, 62, 256, 450, 451, 452]),
        ("3balls_1024px", Geometry(1024, 1024, 50.0, 5e5, K_UM, nf=True),
         np.c_[_three_balls(rng, 250_000), np.zeros(250_000)], rng.uniform(0.5, 1.0, 250_000), [511, 512, 800]),
    ]
    sites30 = _sc_sphere(6e-4, 0.03)
    sc30 = sites30[rng.choice(len(sites30), 20_000, replace=False)]
    cases.append(("sc20k_30nm_256px", Geometry(256, 256, 100.0, 1e5, K_UM, nf=True), np.c_[sc30, np.zeros(len(sc30))],
                  None, [0, 127, 128, 200]))
    for name, g, at, amp, rows in cases:
        ref = np.abs(oracle_mod.field_c(g, at, amp, rows=rows)) ** 2
        h = lib.Handle(g.n0, g.n1, g.pixelsize, g.detz, K_UM, nf=True, precision=lib.PREC_FP32SPLIT)
        d = torch.from_numpy(np.ascontiguousarray(at)).cuda()
        da = None if amp is None else torch.from_numpy(amp).cuda()
        checked = {}
        for tile in [-1] + TILES:
            I, plan = forced(lib, h, d, tile, da)
            if I is None or plan["phase_bound"] > TOL_REV:
                continue
            assert plan["fallback"] == 0 and plan["extent_exceeded"] == 0
            err = rel_l2(I[rows], ref)
            assert err <= COHERENT_RTOL, (name, tile, err, plan)
            checked[tile] = err
        # the q = 0 pixel of the centred detectors holds |sum a_j|^2 up to the nf curvature
        print(name, "peak/mean", ref.max() / ref.mean(), {t: f"{e:.2e}" for t, e in checked.items()})
        assert -1 in checked and len(checked) >= 3, (name, checked)


def test_coherent_crystal_through_simulate(lib, oracle_mod):
    """simobj.sc with rndPhase = False (the CLI's --coherent) through the drop-in entry points:
    hip.simulate with "single nf" and "double nf", and cpu.simulate with the reference CPU
    backend's defaults (far field, FP64), against the oracle on rows through the Bragg peaks."""
    from oracle import Geometry

    from idi_amd.simulation import cpu, hip, simobj

    E = 1.24 / 1.54e-4  # eV, so that simobj's k = 2 pi E / 1.24 is 2 pi / 0.154 nm
    obj = simobj.sc(E, 1000, 2.1e-3)  # 10 x 10 x 10 sites, all chosen
    obj.rndPhase = False
    obj.rndPos = False
    at = obj.get()
    assert np.all(at[:, 3] == 0) and len(at) == 1000
    rows = [61, 256, 451]
    for settings in ("single nf", "double nf"):
        g = Geometry(512, 512, 75.0, 2e5, obj.k, nf=True)
        tol = COHERENT_RTOL if "single" in settings else fp64_tol(g)
        ref = np.abs(oracle_mod.field_c(g, at, rows=rows)) ** 2
        f = hip.simulate(2, obj, 512, 75.0, 2e5, settings=settings, verbose=False)
        assert f.shape == (2, 512, 512) and f.dtype == np.complex128
        for i in range(2):
            e = rel_l2(np.abs(f[i][rows]) ** 2, ref)
            assert e <= tol, (settings, i, e)
    g = Geometry(512, 512, 75.0, 2e5, obj.k)
    ref = np.abs(oracle_mod.field_c(g, at, rows=rows)) ** 2
    f = cpu.simulate(2, obj, 512, 75.0, 2e5, obj.k)
    assert f.shape == (2, 512, 512) and f.dtype == np.complex128
    for i in range(2):
        assert rel_l2(np.abs(f[i][rows]) ** 2, ref) <= fp64_tol(g)
