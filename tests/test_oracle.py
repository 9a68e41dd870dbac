"""Pins the FP64 oracle (oracle/) before it is trusted as the checker of the HIP library.

Each check is chosen so that a plausible mistake fails it:
  golden     50-digit mpmath evaluations (tests/golden/make_golden.py) at real
             geometries: near/far field, scale, secondorder, rotation, offset,
             odd detector, amplitudes;
  K1         single emitter: |E|^2 = a^2, a^2/D^2 (nf scale), a^2/|r_p|^2 (ff scale);
  K2         two emitters: a1^2 + a2^2 + 2 a1 a2 cos(theta1 - theta2), theta by mpmath;
  K3         emitters at the origin: |sum a e^{i phi}|^2 on every pixel;
  K4         Laue closed form of a coherent box lattice (far field);
  K5         invariants: global phase, far-field translation, Friedel symmetry,
             near field -> far field as z grows;
  K6         mean intensity over random phases = sum a^2;
  K8         NumPy restatement == C restatement on random tiny cases;
  K9         analytic FP64 gradient vs central differences; single-emitter gradient;
  mutations  plausible implementation mistakes, planted in the oracle's inputs
             on the CPU, are rejected by the golden vectors.
"""
import mpmath as mp
import numpy as np
import pytest

from conftest import K_UM, fp64_phase_tolerance, golden_field, golden_geometry, golden_term_scale


def _ball(rng, n, r):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v * r * rng.uniform(0, 1, (n, 1)) ** (1 / 3)


def _atoms(rng, n, r, phases=True):
    ph = rng.uniform(0, 2 * np.pi, (n, 1)) if phases else np.zeros((n, 1))
    return np.concatenate([_ball(rng, n, r), ph], 1)


# ----------------------------------------------------------------- golden
@pytest.mark.parametrize("impl", ["c", "numpy"])
def test_golden(oracle_mod, golden, impl):
    for case in golden:
        g = golden_geometry(case)
        atoms = np.array(case["atoms"])
        amp = None if case["amp"] is None else np.array(case["amp"])
        f = oracle_mod.field_c(g, atoms, amp) if impl == "c" else oracle_mod.field_numpy(g, atoms, amp)
        ref = golden_field(case)
        tol = fp64_phase_tolerance(case) * golden_term_scale(case)
        err = np.abs(f - ref).max()
        assert err <= tol, (case["name"], err, tol)


def _golden_err(oracle_mod, case, atoms=None, geom=None):
    g = golden_geometry(case) if geom is None else geom
    atoms = np.array(case["atoms"]) if atoms is None else atoms
    amp = None if case["amp"] is None else np.array(case["amp"])
    f = oracle_mod.field_numpy(g, atoms, amp)
    return np.abs(f - golden_field(case)).max() / (fp64_phase_tolerance(case) * golden_term_scale(case))


def test_mutations_are_caught(oracle_mod, golden):
    """Plausible mistakes, planted on the CPU, exceed the golden tolerance by orders of magnitude."""
    import oracle

    by = {c["name"]: c for c in golden}
    nf = by["nf_ball50nm_128px_z0.1m_wide"]
    ff = by["ff_ball50nm_128px_z0.1m_wide"]
    odd = by["nf_odd_7x5"]
    rot = by["nf_rotated_offset"]
    a = np.array(nf["atoms"])
    # dropped incident-wave term +k z_j (theta - k z_j)
    m = a.copy()
    m[:, 3] -= nf["k"] * m[:, 2]
    assert _golden_err(oracle_mod, nf, m) > 100
    # sign of the far-field q . r
    m = np.array(ff["atoms"]).copy()
    m[:, :3] *= -1
    assert _golden_err(oracle_mod, ff, m) > 100
    # x and y swapped in the pixel coordinates
    m = np.array(nf["atoms"]).copy()
    m[:, [0, 1]] = m[:, [1, 0]]
    assert _golden_err(oracle_mod, nf, m) > 100
    # integer division for the detector centre of an odd detector (sim.cu:108)
    g = golden_geometry(odd)
    g.detoffset = (g.detoffset[0] + 0.5 * g.pixelsize, g.detoffset[1] + 0.5 * g.pixelsize)
    assert _golden_err(oracle_mod, odd, geom=g) > 100
    # transposed rotation matrix
    g = golden_geometry(rot)
    g.M = g.M.T.copy()
    assert _golden_err(oracle_mod, rot, geom=g) > 100
    # k rounded to float32 (cuda.py:102)
    g = golden_geometry(nf)
    g.k = float(np.float32(g.k))
    assert _golden_err(oracle_mod, nf, geom=g) > 100
    assert oracle.relative_l2(np.ones(3), np.ones(3)) == 0


# ----------------------------------------------------------------- K1
@pytest.mark.parametrize("nf", [True, False])
@pytest.mark.parametrize("scale", [False, True])
def test_single_emitter(oracle_mod, nf, scale):
    rng = np.random.default_rng(1)
    g = oracle_mod.Geometry(9, 7, 75.0 * 128 / 9, 1e5, K_UM, nf=nf, scale=scale)
    at = _atoms(rng, 1, 0.05)
    a = np.array([0.7])
    intensity = np.abs(oracle_mod.field_c(g, at, a)) ** 2
    detx, dety, detz = g.pixels()
    if not scale:
        expect = np.full(intensity.shape, 0.49)
    elif nf:
        D2 = (detx - at[0, 0]) ** 2 + (dety - at[0, 1]) ** 2 + (detz - at[0, 2]) ** 2
        expect = 0.49 / D2
    else:
        expect = 0.49 / (detx**2 + dety**2 + detz**2)
    np.testing.assert_allclose(intensity, expect, rtol=1e-12)


# ----------------------------------------------------------------- K2
def test_two_emitters_mpmath(oracle_mod):
    mp.mp.dps = 40
    rng = np.random.default_rng(2)
    g = oracle_mod.Geometry(4, 3, 2400.0, 1e5, K_UM, nf=True)
    at = _atoms(rng, 2, 0.05)
    a = np.array([1.0, 0.6])
    intensity = np.abs(oracle_mod.field_c(g, at, a)) ** 2
    detx, dety, detz = g.pixels()
    for x in range(g.n0):
        for y in range(g.n1):
            th = []
            for j in range(2):
                D = mp.sqrt((mp.mpf(detx[x, y]) - at[j, 0]) ** 2 + (mp.mpf(dety[x, y]) - at[j, 1]) ** 2
                            + (mp.mpf(detz[x, y]) - at[j, 2]) ** 2)
                th.append((D - detz[x, y] + at[j, 2]) * mp.mpf(g.k) + at[j, 3])
            expect = 1.0 + 0.36 + 2 * 0.6 * float(mp.cos(th[0] - th[1]))
            tol = 2 * 2 * 0.6 * fp64_phase_tolerance({"detz": 1e5, "n0": 4, "n1": 3, "pixelsize": 2400.0,
                                                      "detoffset": (0, 0), "k": K_UM})
            assert abs(intensity[x, y] - expect) <= tol + 1e-12


# ----------------------------------------------------------------- K3
@pytest.mark.parametrize("nf", [True, False])
def test_emitters_at_origin(oracle_mod, nf):
    rng = np.random.default_rng(3)
    n = 50
    at = np.zeros((n, 4))
    at[:, 3] = rng.uniform(0, 2 * np.pi, n)
    a = rng.uniform(0.5, 1, n)
    g = oracle_mod.Geometry(8, 6, 2000.0, 1e5, K_UM, nf=nf, detangles=(0.2, -0.1), detoffset=(300.0, 100.0))
    intensity = np.abs(oracle_mod.field_c(g, at, a)) ** 2
    E = np.sum(a * np.exp(1j * at[:, 3]))
    expect = abs(E) ** 2
    # literal FP64 formula: per-term phase error <= a few ulps of k |r_p| (near field)
    dE = a.sum() * (fp64_phase_tolerance({"detz": 1e5, "n0": 8, "n1": 6, "pixelsize": 2000.0,
                                          "detoffset": (300.0, 100.0), "k": K_UM}) if nf else 1e-13)
    assert np.abs(intensity - expect).max() <= 2 * abs(E) * dE + dE**2


# ----------------------------------------------------------------- K4
def test_laue_box_lattice(oracle_mod):
    n1, n2, n3, a = 4, 3, 5, 0.0003 * 7  # 2.1 nm spacing
    grid = np.stack(np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij"), -1).reshape(-1, 3) * a
    grid -= grid.mean(0)
    at = np.concatenate([grid, np.zeros((len(grid), 1))], 1)
    g = oracle_mod.Geometry(16, 12, 1500.0, 1e5, K_UM)
    intensity = np.abs(oracle_mod.field_c(g, at)) ** 2
    detx, dety, detz = g.pixels()
    R = np.sqrt(detx**2 + dety**2 + detz**2)
    q = np.stack([K_UM * detx / R, K_UM * dety / R, K_UM * (detz / R - 1)])

    def laue(n, qa):
        s = np.sin(qa / 2)
        with np.errstate(invalid="ignore", divide="ignore"):
            v = np.sin(n * qa / 2) ** 2 / s**2
        return np.where(np.abs(s) < 1e-12, n * n, v)

    expect = laue(n1, q[0] * a) * laue(n2, q[1] * a) * laue(n3, q[2] * a)
    np.testing.assert_allclose(intensity, expect, rtol=1e-8, atol=1e-8 * expect.max())


# ----------------------------------------------------------------- K5
def test_invariants(oracle_mod):
    rng = np.random.default_rng(5)
    at = _atoms(rng, 40, 0.05)
    gff = oracle_mod.Geometry(10, 8, 1200.0, 1e5, K_UM)
    gnf = oracle_mod.Geometry(10, 8, 1200.0, 1e5, K_UM, nf=True)
    base_ff = np.abs(oracle_mod.field_c(gff, at)) ** 2
    base_nf = np.abs(oracle_mod.field_c(gnf, at)) ** 2
    # global phase shift
    sh = at.copy()
    sh[:, 3] += 1.234
    np.testing.assert_allclose(np.abs(oracle_mod.field_c(gff, sh)) ** 2, base_ff, rtol=1e-9)
    np.testing.assert_allclose(np.abs(oracle_mod.field_c(gnf, sh)) ** 2, base_nf, rtol=1e-5)
    # far field: translation of the object
    tr = at.copy()
    tr[:, :3] += np.array([0.02, -0.01, 0.015])
    np.testing.assert_allclose(np.abs(oracle_mod.field_c(gff, tr)) ** 2, base_ff, rtol=1e-8)
    # permutation
    np.testing.assert_allclose(np.abs(oracle_mod.field_c(gff, at[::-1].copy())) ** 2, base_ff, rtol=1e-9)
    # Friedel symmetry of a planar object in the far field: I(x, y) = I(N0-x, N1-y)
    pl = at.copy()
    pl[:, 2] = 0
    pl[:, 3] = 0
    I = np.abs(oracle_mod.field_c(gff, pl)) ** 2
    np.testing.assert_allclose(I[1:, 1:], I[1:, 1:][::-1, ::-1], rtol=1e-9)
    # near field tends to far field as the detector recedes at fixed angles
    errs = []
    for z in (1e4, 1e5, 1e6):
        p = 1200.0 * z / 1e5
        I1 = np.abs(oracle_mod.field_c(oracle_mod.Geometry(10, 8, p, z, K_UM, nf=True), at)) ** 2
        I2 = np.abs(oracle_mod.field_c(oracle_mod.Geometry(10, 8, p, z, K_UM), at)) ** 2
        errs.append(oracle_mod.relative_l2(I1, I2))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 40 * K_UM * 0.05**2 / (2 * 1e6)


# ----------------------------------------------------------------- K6
def test_mean_over_random_phases(oracle_mod):
    rng = np.random.default_rng(6)
    pos = _ball(rng, 30, 0.05)
    a = rng.uniform(0.5, 1, 30)
    g = oracle_mod.Geometry(6, 6, 2000.0, 1e5, K_UM, nf=True)
    acc = np.zeros((6, 6))
    shots = 400
    for _ in range(shots):
        at = np.concatenate([pos, rng.uniform(0, 2 * np.pi, (30, 1))], 1)
        acc += np.abs(oracle_mod.field_c(g, at, a)) ** 2
    mean = acc / shots
    # per pixel the mean has relative std ~1/sqrt(shots); over 36 pixels
    assert abs(mean.mean() / np.sum(a * a) - 1) < 0.03


# ----------------------------------------------------------------- K8
def test_numpy_equals_c(oracle_mod):
    rng = np.random.default_rng(8)
    for trial in range(12):
        nf, scale, so = bool(trial & 1), bool(trial & 2), bool(trial & 4)
        ang = (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3)) if trial % 3 == 0 else (0.0, 0.0)
        g = oracle_mod.Geometry(int(rng.integers(3, 9)), int(rng.integers(3, 9)), rng.uniform(50, 3000), 1e5, K_UM,
                                detoffset=tuple(rng.uniform(-500, 500, 2)), detangles=ang, nf=nf, scale=scale,
                                secondorder=so)
        at = _atoms(rng, 25, 0.05)
        a = rng.uniform(0.5, 1, 25) if trial % 2 else None
        f1 = oracle_mod.field_numpy(g, at, a)
        f2 = oracle_mod.field_c(g, at, a)
        scale_ = np.abs(f1).max() + 1e-300
        assert np.abs(f1 - f2).max() <= 1e-12 * 25 * (1 if not scale else scale_)
        rows = [g.n0 - 1, 0]
        np.testing.assert_array_equal(oracle_mod.field_c(g, at, a, rows=rows), f2[rows])


# ----------------------------------------------------------------- K9
@pytest.mark.parametrize("mode", ["nf", "ff", "ff_secondorder", "nf_scale"])
def test_adjoint_vs_central_differences(oracle_mod, mode):
    rng = np.random.default_rng(9)
    g = oracle_mod.Geometry(12, 10, 1500.0, 1e5, K_UM, nf="nf" in mode, scale="scale" in mode,
                            secondorder="secondorder" in mode)
    at = _atoms(rng, 6, 0.02)
    a = rng.uniform(0.5, 1, 6)
    target = np.abs(oracle_mod.field_c(g, _atoms(rng, 6, 0.02), a)) ** 2

    # The loss for the differences is evaluated in 80-bit long double: in FP64 the
    # rounding of D ~ 1e5 um (ulp 1.5e-11 um) changes at every step of h = 1e-4 nm
    # and would dominate the difference quotient.
    detx, dety, detz = (v.astype(np.longdouble) for v in g.pixels())
    kk = np.longdouble(g.k)
    R = np.sqrt(detx**2 + dety**2 + detz**2)

    def field_ld(atoms, amp):
        E = np.zeros(detx.shape, dtype=np.clongdouble)
        for (x, y, z, ph), aj in zip(atoms.astype(np.longdouble), amp.astype(np.longdouble)):
            if g.nf:
                D = np.sqrt((detx - x) ** 2 + (dety - y) ** 2 + (detz - z) ** 2)
                th = (D - detz + z) * kk + ph
                m = aj / D if g.scale else aj
            else:
                th = -(x * kk * detx / R + y * kk * dety / R + z * kk * (detz / R - 1)) + ph
                if g.secondorder:
                    th = th + kk / (2 * R) * (x * x + y * y)
                m = aj / R if g.scale else aj
            E += m * (np.cos(th) + 1j * np.sin(th))
        return E

    def loss(atoms, amp):
        I = np.abs(field_ld(atoms, amp)) ** 2
        return np.sum((I - target) ** 2)

    field = oracle_mod.field_c(g, at, a)
    ref = field_ld(at, a)
    assert np.abs(field - ref).max() <= 6 * fp64_phase_tolerance(
        {"detz": 1e5, "n0": 12, "n1": 10, "pixelsize": 1500.0, "detoffset": (0, 0), "k": K_UM}) * np.abs(ref).max()
    dLdI = 2 * (np.abs(field) ** 2 - target)
    gpos, gph, gamp = oracle_mod.adjoint_c(g, at, dLdI, field=field, amp=a)
    # h = 1e-4 nm (k h |grad| ~ 4e-3 rad): truncation ~1e-6 relative
    coords = [(j, c) for j in range(6) for c in range(4)]
    floor = np.abs(gpos).max() * 1e-2 + 1e-300
    for h, rtol in ((1e-7, 2e-5), (1e-6, 1e-4)):
        steps = [h if c < 3 else 1e-4 for _, c in coords]
        # the C long-double differences (the base field once, the moved term's change in closed form)
        cfd = oracle_mod.loss_central_diff(g, at, target, coords, steps, amp=a)
        for i, ((j, c), hh) in enumerate(zip(coords, steps)):
            p, m = at.copy(), at.copy()
            p[j, c] += hh
            m[j, c] -= hh
            fd = (loss(p, a) - loss(m, a)) / (2 * hh)
            an = gpos[j, c] if c < 3 else gph[j]
            assert abs(fd - an) <= rtol * (abs(an) + floor), (h, j, c, fd, an)
            assert abs(cfd[i] - an) <= rtol * (abs(an) + floor), (h, j, c, cfd[i], an)
            if h == 1e-6:  # where the direct NumPy quotient's own rounding noise is ~1e-6 of it
                assert abs(cfd[i] - fd) <= 1e-5 * (abs(fd) + floor), (h, j, c, cfd[i], fd)
    # the C quotient converges as h^2 (no rounding noise down to h = 1e-5 nm)
    pos = [(j, c) for j, c in coords if c < 3]
    d7, d8 = (oracle_mod.loss_central_diff(g, at, target, pos, h, amp=a) for h in (1e-7, 1e-8))
    d6 = oracle_mod.loss_central_diff(g, at, target, pos, 1e-6, amp=a)
    assert np.all(np.abs(d7 - d8) <= 2e-2 * np.abs(d6 - d8) + 1e-9 * floor)
    for j in range(6):
        ap, am = a.copy(), a.copy()
        ap[j] += 1e-6
        am[j] -= 1e-6
        fd = (loss(at, ap) - loss(at, am)) / 2e-6
        assert abs(fd - gamp[j]) <= 1e-5 * (abs(gamp[j]) + np.abs(gamp).max()), (j, fd, gamp[j])


def test_adjoint_single_emitter_closed_form(oracle_mod):
    """One emitter, nf: dL/dphi = 0 and dL/dr = 0 for L = sum I (I = a^2 everywhere);
    for L = sum_p c_p I_p with a second fixed emitter, the phase gradient is the closed form."""
    g = oracle_mod.Geometry(5, 4, 3000.0, 1e5, K_UM, nf=True)
    at = np.array([[0.01, -0.02, 0.005, 0.3], [0.0, 0.0, 0.0, 0.0]])
    field = oracle_mod.field_c(g, at)
    w = np.linspace(0.5, 1.5, 20).reshape(5, 4)
    gpos, gph, _ = oracle_mod.adjoint_c(g, at, w, field=field)
    # I = 2 + 2 cos(theta_0 - theta_1): dI/dphi_0 = -2 sin(theta_0 - theta_1)
    detx, dety, detz = g.pixels()
    th = []
    for j in range(2):
        D = np.sqrt((detx - at[j, 0]) ** 2 + (dety - at[j, 1]) ** 2 + (detz - at[j, 2]) ** 2)
        th.append((D - detz + at[j, 2]) * K_UM + at[j, 3])
    expect = np.sum(w * (-2 * np.sin(th[0] - th[1])))
    assert abs(gph[0] - expect) <= 1e-6 * np.sum(np.abs(w))
    assert abs(gph[1] + expect) <= 1e-6 * np.sum(np.abs(w))
