"""The host side of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU).

tests/host_asan/planner_check.cpp compiles the FP32-split error planner (idi_amd/csrc/planner.hpp)
and the entry points' argument checks (idi_amd/csrc/checks.hpp), the same headers the library
compiles, with g++ -fsanitize=address,undefined -fno-sanitize-recover=all.  Any memory error or
undefined behaviour aborts the driver, and the test fails.  Checked here:
  * every argument check returns the documented code on good and bad arguments;
  * the planner's tables on the detectors of tests/test_planner_cpu.py equal that file's NumPy
    restatement of the same coefficients (mirror-pair remainder, 2 x 4 block spread, wave
    reference offset, linear-reference residual);
  * in the far field the planner's bounds grow with the extent; the eligible tiles are exactly those within the
    tolerance (none of the mirror-pair tiles with per-term Kahan sums, never the reserved index 4),
    and the bound of the linear-reference tile on the 100k-site image equals the one the library
    reported on the GPU for that workload.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import K_UM
from test_planner_cpu import (CASES, _block_bound_coef, _bound_coef, _cent_true_coef, _lin_residual, _ref_bound_coef,
                               _ref_true_coef)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 5e-6


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"):
        pytest.skip("needs g++ and the ROCm headers")
    build = tmp_path_factory.mktemp("host_asan")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "host_asan"), f"BUILD={build}"], check=True)
    return str(build / "planner_check")


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    return env


def _plan(driver, geoms, rhos, summode=0):
    lines = []
    for g in geoms:
        M = np.asarray(g.M, dtype=np.float64).ravel()
        lines.append(" ".join(str(v) for v in (
            g.n0, g.n1, repr(float(g.pixelsize)), repr(float(g.detz)), repr(float(g.k)), repr(float(g.detoffset[0])),
            repr(float(g.detoffset[1])), int(g.nf), int(g.scale), int(g.secondorder), int(g.rot), *(repr(float(v)) for v in M),
            TOL, summode, len(rhos), *(repr(float(v)) for v in rhos))))
    out = subprocess.run([driver], input="\n".join(lines) + "\n", capture_output=True, text=True, env=_env())
    assert out.returncode == 0, out.stderr[-3000:]
    return [json.loads(s) for s in out.stdout.splitlines()]


def test_argument_checks(driver):
    out = subprocess.run([driver, "checks"], capture_output=True, text=True, env=_env())
    assert out.returncode == 0, out.stderr[-3000:]
    codes = json.loads(out.stdout)
    assert len(codes) >= 30
    for name, code in codes.items():
        if name.endswith("_ok"):
            assert code == 0, name
        elif name == "option_tile_reserved":
            assert code == -5, name  # IDISIM_E_UNSUPPORTED
        else:
            assert code == -1, name  # IDISIM_E_INVALID


def _geom(args, nf=False):
    from oracle import Geometry

    return Geometry(args["n0"], args["n1"], args["pixelsize"], args["detz"], K_UM, nf=nf,
                    detoffset=args.get("detoffset", (0.0, 0.0)), detangles=args.get("detangles", (0.0, 0.0)))


def test_planner_tables_match_the_numpy_restatement(driver):
    geoms = [_geom(a) for a, _ in CASES]
    res = _plan(driver, geoms, [0.01])
    for (args, _), g, r in zip(CASES, geoms, res):
        # the compiled planner takes R_min over the pixel range its tiles reach (one pixel before the
        # first, to the end of the last 32 x 32 tile), the restatement over the detector's pixels: the
        # same on the centred detectors, no larger elsewhere, so the compiled coefficients are at
        # least the restatement's
        for key, ref in (("pair_coef", _bound_coef(g)), ("block_coef", _block_bound_coef(g)),
                         ("ref_coef", _ref_bound_coef(g))):
            if not any(args.get(k) for k in ("detoffset", "detangles")):
                assert r[key] == pytest.approx(ref, rel=1e-12), (args, key)
            assert ref * (1 - 1e-12) <= r[key] <= 1.2 * ref, (args, key, r[key], ref)
    small = [_geom(a) for a, _ in CASES if a["n0"] * a["n1"] <= 64 * 64]
    for g, r in zip(small, _plan(driver, small, [0.01])):
        # the FP32 roundings of D and J may fall differently (FP64 formed in another order)
        assert r["lin_coef"] == pytest.approx(1.05 * _lin_residual(g), rel=1e-4)
    # the exact table of the sub-block reference offset (variant 14): the restatement's maximum with
    # FP64 references (the compiled table uses the FP32 D_k of the kernel, 1e-5 of the offset apart)
    mid = [_geom(a) for a, _ in CASES if a["n0"] * a["n1"] <= 512 * 512]
    for g, r in zip(mid, _plan(driver, mid, [0.01])):
        assert r["refs_coef"] == pytest.approx(1.05 * _ref_true_coef(g, 8), rel=1e-3), (g.n0, g.n1)
        assert r["refs_coef"] <= 1.05 * _ref_bound_coef(g, 8)
        # the pair-centre residual of variant 15 (its FP32 reference rounding is 1e-7 of m_0 - m_1,
        # which is up to 10x the residual: compared to 1e-4 of the restatement's value plus that)
        assert r["cent_coef"] == pytest.approx(1.05 * _cent_true_coef(g), rel=1e-4, abs=1e-5 * r["cent_coef"]), (
            g.n0, g.n1)


@pytest.mark.parametrize("nf", [False, True])
def test_planner_bounds_and_eligibility(driver, nf):
    rhos = [0.002, 0.005, 0.01, 0.02, 0.03, 0.05, 0.1]
    geoms = [_geom(a, nf=nf) for a, _ in CASES]
    for summode in (0, 1):
        for r in _plan(driver, geoms, rhos, summode):
            prev = None
            for e in r["extents"]:
                b = [np.inf if v is None else v for v in e["bound"]]
                want = [v for v in range(16) if v != 4 and b[v] <= TOL and not (summode and v >= 5)]
                assert e["eligible"] == want, (e["rho"], e["eligible"], want)
                if prev is not None and not nf:
                    # no curvature term to fold: a larger extent gives a larger bound (in the near
                    # field the bound drops where W stops being folded and its fold error with it)
                    assert all(x >= y for x, y in zip(b, prev)), e["rho"]
                prev = b


def test_subblock_scale_halves_the_truncation(driver):
    """Variant 14 rotates its reference phasors to first order, and a lane whose own angle bound T_k exceeds
    T / sqrt 2 (T the planner's bound, reported as the centre c = T^2 / 4) scales its sums by
    1 / (1 + T_k^2 / 2 - c).  Checked on a grid of T_k <= T and |t| <= T_k in FP64: the error of the scaled
    first-order rotation stays within T^2 / 4 + T^3 / 2, half the T^2 / 2 of the unscaled one, and the lanes
    with T_k <= T / sqrt 2 are left unscaled."""
    g = _geom(dict(n0=256, n1=256, pixelsize=100.0, detz=1e5), nf=True)
    (r,) = _plan(driver, [g], [0.03])
    e = r["extents"][0]
    T, c = e["ref_angle"][14], e["ref_centre"][14]
    assert 1e-3 < T < 1e-2 and c == pytest.approx(T * T / 4, rel=1e-12)
    assert e["ref_centre"][15] == c and e["ref_angle"][15] == T
    assert all(v == 0.0 for i, v in enumerate(e["ref_centre"]) if i not in (14, 15))
    worst, scaled = 0.0, 0
    for tk in np.linspace(0.0, T, 201):
        ck = max(0.0, tk * tk / 2 - c)
        scaled += ck > 0
        t = np.linspace(-tk, tk, 401)
        err = np.abs((1 + 1j * t) / (1 + ck) - np.exp(1j * t))
        worst = max(worst, err.max())
    assert worst <= T * T / 4 + T**3 / 2
    assert worst >= 0.9 * T * T / 4  # attained at T_k = T
    assert 0.25 < scaled / 201 < 0.35  # only T_k > T / sqrt 2
    t = np.linspace(-T, T, 401)
    assert np.abs((1 + 1j * t) - np.exp(1j * t)).max() >= 0.99 * T * T / 2
    # variants 14 and 15 are within the tolerance for the batched shots' detector and 30 nm ball, variant 11
    # (the same rotation from the 16 x 16 wave reference) is not; 15 adds its pair-centre residual to 14's bound
    assert 14 in e["eligible"] and 15 in e["eligible"] and 11 not in e["eligible"]
    # (times 1 + sigma, sigma ~ rho / R = 3e-7 here)
    assert e["bound"][15] - e["bound"][14] == pytest.approx(r["cent_coef"] * 0.03, rel=1e-5)


def test_planner_bound_of_the_benchmark_image(driver):
    """The bound that the library's plan reported on the GPU for the 100k-site image (tile 13,
    extent = max |r_j| of the benchmark's emitters) is what the host build computes."""
    from oracle import Geometry

    rng = np.random.default_rng(2000)
    m = int(np.ceil(0.01 / 3e-4))
    grid = np.arange(-m, m + 1) * 3e-4
    pts = np.stack(np.meshgrid(grid, grid, grid, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[np.linalg.norm(pts, axis=1) <= 0.01 * (1 + 1e-12)]
    sel = pts[rng.choice(len(pts), 100_000, replace=False)]
    rho = float(np.linalg.norm(sel, axis=1).max())
    g = Geometry(512, 512, 75.0, 2e5, K_UM, nf=True)
    (r,) = _plan(driver, [g], [rho])
    assert r["extents"][0]["bound"][13] == pytest.approx(1.6306114528647848e-06, rel=1e-12)


def test_fp64_sincos_against_mpmath(driver):
    """sincos_fp64 of the FP64 kernels (idi_amd/csrc/sincos64.hpp), compiled for the host, against
    40-digit mpmath: random phases of every magnitude up to 1e8 rad, arguments next to multiples
    of pi/2 (where the reduction cancels), zero and both sides of the library fallback at
    2^26 C1.  Each value within 2^-52 of the true sine and cosine (measured: 1.3 x 2^-53; libm 0.5);
    a dropped C3 (1.5e6 x 2^-52), a dropped r^15 term (92 x) or a wrong quadrant sign exceeds it."""
    import mpmath as mp

    mp.mp.dps = 40
    rng = np.random.default_rng(5)
    lim = float.fromhex("0x1.921fb54p+26")
    xs = np.concatenate([
        [0.0, -0.0, 1e-300, 1e-10, 0.5, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2, -np.pi / 2],
        rng.uniform(-10, 10, 1000),
        10 ** rng.uniform(0, 8, 1000) * rng.choice([-1, 1], 1000),
        np.arange(-500, 500) * (np.pi / 2),
        np.arange(1, 500) * (np.pi / 2) * 1e4,
        [np.nextafter(lim, 0), lim, 1.2e8, -1.2e8],
    ])
    out = subprocess.run([driver, "sincos"], input="\n".join(repr(float(x)) for x in xs) + "\n",
                         capture_output=True, text=True, env=_env())
    assert out.returncode == 0, out.stderr[-3000:]
    lines = out.stdout.split()
    assert len(lines) == 2 * len(xs)
    worst = 0.0
    for i, x in enumerate(xs):
        s, c = float.fromhex(lines[2 * i]), float.fromhex(lines[2 * i + 1])
        X = mp.mpf(float(x))
        err = max(abs(mp.mpf(s) - mp.sin(X)), abs(mp.mpf(c) - mp.cos(X)))
        worst = max(worst, float(err))
        assert err <= 2.0**-52, (x, s, c)
    print("sincos_fp64 max abs error / 2^-53:", worst / 2.0**-53)
