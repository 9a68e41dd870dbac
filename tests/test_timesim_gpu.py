"""Time-dependent simulation on the GPU (idisim_time_forward) vs the FP64 restatement."""
import numpy as np
import pytest

from conftest import K_UM
from gpu_util import fp64_tol, rel_l2

pytestmark = pytest.mark.gpu

C_UM_FS = 0.299792458


def _atoms(rng, n, r):
    u = rng.normal(size=(n, 3))
    u *= (r * rng.uniform(0, 1, n) ** (1 / 3) / np.linalg.norm(u, axis=1))[:, None]
    return np.concatenate([u, rng.uniform(0, 2 * np.pi, (n, 1))], 1)


def _run(lib, g, at, times, tau, precision):
    import torch

    from gpu_util import make_handle

    h = make_handle(lib, g, precision)
    out = h.time_forward(torch.from_numpy(at).cuda(), torch.from_numpy(times).cuda(), C_UM_FS, tau)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.fixture(scope="module")
def lib():
    from idi_amd import _lib

    _lib.load()
    return _lib


@pytest.mark.parametrize("nf,scale", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("n", [1, 2, 700, 5000, 12000, 20000])
def test_time_forward_matches_restatement(lib, oracle_mod, nf, scale, n):
    """Both precisions, N across the 2048-arrival chunk boundary, arrivals interleaved by the
    path differences (pulse comparable to r/c).  12000 and 20000 emitters take the 512-lane
    workgroups (float64 keys at 20000: the batched sort path)."""
    import oracle

    rng = np.random.default_rng(n)
    at = _atoms(rng, n, 0.02)
    times = rng.normal(0, 5e-4, n)  # fs; the path differences spread arrivals by ~1e-3 fs
    g = oracle.Geometry(9, 7, 300.0, 1e5, K_UM, nf=nf, scale=scale)
    ref = oracle_mod.time_c(g, at, times, C_UM_FS, 1e-3)
    got64 = _run(lib, g, at, times, 1e-3, lib.PREC_FP64)
    got32 = _run(lib, g, at, times, 1e-3, lib.PREC_FP32SPLIT)
    e64, e32 = rel_l2(got64, ref), rel_l2(got32, ref)
    print("n", n, "nf", nf, "scale", scale, "relL2 fp64 keys", e64, "float32 keys", e32)
    # near field: the restatement's |r_j - r_p| - |r_p| carries the FP64 rounding of |r_p| in the phase
    assert e64 <= (fp64_tol(g) if nf else 1e-10)
    assert e32 <= 1e-4


def test_time_forward_pixel_batches_and_limits(lib, oracle_mod):
    """A detector large enough to need several pixel batches, and the long-decay limit
    (2 I / tau -> the time-independent intensity of the same emitters)."""
    import oracle

    rng = np.random.default_rng(11)
    n = 12000  # 64^2 pixels x 12000 emitters: two pixel batches of the 1 GiB scratch budget
    at = _atoms(rng, n, 0.02)
    times = rng.normal(0, 5e-4, n)
    g = oracle.Geometry(64, 64, 150.0, 1e5, K_UM)
    got = _run(lib, g, at, times, 1e-3, lib.PREC_FP64)
    rows = [0, 17, 63]
    ref = oracle_mod.time_c(g, at, times, C_UM_FS, 1e-3)
    assert rel_l2(got[rows], ref[rows]) <= 1e-10
    g2 = oracle.Geometry(16, 16, 300.0, 1e5, K_UM, nf=True)
    at2 = at[:3000]
    I = _run(lib, g2, at2, np.zeros(3000), 1e9, lib.PREC_FP64)
    static = np.abs(oracle_mod.field_c(g2, at2)) ** 2
    assert rel_l2(2 * I / 1e9, static) <= 1e-6


def test_dropin_time_modules(lib, oracle_mod):
    """cutime.simulate / time.simulate signatures (cutime.py:9-22, time.py:47-61) with fixed times."""
    import oracle

    from idi_amd.simulation import cutime, simobj
    from idi_amd.simulation import time as tsim

    obj = simobj.sphere(8048.0, 500, 0.02)
    obj.rng = np.random.default_rng(3)
    obj.get()
    obj.rndPos = False  # keep these positions; phases are then 0 (simobj.py:56-57)
    obj.rndPhase = False
    at = obj.get()
    times = np.random.default_rng(4).normal(0, 5e-4, 500)
    g = oracle.Geometry(12, 10, 300.0, 1e5, K_UM, nf=True, scale=True)
    ref = oracle_mod.time_c(g, at, times, C_UM_FS, 1e-3)
    a = cutime.simulate(obj, (12, 10), 300.0, 1e5, K_UM, C_UM_FS, 1e-3, 1.0, settings="double nf scale", times=times)
    assert a.shape == (12, 10) and a.dtype == np.float64
    assert rel_l2(a, ref) <= fp64_tol(g)
    b = cutime.simulate(obj, (12, 10), 300.0, 1e5, K_UM, C_UM_FS, 1e-3, 1.0, settings="single nf scale", times=times)
    assert b.dtype == np.float32 and rel_l2(b, ref) <= 1e-4
    gff = oracle.Geometry(12, 10, 300.0, 1e5, K_UM)
    c = tsim.simulate(obj, (12, 10), 300.0, 1e5, K_UM, C_UM_FS, 1e-3, 1.0, times=times)
    assert rel_l2(c, oracle_mod.time_c(gff, at, times, C_UM_FS, 1e-3)) <= 1e-10
    imgs = list(cutime.simulate_gen(obj, 8, 300.0, 1e5, K_UM, C_UM_FS, 1e-3, 1e-3, maximg=2))
    assert len(imgs) == 2 and imgs[0].shape == (8, 8) and not np.allclose(imgs[0], imgs[1])


def test_time_paths_agree_and_empty_input(lib):
    """The one-workgroup-per-pixel path and the batched sort path give the same image; no emitters
    give a zero image."""
    import torch

    import oracle
    from gpu_util import make_handle

    rng = np.random.default_rng(21)
    at = _atoms(rng, 3000, 0.02)
    times = rng.normal(0, 5e-4, 3000)
    g = oracle.Geometry(20, 17, 300.0, 1e5, K_UM, nf=True)
    h = make_handle(lib, g, lib.PREC_FP32SPLIT)
    d, t = torch.from_numpy(at).cuda(), torch.from_numpy(times).cuda()
    res = []
    for path in (1, 2):
        h.set_option(lib.OPT_TIME_PATH, path)
        res.append(h.time_forward(d, t, C_UM_FS, 1e-3).cpu().numpy())
    assert rel_l2(res[0], res[1]) <= 1e-12
    h.set_option(lib.OPT_TIME_PATH, 0)
    empty = h.time_forward(d[:0], t[:0], C_UM_FS, 1e-3).cpu().numpy()
    assert np.all(empty == 0)
    with pytest.raises(lib.IdisimError):
        h.time_forward(d, t, C_UM_FS, -1.0)


@pytest.mark.parametrize("n,precision", [(10500, 0), (12288, 0), (12289, 0), (16000, 1), (20480, 1), (20481, 1)])
def test_time_paths_agree_across_fused_capacity(lib, n, precision):
    """The 512-lane one-workgroup-per-pixel kernels against the batched sort path on both sides of
    the fused capacity (12288 emitters with float64 keys, 20480 with float32 keys)."""
    import torch

    import oracle
    from gpu_util import make_handle

    rng = np.random.default_rng(n)
    at = _atoms(rng, n, 0.02)
    times = rng.normal(0, 5e-4, n)
    g = oracle.Geometry(6, 5, 300.0, 1e5, K_UM, nf=True)
    h = make_handle(lib, g, precision)
    d, t = torch.from_numpy(at).cuda(), torch.from_numpy(times).cuda()
    res = []
    for path in (0, 2):
        h.set_option(lib.OPT_TIME_PATH, path)
        res.append(h.time_forward(d, t, C_UM_FS, 1e-3).cpu().numpy())
    assert rel_l2(res[0], res[1]) <= 1e-12
