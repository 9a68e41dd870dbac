"""The large configuration at its full size (one shot): 1e6 emitters in three balls (r = 15/20/25 nm,
counts 125000 / 296296 / 578704), amplitudes U[0.5, 1], 1024 x 1024 px of 50 um at z = 0.5 m, near
field, FP32-split.  Size-independent checks: linearity over the balls and FP64 oracle rows at full N."""
import numpy as np
import pytest

from conftest import K_UM
from gpu_util import ball, rel_l2

pytestmark = pytest.mark.gpu


def _three_balls(rng):
    radii = np.array([0.015, 0.020, 0.025])
    centres = np.array([[-0.040, 0, 0], [0.0, 0, 0], [0.050, 0, 0]])
    counts = np.rint(1_000_000 * radii**3 / np.sum(radii**3)).astype(int)
    counts[-1] = 1_000_000 - counts[:-1].sum()
    pos = np.concatenate([c + ball(rng, k, r) for c, k, r in zip(centres, counts, radii)])
    return np.concatenate([pos, rng.uniform(0, 2 * np.pi, (len(pos), 1))], 1), counts


def test_large_config_full_size(oracle_mod):
    import torch

    import oracle
    from idi_amd import _lib

    rng = np.random.default_rng(5000)
    at, counts = _three_balls(rng)
    amp = rng.uniform(0.5, 1.0, len(at))
    n0 = 1024
    h = _lib.Handle(n0, n0, 50.0, 5e5, K_UM, nf=True, precision=_lib.PREC_FP32SPLIT)
    d, da = torch.from_numpy(at).cuda(), torch.from_numpy(amp).cuda()
    full = h.forward(d, da).cpu().numpy()
    assert h.plan()["fallback"] == 0
    # linearity: the field of the union is the sum of the balls' fields
    parts = np.zeros_like(full)
    o = np.concatenate([[0], np.cumsum(counts)])
    for b in range(3):
        parts += h.forward(d[o[b]:o[b + 1]].contiguous(), da[o[b]:o[b + 1]].contiguous()).cpu().numpy()
    lin = rel_l2(full, parts)
    # FP64 oracle on two rows at the full emitter count
    g = oracle.Geometry(n0, n0, 50.0, 5e5, K_UM, nf=True)
    rows = [0, 700]
    ref = oracle_mod.field_c(g, at, amp, rows=rows)
    e_rows = rel_l2(np.abs(full[rows]) ** 2, np.abs(ref) ** 2)
    print("1e6 emitters, 1024^2: linearity relL2", lin, "oracle rows relL2(I)", e_rows, h.plan())
    assert lin <= 1e-5
    assert e_rows <= 1e-4


def test_large_config_batched_two_shots(oracle_mod):
    """Two shots of the large configuration (1e6 emitters each, fresh phases, amplitudes) through
    the persistent batched launch at 1024^2: each shot equals the single-shot forward of the same
    emitters to FP32 summation order, and the FP64 oracle on two rows of the second shot."""
    import torch

    import oracle
    from idi_amd import _lib

    rng = np.random.default_rng(5001)
    at, _ = _three_balls(rng)
    n = len(at)
    shots = [at.copy(), at.copy()]
    shots[1][:, 3] = rng.uniform(0, 2 * np.pi, n)
    amp = rng.uniform(0.5, 1.0, 2 * n)
    n0 = 1024
    h = _lib.Handle(n0, n0, 50.0, 5e5, K_UM, nf=True, precision=_lib.PREC_FP32SPLIT)
    d = torch.from_numpy(np.concatenate(shots)).cuda()
    da = torch.from_numpy(amp).cuda()
    ext = float(np.linalg.norm(at[:, :3], axis=1).max())
    offs = np.array([0, n, 2 * n], dtype=np.int64)
    print("large batched: 2 shots of 1e6 emitters, 1024^2", flush=True)
    out = h.forward_batched(d, offs, da, None, _lib.OUT_INTENSITY, ext)
    torch.cuda.synchronize()
    plan = h.plan()
    assert plan["fallback"] == 0 and plan["extent_exceeded"] == 0
    batched = out.cpu().numpy()
    for s in range(2):
        single = h.forward(d[s * n:(s + 1) * n].contiguous(), da[s * n:(s + 1) * n].contiguous(), None,
                           _lib.OUT_INTENSITY, ext).cpu().numpy()
        e = rel_l2(batched[s], single)
        print("shot", s, "batched vs single relL2(I)", e, plan)
        assert e <= 1e-5
    g = oracle.Geometry(n0, n0, 50.0, 5e5, K_UM, nf=True)
    rows = [200, 512]
    ref = np.abs(oracle_mod.field_c(g, shots[1], amp[n:], rows=rows)) ** 2
    e = rel_l2(batched[1][rows], ref)
    print("shot 1 vs oracle rows relL2(I)", e)
    assert e <= 1e-4
