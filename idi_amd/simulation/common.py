"""Photon sampling of simulated intensities (drop-in for idi/simulation/common.py:11-23)."""
import numpy as _np


def randomphotons(probs, Nphotons, rng=None, dtype=float):
    """Poisson photon counts with expectation probs * Nphotons / sum(probs) per image.

    The sum runs over the last two axes (one detector image).  ``probs`` may be
    a numpy array or a torch tensor (sampled on its device with torch.poisson).
    Nphotons = inf returns ``probs`` unchanged.
    """
    if not _np.isfinite(Nphotons):
        return probs
    try:
        import torch

        if isinstance(probs, torch.Tensor):
            lam = probs * (Nphotons / probs.sum(dim=(-2, -1), keepdim=True))
            gen = None
            if rng is not None:
                gen = torch.Generator(device=probs.device)
                gen.manual_seed(int(rng.integers(2**63)) if isinstance(rng, _np.random.Generator) else int(rng))
            tdt = torch.float64 if dtype is float else dtype
            return torch.poisson(lam.to(torch.float64), generator=gen).to(tdt)
    except ImportError:
        pass
    if rng is None:
        rng = _np.random.default_rng()
    probs = _np.asarray(probs)
    return rng.poisson(probs * (Nphotons / probs.sum(axis=(-2, -1), keepdims=True))).astype(dtype)
