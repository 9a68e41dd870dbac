"""Photon sampling and detector charge sharing (drop-in for idi/simulation/common.py:11-67)."""
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


def _seed(rng):
    if rng is None:
        return int(_np.random.randint(2**63, dtype=_np.int64))
    if isinstance(rng, _np.random.Generator):
        return int(rng.integers(2**63))
    return int(rng)


def chargesharing(img, s, ns=0, rng=None):
    """Charge sharing between neighbouring pixels, on the GPU (drop-in for common.py:26-48).

    Every count of the photon image ``img`` (2-d; counts are truncated toward zero)
    is a hit at a uniform random position in its pixel; its charge, a Gaussian of
    width ``s`` pixels, is spread over the pixels within w = ceil(2 s); charge that
    leaves the detector is lost.  ``ns`` > 0 scales each hit by 1 + ns N(0, 1).  With
    s <= 0 the image is returned as is (ns == 0) or multiplied per pixel by
    1 + ns N(0, 1).  Returns float64: a CUDA tensor for a torch input, else numpy.
    ``rng`` (numpy Generator or int seed) seeds the device's Philox4x32-10 stream.
    """
    if not s > 0 and ns == 0:
        return img
    import torch

    from .. import _lib

    _lib._require_cuda()
    is_t = isinstance(img, torch.Tensor)
    t = img if is_t else torch.from_numpy(_np.ascontiguousarray(img, dtype=_np.float64))
    if t.dim() != 2:
        raise ValueError("chargesharing takes one 2-d detector image")
    dev = t.device if t.is_cuda else torch.device("cuda", torch.cuda.current_device())
    src = t.to(device=dev, dtype=torch.float64).contiguous()
    if bool((src < 0).any()):
        raise ValueError("photon counts must be non-negative")
    if s > 0 and float(src.sum()) >= 2.0**31:
        raise ValueError("chargesharing expects photon counts (at most 2^31 hits per image); got a larger total")
    out = torch.empty_like(src)
    _lib.chargesharing(src, out, s, ns, _seed(rng))
    return out if is_t else out.cpu().numpy()
