"""Drop-in for idi/simulation/cpu.py: the reference's CPU backend interface, computed on the GPU.

    simulate(Nimg, simobject, Ndet, pixelsize, detz, k=None, settings="", verbose=False,
             detangles=(0, 0), detoffset=(0, 0), *args, **kwargs)
    simulate_gen(simobject, Ndet, pixelsize, detz, k=None, settings="", maximg=inf,
                 detangles=(0, 0), detoffset=(0, 0), *args, **kwargs)

keep the signatures, defaults and return conventions of cpu.py:112-212:
  - ``settings`` defaults to "" (far field, no 1/r scaling); "nf" and "scale" are
    the only tokens that change the result (cpu.py:189-190), every other token is
    ignored, so the arithmetic is always FP64 (complex128), as the numba kernel's;
  - ``simulate`` returns complex128 (Nimg, N0, N1) and prints "n. " per image with
    ``verbose`` (cpu.py:146-156);
  - ``simulate_gen`` yields ONE complex128 (N0, N1) array, overwritten by every next
    image, as cpu.py:195,212 yields its ``result`` buffer; copy an image to keep it.

The images are computed by the literal FP64 HIP kernels (the reference formula,
Kahan per term) through ``idi_amd.simulation.hip``; there is no CPU kernel, and
without a ROCm GPU every call raises RuntimeError.
"""
from __future__ import annotations

import numpy as _np

from . import hip as _hip

__all__ = ["simulate", "simulate_gen"]


def _settings(settings):
    """The tokens cpu.py reads (cpu.py:189-190), on the FP64 kernels."""
    s = settings or ""
    return "double" + (" nf" if "nf" in s else "") + (" scale" if "scale" in s else "")


def simulate_gen(simobject, Ndet, pixelsize, detz, k=None, settings="", maximg=_np.inf, detangles=(0, 0),
                 detoffset=(0, 0), *args, **kwargs):
    """Generator of wavefields (cpu.py:159-212): yields the same complex128 buffer per image."""
    if k is None:
        k = simobject.k  # cpu.py:186-187
    n0, n1 = _hip._ndet(Ndet)
    gen = _hip.simulate_gen(simobject, (n0, n1), pixelsize, detz, k, settings=_settings(settings), maximg=maximg,
                            detangles=detangles, detoffset=detoffset, amplitudes=kwargs.get("amplitudes"))
    result = _np.empty((n0, n1), dtype=_np.complex128)

    def _gen():
        for img in gen:
            result[...] = img
            yield result

    return _gen()


def simulate(Nimg, simobject, Ndet, pixelsize, detz, k=None, settings="", verbose=False, detangles=(0, 0),
             detoffset=(0, 0), *args, **kwargs):
    """Array of Nimg wavefields, complex128 (Nimg, N0, N1) (cpu.py:112-156)."""
    if k is None:
        k = simobject.k  # cpu.py:142-143
    n0, n1 = _hip._ndet(Ndet)  # a scalar Ndet is a square detector (cpu.py:144-145)
    result = _np.empty((Nimg, n0, n1), dtype=_np.complex128)
    gen = simulate_gen(simobject, (n0, n1), pixelsize, detz, k, settings, Nimg, detangles, detoffset, **kwargs)
    for n in range(0, Nimg):
        if verbose:
            print(n, end="", flush=True)
        result[n] = next(gen)
        if verbose:
            print(". ", end="", flush=True)
    return result
