"""Drop-in for idi/simulation/time.py (the reference's CPU time-dependent simulation).

The same entry points run on the GPU (cutime); ``settings`` defaults to the far field
with float64 arrival times, the reference CPU path's behaviour ('' or 'scale').
"""
from __future__ import annotations

import numpy as _np

from . import cutime as _cutime


def simulate(simobject, Ndet, pixelsize, detz, k, c, tau, pulsewidth, detangles=(0, 0), settings="", threads=None,
             *args, **kwargs):
    """Time-integrated intensity (time.py:47-98), computed on the GPU in FP64."""
    return _cutime.simulate(simobject, Ndet, pixelsize, detz, k, c, tau, pulsewidth, detangles,
                            "double " + (settings or ""), threads, *args, **kwargs)


def simulate_gen(simobject, Ndet, pixelsize, detz, k, c, tau, pulsewidth, detangles=(0, 0), settings="",
                 maximg=_np.inf, threads=None, *args, **kwargs):
    """Generator of ``simulate`` images (time.py:101-122)."""
    i = 0
    while i < maximg:
        yield simulate(simobject, Ndet, pixelsize, detz, k, c, tau, pulsewidth, detangles, settings, threads, *args,
                       **kwargs)
        i += 1
