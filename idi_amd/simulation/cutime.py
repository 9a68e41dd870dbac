"""Time-dependent simulation with decaying emitters on the GPU (drop-in for
idi/simulation/cutime.py:9-138; the CPU module idi/simulation/time.py maps here too).

Each emitter is excited at a random time (Gaussian pulse of FWHM ``pulsewidth``) and
radiates a field that decays with time constant ``tau`` from its arrival at a pixel
on; the detector records the time integral of |E(t)|^2.  The work is
idisim_time_forward: arrival times and amplitudes per (pixel, emitter) in FP64, a
per-pixel segmented sort by arrival time, and one fused decay-scan + integral kernel.
"""
from __future__ import annotations

import numpy as _np

from .. import _lib
from .hip import get_handle


def _precision_settings(settings):
    """'double' -> FP64 handle (float64 arrival times); 'single' / 'mixed' (default) -> float32 times."""
    s = settings or ""
    flags = " ".join(t for t in ("nf", "scale") if t in s)
    return ("single " if "double" not in s else "double ") + flags


def simulate(simobject, Ndet, pixelsize, detz, k, c, tau, pulsewidth, detangles=(0, 0), settings="mixed",
             threads=None, *args, times=None, **kwargs):
    """Time-integrated intensity (Ndet) of ``simobject``'s emitters (cutime.py:9-138).

    simobject: emitters (lengthunit); pixelsize, detz: lengthunit; k: 1/lengthunit;
    c: lengthunit/timeunit; tau: decay time and pulsewidth: FWHM of the excitation
    pulse (timeunit); detangles: detector rotation (theta, phi).
    settings: 'double' | 'single' | 'mixed' (default) precision of the arrival times,
    'nf' for the exact distance, 'scale' for 1/r amplitudes.  ``threads`` is accepted
    for signature compatibility and ignored.  ``times`` (extension) fixes the
    excitation times instead of drawing them with numpy's global generator as the
    reference does.  Returns float64 (float32 for 'single') numpy (N0, N1).
    """
    import torch

    torch_ = _lib._require_cuda()
    h = get_handle(Ndet, pixelsize, detz, k, _precision_settings(settings), detangles)
    dev = torch_.device("cuda", h.device)
    data = _np.ascontiguousarray(simobject.get(), dtype=_np.float64)
    if times is None:
        times = _np.random.randn(simobject.N) * (pulsewidth / 2.35)
    d = torch.from_numpy(data).to(dev)
    t = torch.from_numpy(_np.ascontiguousarray(times, dtype=_np.float64)).to(dev)
    out = h.time_forward(d, t, c, tau)
    res = out.cpu().numpy()
    if "single" in (settings or ""):
        res = res.astype(_np.float32)
    return res


def simulate_gen(simobject, Ndet, pixelsize, detz, k, c, tau, pulsewidth, detangles=(0, 0), settings="mixed",
                 maximg=_np.inf, threads=None, *args, **kwargs):
    """Generator of time-integrated images (time.py:101-122): one ``simulate`` per image."""
    i = 0
    while i < maximg:
        yield simulate(simobject, Ndet, pixelsize, detz, k, c, tau, pulsewidth, detangles, settings, threads, *args,
                       **kwargs)
        i += 1
