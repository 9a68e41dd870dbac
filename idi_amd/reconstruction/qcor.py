"""q-corrected direct autocorrelation on the GPU (drop-in for idi/reconstruction/cucor.py:9-133).

A pixel pair is binned by the difference of the pixels' momentum transfers on an
Ewald sphere z pixels from the sample, (dqz, dqy, dqx), not by the pixel offset.
The pair loop and the symmetrisation are the HIP kernels behind idisim_qcorr.
"""
from __future__ import annotations

import contextlib
import math
import warnings

import numpy as _np

from .. import _lib


def geometry(shape, z, maxq, xcenter=None, ycenter=None):
    """(zd float32 (Nr, Nc), xcenter, ycenter, qmax, maxdqz) as cucor.py:36-49 sets them up."""
    nr, nc = int(shape[0]), int(shape[1])
    if not 32 <= maxq <= 1024:
        warnings.warn("maxq will be clamped between 32 and 1024")
    qmax = int(max(32, min(1024, maxq, nr, nc)))
    y, x = _np.meshgrid(_np.arange(nc, dtype=_np.float64), _np.arange(nr, dtype=_np.float64))
    xc = _np.float32(xcenter or nr / 2)
    yc = _np.float32(ycenter or nc / 2)
    x -= xc
    y -= yc
    d = _np.sqrt(x**2 + y**2 + z**2)
    zd = _np.array(z / d, _np.float32, order="C")
    maxdqz = int(math.ceil(z * float(_np.max(zd) - _np.min(zd)))) | 1
    return zd, float(xc), float(yc), qmax, maxdqz


@contextlib.contextmanager
def corrfunction(shape, z, maxq, xcenter=None, ycenter=None):
    """GPU autocorrelation with q correction (cucor.py:9-26).

    shape: (Nr, Nc) of the inputs in pixels; z: detector distance in pixels;
    maxq: largest pixel offset (clamped to [32, 1024]); xcenter, ycenter: beam
    centre (default shape / 2).  Yields corr(image) -> float64 array
    val[dqz, dqy, dqx] = sum over pixel pairs image(q) image(q + dq).
    """
    import torch

    torch_ = _lib._require_cuda()
    zd, xc, yc, qmax, maxdqz = geometry(shape, z, maxq, xcenter, ycenter)
    dev = torch_.device("cuda", torch_.cuda.current_device())
    dzd = torch.from_numpy(zd).to(dev)
    out = torch.empty((2 * maxdqz, 2 * qmax, 2 * qmax), dtype=torch.float64, device=dev)
    closed = False

    def corr(input):
        if closed:
            raise ValueError("already closed, use within with statement")
        img = torch.as_tensor(_np.asarray(input) if not isinstance(input, torch.Tensor) else input)
        if tuple(img.shape) != tuple(zd.shape):
            raise ValueError("input has not the shape specified in init!")
        img = img.to(device=dev, dtype=torch.float32).contiguous()
        _lib.qcorr(dzd, img, xc, yc, z, qmax, maxdqz, out)
        return out.cpu().numpy()

    try:
        yield corr
    finally:
        closed = True
        torch.cuda.synchronize(dev)
        del dzd, out
