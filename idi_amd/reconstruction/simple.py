"""FFT autocorrelation on the GPU (drop-in for idi/reconstruction/cusimple.py:6-121).

The transforms are rocFFT calls through torch.fft.  The per-image accumulation of
``correlator2d`` (masked division and running mean / variance) is one fused HIP
kernel, idisim_corr_accumulate.  The inverse transform uses the forward
transform's length on every axis.  The reference's irfftn does not pass it, so
the two agree when the fast length of the last axis is even, which includes
every case the reference's tests use.
"""
from __future__ import annotations

import numpy as _np

from .. import _lib
from ..util import fastlen as _fastlen


def _to_device(x, dtype=None):
    import torch

    torch_ = _lib._require_cuda()
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(_np.asarray(x))
    if dtype is None:
        dtype = t.dtype if t.dtype in (torch.float32, torch.float64) else torch.float64
    dev = t.device if t.is_cuda else torch_.device("cuda", torch_.cuda.current_device())
    return t.to(device=dev, dtype=dtype)


def _autocorr(x, axes, fftshape):
    import torch

    f = torch.fft.rfftn(x, s=fftshape, dim=axes)
    p = f.real * f.real + f.imag * f.imag
    r = torch.fft.irfftn(p, s=fftshape, dim=axes)
    r = torch.fft.fftshift(r, dim=axes)
    sl = [slice(None)] * x.dim()
    for ax, ps in zip(axes, fftshape):
        sl[ax] = slice(ps // 2 - x.shape[ax], ps // 2 + x.shape[ax])
    return r[tuple(sl)]


def corr(input, axes=(-1, -2), norm=False, returngpu=False, **kwargs):
    """Autocorrelation of ``input`` along ``axes`` (default: last two) on the GPU (cusimple.py:6-38).

    The output has 2 n entries along each correlated axis of length n; entry n is
    lag 0 and entry 0 is lag -n (always zero).  ``norm`` divides the input by its
    mean over the other axes and the result by the number of pixel pairs per lag,
    with NaN where fewer than 0.9 pairs exist.  Returns numpy, or a CUDA tensor
    with ``returngpu``.
    """
    import torch

    x = _to_device(input)
    axes = sorted(a + x.dim() if a < 0 else a for a in axes)
    fftshape = [_fastlen(2 * x.shape[ax]) for ax in axes]
    if norm:
        other = [i for i in range(x.dim()) if i not in axes]
        x = x * (1 / (x.mean(dim=other) if other else x.mean()))
    ret = _autocorr(x, axes, fftshape)
    if norm:
        n = corr(torch.ones(tuple(x.shape[ax] for ax in axes), dtype=x.dtype, device=x.device), returngpu=True)
        ret = ret / n
        ret[(...,) + tuple(torch.nonzero(n < 0.9, as_tuple=True))] = float("nan")
    return ret if returngpu else ret.cpu().numpy()


class correlator2d:
    """Streaming 2-d autocorrelation divided by the mask autocorrelation (cusimple.py:44-121).

    mask: mask to use (float32) or None for an all-ones mask
    norm: divide every input by its mean over all pixels
    ``add`` accumulates one image; ``mean``, ``std`` (population) and ``n`` give the statistics.
    """

    def __init__(self, mask=None, norm=True) -> None:
        import torch

        self._fftshape = None
        if mask is not None:
            m = _to_device(mask, torch.float32).contiguous()
            self._maskcorr = self._corr(m)
            self._mask = m
        else:
            self._mask = None
            self._maskcorr = None
        self._mean = None
        self._n = 0
        self._nvar = None
        self._norm = norm

    def _corr(self, x):
        if self._fftshape is None:
            self._fftshape = [_fastlen(2 * s) for s in x.shape[-2:]]
        return _autocorr(x, [x.dim() - 2, x.dim() - 1], self._fftshape).contiguous()

    def add(self, input, mask=None) -> None:
        import torch

        if _np.ndim(input) != 2:
            raise ValueError("input must be 2D.")
        x = _to_device(input, torch.float32).clone()
        if self._norm:
            x /= torch.mean(x)
        if mask is not None:
            m = _to_device(mask, torch.float32)
            x *= m
            maskcorr = self._corr(m)
        elif self._mask is not None:
            x *= self._mask
            maskcorr = self._maskcorr
        elif self._maskcorr is not None:
            maskcorr = self._maskcorr
        else:
            maskcorr = self._corr(torch.ones_like(x))
            self._maskcorr = maskcorr
        c = self._corr(x)
        if self._mean is None:
            self._mean = torch.empty_like(c)
            self._nvar = torch.empty_like(c)
        self._n += 1
        _lib.corr_accumulate(c, maskcorr.contiguous(), self._n, self._mean, self._nvar)

    @property
    def mean(self) -> _np.ndarray:
        return self._mean.cpu().numpy()

    @property
    def n(self) -> int:
        return self._n

    @property
    def std(self) -> _np.ndarray:
        import torch

        return torch.sqrt(self._nvar / self._n).cpu().numpy()
