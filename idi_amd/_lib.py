"""ctypes binding of libidisim.so (the C ABI in include/idisim.h).

Device memory and streams come from PyTorch (``torch.cuda``): tensors are
passed to the library as raw device pointers, and work is enqueued on the
current torch stream.  There is no CPU fallback: every entry point raises
``RuntimeError`` when the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libidisim.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "idisim.h")

# constants of include/idisim.h
NF, SCALE, SECONDORDER, ROTATE = 1, 2, 4, 8
PREC_FP64, PREC_FP32SPLIT = 0, 1
OUT_FIELD, OUT_INTENSITY = 0, 1
OPT_AUTOTUNE, OPT_TILE, OPT_SPLIT, OPT_SUMMODE, OPT_GRAPH, OPT_TOL_PPB, OPT_FALLBACK = 1, 2, 3, 4, 5, 6, 7
OPT_TIME_PATH = 8
E_PRECISION = -4
ABI_VERSION = 2


class GeometryT(ctypes.Structure):
    _fields_ = [
        ("n0", ctypes.c_int64),
        ("n1", ctypes.c_int64),
        ("pixelsize", ctypes.c_double),
        ("detz", ctypes.c_double),
        ("k", ctypes.c_double),
        ("detoffset", ctypes.c_double * 2),
        ("rot", ctypes.c_double * 9),
        ("flags", ctypes.c_uint32),
    ]


class PlanT(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("tile", ctypes.c_int32),
        ("tile_rows", ctypes.c_int32),
        ("tile_cols", ctypes.c_int32),
        ("split", ctypes.c_int32),
        ("summode", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("fallback", ctypes.c_int32),
        ("phase_bound", ctypes.c_double),
        ("tuned_ms", ctypes.c_double),
        ("extent_exceeded", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


_lib = None


def declared_functions(header=HEADER):
    """Names of the functions include/idisim.h declares."""
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:IDISIM_API\s+)?(?:const\s+)?\w+\s*\*?\s*(idisim_\w+)\s*\(", text, re.M)))


def load():
    """Load libidisim.so (after torch, so that both use one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401  (torch's libamdhip64 must be the process's HIP runtime)
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m idi_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i64, dbl, ip = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    lib.idisim_version.restype = ip
    lib.idisim_last_error.restype = ctypes.c_char_p
    lib.idisim_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(GeometryT), ip, ip]
    lib.idisim_destroy.argtypes = [vp]
    lib.idisim_set_option.argtypes = [vp, ip, i64]
    lib.idisim_get_plan.argtypes = [vp, ctypes.POINTER(PlanT)]
    lib.idisim_forward.argtypes = [vp, vp, vp, i64, dbl, vp, ip, vp]
    lib.idisim_forward_noalloc.argtypes = [vp, vp, vp, i64, dbl, vp, ip, vp]
    lib.idisim_forward_batched.argtypes = [vp, vp, vp, ctypes.POINTER(ctypes.c_int64), i64, dbl, vp, ip, vp]
    lib.idisim_adjoint.argtypes = [vp, vp, vp, i64, dbl, vp, vp, vp, vp, vp, vp]
    lib.idisim_adjoint_l2.argtypes = [vp, vp, vp, i64, dbl, vp, vp, vp, vp, vp, vp, vp]
    lib.idisim_loss_grad_l2.argtypes = [vp, vp, vp, i64, dbl, vp, vp, vp, vp, vp, vp, vp]
    lib.idisim_extent.argtypes = [vp, i64, ctypes.POINTER(ctypes.c_double), vp]
    lib.idisim_generate_ball.argtypes = [ctypes.c_uint64, ctypes.c_uint64, i64, dbl,
                                         ctypes.POINTER(ctypes.c_double), ip, ip, vp, vp]
    lib.idisim_chargesharing.argtypes = [vp, i64, i64, dbl, dbl, ctypes.c_uint64, vp, vp]
    lib.idisim_generate_choice.argtypes = [ctypes.c_uint64, ctypes.c_uint64, i64, i64, vp, vp,
                                           ctypes.POINTER(ctypes.c_double), ip, vp, vp,
                                           ctypes.POINTER(ctypes.c_size_t), vp]
    lib.idisim_time_forward.argtypes = [vp, vp, vp, i64, dbl, dbl, vp, vp]
    lib.idisim_corr_accumulate.argtypes = [vp, vp, i64, i64, vp, vp, vp]
    lib.idisim_qcorr.argtypes = [vp, vp, i64, i64, dbl, dbl, dbl, i64, i64, vp, vp]
    lib.idisim_generate_balls.argtypes = [ctypes.c_uint64, ctypes.c_uint64, i64, dbl, i64, vp, vp, ip, ip, vp, vp]
    lib.idisim_issue_rates.argtypes = [ip] + [ctypes.POINTER(ctypes.c_double)] * 2
    lib.idisim_generate_gauss.argtypes = [ctypes.c_uint64, ctypes.c_uint64, i64, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double), ip, ip, vp, vp]
    for name in ("idisim_create", "idisim_destroy", "idisim_set_option", "idisim_get_plan", "idisim_forward",
                 "idisim_forward_noalloc", "idisim_adjoint_l2", "idisim_loss_grad_l2",
                 "idisim_forward_batched", "idisim_adjoint", "idisim_extent", "idisim_generate_ball",
                 "idisim_chargesharing", "idisim_generate_balls", "idisim_generate_gauss",
                 "idisim_corr_accumulate", "idisim_qcorr", "idisim_time_forward", "idisim_generate_choice",
                 "idisim_issue_rates"):
        getattr(lib, name).restype = ip
    if lib.idisim_version() != ABI_VERSION:
        raise RuntimeError("libidisim.so ABI version mismatch; rebuild with `python -m idi_amd.build`")
    _lib = lib
    return lib


class IdisimError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"idisim error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != 0:
        raise IdisimError(rc, _lib.idisim_last_error().decode(errors="replace"))


def _require_cuda():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("idi_amd needs a ROCm GPU (torch.cuda.is_available() is False); there is no CPU path")
    return torch


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def generate_ball(out, seed, rng_stream=0, radius=1.0, centre=(0.0, 0.0, 0.0), positions=True, phases=True):
    """Fill a contiguous (N, 4) float64 CUDA tensor with emitters uniform in a ball (Philox4x32-10)."""
    import torch

    _require_cuda()
    lib = load()
    if out.dtype != torch.float64 or out.dim() != 2 or out.shape[1] != 4 or not out.is_contiguous() or not out.is_cuda:
        raise ValueError("out must be a contiguous (N, 4) float64 CUDA tensor")
    c = (ctypes.c_double * 3)(*[float(v) for v in centre])
    check(lib.idisim_generate_ball(ctypes.c_uint64(int(seed) & (2**64 - 1)),
                                   ctypes.c_uint64(int(rng_stream) & (2**64 - 1)), int(out.shape[0]),
                                   float(radius), c, int(bool(positions)), int(bool(phases)), _ptr(out), _stream()))
    return out


def _check_emitter_out(out):
    import torch

    if out.dtype != torch.float64 or out.dim() != 2 or out.shape[1] != 4 or not out.is_contiguous() or not out.is_cuda:
        raise ValueError("out must be a contiguous (N, 4) float64 CUDA tensor")


def generate_balls(out, seed, centres, counts, rng_stream=0, radius=1.0, positions=True, phases=True):
    """Fill (N, 4) emitters: counts[b] of them uniform in the ball of `radius` about centres[b] (Philox4x32-10)."""
    import torch

    _require_cuda()
    lib = load()
    _check_emitter_out(out)
    centres = torch.as_tensor(np.asarray(centres, dtype=np.float64).reshape(-1, 3)).to(out.device).contiguous()
    counts = np.asarray(counts, dtype=np.int64).ravel()
    if len(counts) != len(centres) or np.any(counts < 0) or int(counts.sum()) != out.shape[0]:
        raise ValueError("counts must be non-negative, one per centre, summing to N")
    offsets = torch.as_tensor(np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)).to(out.device)
    check(lib.idisim_generate_balls(ctypes.c_uint64(int(seed) & (2**64 - 1)),
                                    ctypes.c_uint64(int(rng_stream) & (2**64 - 1)), int(out.shape[0]), float(radius),
                                    int(len(counts)), _ptr(centres), _ptr(offsets), int(bool(positions)),
                                    int(bool(phases)), _ptr(out), _stream()))
    return out


def generate_gauss(out, seed, sigma, rot=None, rng_stream=0, positions=True, phases=True):
    """Fill (N, 4) emitters with a Gaussian cloud of per-axis standard deviations `sigma`, rotated by `rot`."""
    _require_cuda()
    lib = load()
    _check_emitter_out(out)
    sig = (ctypes.c_double * 3)(*[float(v) for v in np.broadcast_to(np.asarray(sigma, dtype=np.float64), (3,))])
    r = None if rot is None else (ctypes.c_double * 9)(*[float(v) for v in np.asarray(rot, dtype=np.float64).ravel()])
    check(lib.idisim_generate_gauss(ctypes.c_uint64(int(seed) & (2**64 - 1)),
                                    ctypes.c_uint64(int(rng_stream) & (2**64 - 1)), int(out.shape[0]), sig, r,
                                    int(bool(positions)), int(bool(phases)), _ptr(out), _stream()))
    return out


def _check_cuda(name, t, dtype, shape=None):
    if t.dtype != dtype or not t.is_contiguous() or not t.is_cuda or (shape is not None and tuple(t.shape) != shape):
        raise ValueError(f"{name} must be a contiguous CUDA {dtype} tensor" + (f" of shape {shape}" if shape else ""))


def corr_accumulate(corr, maskcorr, count, mean, m2):
    """Masked division and Welford update of (mean, m2) with the count-th correlation (float32 tensors)."""
    import torch

    _require_cuda()
    lib = load()
    for name, t in (("corr", corr), ("maskcorr", maskcorr), ("mean", mean), ("m2", m2)):
        _check_cuda(name, t, torch.float32, tuple(corr.shape))
    check(lib.idisim_corr_accumulate(_ptr(corr), _ptr(maskcorr), int(corr.numel()), int(count), _ptr(mean), _ptr(m2),
                                     _stream()))


def qcorr(zd, img, xcenter, ycenter, z, qmax, maxdqz, out):
    """q-corrected autocorrelation of one float32 image into out (2 maxdqz, 2 qmax, 2 qmax) float64."""
    import torch

    _require_cuda()
    lib = load()
    _check_cuda("zd", zd, torch.float32)
    _check_cuda("img", img, torch.float32, tuple(zd.shape))
    _check_cuda("out", out, torch.float64, (2 * int(maxdqz), 2 * int(qmax), 2 * int(qmax)))
    check(lib.idisim_qcorr(_ptr(zd), _ptr(img), int(zd.shape[0]), int(zd.shape[1]), float(xcenter), float(ycenter),
                           float(z), int(qmax), int(maxdqz), _ptr(out), _stream()))
    return out


class ChoiceSampler:
    """Draws (N, 4) emitters as a choice of N of the given sites without replacement, on the device.

    sites: (M, 3) positions; weights: (M,) non-negative or None (uniform).  The sites,
    weights and workspace stay on the device between draws.
    """

    def __init__(self, sites, weights=None, device=None):
        import torch

        torch_ = _require_cuda()
        self._lib = load()
        dev = torch_.device("cuda", torch_.cuda.current_device() if device is None else int(device))
        self.sites = torch.as_tensor(np.ascontiguousarray(sites, dtype=np.float64).reshape(-1, 3)).to(dev)
        self.set_weights(weights)
        need = ctypes.c_size_t(0)
        check(self._lib.idisim_generate_choice(0, 0, 0, len(self.sites), None, None, None, 0, None, None,
                                               ctypes.byref(need), None))
        self.workspace = torch.empty(int(need.value), dtype=torch.uint8, device=dev)

    def set_weights(self, weights):
        import torch

        if weights is None:
            self.weights = None
        else:
            w = np.ascontiguousarray(weights, dtype=np.float64).ravel()
            if len(w) != len(self.sites):
                raise ValueError("one weight per site")
            self.weights = torch.as_tensor(w).to(self.sites.device)

    def draw(self, out, seed, rng_stream=0, rot=None, phases=True):
        _check_emitter_out(out)
        r = None if rot is None else (ctypes.c_double * 9)(*[float(v) for v in np.asarray(rot, np.float64).ravel()])
        nb = ctypes.c_size_t(self.workspace.numel())
        check(self._lib.idisim_generate_choice(ctypes.c_uint64(int(seed) & (2**64 - 1)),
                                               ctypes.c_uint64(int(rng_stream) & (2**64 - 1)), int(out.shape[0]),
                                               len(self.sites), _ptr(self.sites), _ptr(self.weights), r,
                                               int(bool(phases)), _ptr(out), _ptr(self.workspace), ctypes.byref(nb),
                                               _stream()))
        return out


def issue_rates(device=None):
    """(v_fma_f32, v_sin_f32/v_cos_f32) lane operations per second measured on `device` (default: the
    current one): the issue rates that bound the FP32-split kernels."""
    torch = _require_cuda()
    lib = load()
    dev = torch.cuda.current_device() if device is None else int(device)
    f, t = ctypes.c_double(), ctypes.c_double()
    check(lib.idisim_issue_rates(dev, ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def chargesharing(counts, out, sigma, gain_noise, seed):
    """Charge sharing of a contiguous (n0, n1) float64 CUDA count image into `out` (same shape, no overlap)."""
    import torch

    _require_cuda()
    lib = load()
    for name, t in (("counts", counts), ("out", out)):
        if t.dtype != torch.float64 or t.dim() != 2 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"{name} must be a contiguous 2-d float64 CUDA tensor")
    if tuple(out.shape) != tuple(counts.shape):
        raise ValueError("out must have the shape of counts")
    check(lib.idisim_chargesharing(_ptr(counts), int(counts.shape[0]), int(counts.shape[1]), float(sigma),
                                   float(gain_noise), ctypes.c_uint64(int(seed) & (2**64 - 1)), _ptr(out),
                                   _stream()))
    return out


class Handle:
    """One geometry + precision on one device; owns the library handle."""

    def __init__(self, n0, n1, pixelsize, detz, k, detoffset=(0.0, 0.0), rot=None, nf=False, scale=False,
                 secondorder=False, precision=PREC_FP32SPLIT, device=None):
        torch = _require_cuda()
        lib = load()
        self.device = torch.cuda.current_device() if device is None else int(device)
        g = GeometryT()
        g.n0, g.n1 = int(n0), int(n1)
        g.pixelsize, g.detz, g.k = float(pixelsize), float(detz), float(k)
        g.detoffset[0], g.detoffset[1] = float(detoffset[0]), float(detoffset[1])
        flags = (NF if nf else 0) | (SCALE if scale else 0) | (SECONDORDER if secondorder else 0)
        M = np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64).reshape(3, 3)
        if rot is not None:
            flags |= ROTATE
        for i, v in enumerate(M.ravel()):
            g.rot[i] = float(v)
        g.flags = flags
        self.geom = g
        self.shape = (int(n0), int(n1))
        self.precision = precision
        self._h = ctypes.c_void_p()
        check(lib.idisim_create(ctypes.byref(self._h), ctypes.byref(g), int(precision), self.device))
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.idisim_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def field_dtype(self):
        import torch

        return torch.complex128 if self.precision == PREC_FP64 else torch.complex64

    @property
    def real_dtype(self):
        import torch

        return torch.float64 if self.precision == PREC_FP64 else torch.float32

    def set_option(self, option, value):
        check(self._lib.idisim_set_option(self._h, int(option), int(value)))

    def plan(self):
        p = PlanT()
        check(self._lib.idisim_get_plan(self._h, ctypes.byref(p)))
        return p.as_dict()

    def _on_device(self, **tensors):
        """Every given tensor (None skipped) must live on this handle's GPU: the library reads them
        as device pointers, and a host or other-device pointer would fault the kernel."""
        for name, t in tensors.items():
            if t is not None and (not t.is_cuda or t.device.index != self.device):
                raise ValueError(f"{name} must be a CUDA tensor on device cuda:{self.device} (got {t.device})")

    def forward(self, emit, amp=None, out=None, kind=OUT_FIELD, extent=0.0):
        """emit: (N, 4) float64 CUDA tensor; amp: (N,) float64 or None; returns `out`."""
        import torch

        if emit.dtype != torch.float64 or emit.dim() != 2 or emit.shape[1] != 4 or not emit.is_contiguous():
            raise ValueError("emitters must be a contiguous (N, 4) float64 tensor")
        if amp is not None and (amp.dtype != torch.float64 or amp.shape != (emit.shape[0],) or not amp.is_contiguous()):
            raise ValueError("amplitudes must be a contiguous (N,) float64 tensor")
        self._on_device(emitters=emit, amplitudes=amp, out=out)
        if out is None:
            out = torch.empty(self.shape, dtype=self.field_dtype if kind == OUT_FIELD else self.real_dtype,
                              device=emit.device)
        self._check_out(out, kind, 1)
        check(self._lib.idisim_forward(self._h, _ptr(emit), _ptr(amp), int(emit.shape[0]), float(extent), _ptr(out),
                                       int(kind), _stream()))
        return out

    def forward_noalloc(self, emit, out, amp=None, kind=OUT_FIELD, extent=0.0):
        """idisim_forward_noalloc: as forward(), into a caller-given `out`, with the library's
        static device scratch (no device allocation in the call; calls on one device must not
        overlap)."""
        import torch

        if emit.dtype != torch.float64 or emit.dim() != 2 or emit.shape[1] != 4 or not emit.is_contiguous():
            raise ValueError("emitters must be a contiguous (N, 4) float64 tensor")
        if amp is not None and (amp.dtype != torch.float64 or amp.shape != (emit.shape[0],) or not amp.is_contiguous()):
            raise ValueError("amplitudes must be a contiguous (N,) float64 tensor")
        self._on_device(emitters=emit, amplitudes=amp)
        self._check_out(out, kind, 1)
        check(self._lib.idisim_forward_noalloc(self._h, _ptr(emit), _ptr(amp), int(emit.shape[0]), float(extent),
                                               _ptr(out), int(kind), _stream()))
        return out

    def forward_batched(self, emit, offsets, amp=None, out=None, kind=OUT_FIELD, extent=0.0):
        """emit: concatenated (sum N_s, 4) float64; offsets: host int64 (nshots + 1)."""
        import torch

        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        nshots = len(offs) - 1
        if nshots < 1 or offs[0] != 0 or offs[-1] != emit.shape[0]:
            raise ValueError("offsets must start at 0 and end at the number of emitters")
        if emit.dtype != torch.float64 or emit.dim() != 2 or emit.shape[1] != 4 or not emit.is_contiguous():
            raise ValueError("emitters must be a contiguous (N, 4) float64 tensor")
        if amp is not None and (amp.dtype != torch.float64 or amp.shape != (emit.shape[0],) or not amp.is_contiguous()):
            raise ValueError("amplitudes must be a contiguous (N,) float64 tensor")
        self._on_device(emitters=emit, amplitudes=amp, out=out)
        if out is None:
            out = torch.empty((nshots,) + self.shape,
                              dtype=self.field_dtype if kind == OUT_FIELD else self.real_dtype, device=emit.device)
        self._check_out(out, kind, nshots)
        self._offs = offs  # keep alive for the call
        check(self._lib.idisim_forward_batched(
            self._h, _ptr(emit), _ptr(amp), offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), nshots,
            float(extent), _ptr(out), int(kind), _stream()))
        return out

    def adjoint(self, emit, field, dLdI, amp=None, extent=0.0, want_amp=True):
        """Gradients (pos (N,3), phase (N,), amp (N,) or None) in float32 for a loss L(|E|^2)."""
        import torch

        n = emit.shape[0]
        if emit.dtype != torch.float64 or emit.dim() != 2 or emit.shape[1] != 4 or not emit.is_contiguous():
            raise ValueError("emitters must be a contiguous (N, 4) float64 tensor")
        if amp is not None and (amp.dtype != torch.float64 or amp.shape != (n,) or not amp.is_contiguous()):
            raise ValueError("amplitudes must be a contiguous (N,) float64 tensor")
        if field.dtype != torch.complex64 or tuple(field.shape) != self.shape or not field.is_contiguous():
            raise ValueError("field must be the contiguous complex64 (n0, n1) output of forward")
        if dLdI.dtype != torch.float32 or tuple(dLdI.shape) != self.shape or not dLdI.is_contiguous():
            raise ValueError("dLdI must be a contiguous float32 (n0, n1) tensor")
        self._on_device(emitters=emit, amplitudes=amp, field=field, dLdI=dLdI)
        gpos = torch.empty((n, 3), dtype=torch.float32, device=emit.device)
        gph = torch.empty((n,), dtype=torch.float32, device=emit.device)
        gamp = torch.empty((n,), dtype=torch.float32, device=emit.device) if want_amp else None
        check(self._lib.idisim_adjoint(self._h, _ptr(emit), _ptr(amp), int(n), float(extent), _ptr(field),
                                       _ptr(dLdI), _ptr(gpos), _ptr(gph), _ptr(gamp), _stream()))
        return gpos, gph, gamp

    def adjoint_l2(self, emit, field, target, amp=None, extent=0.0, want_amp=True):
        """The squared-intensity loss L = sum (|E|^2 - target)^2 and its gradients, from the
        complex64 field of forward(): returns (L as a 0-d float64 CUDA tensor, pos (N, 3),
        phase (N,), amp (N,) or None), float32 gradients; dL/dI is formed inside the kernels."""
        import torch

        n = emit.shape[0]
        if n < 1:
            raise ValueError("the loss needs at least one emitter")
        if emit.dtype != torch.float64 or emit.dim() != 2 or emit.shape[1] != 4 or not emit.is_contiguous():
            raise ValueError("emitters must be a contiguous (N, 4) float64 tensor")
        if amp is not None and (amp.dtype != torch.float64 or amp.shape != (n,) or not amp.is_contiguous()):
            raise ValueError("amplitudes must be a contiguous (N,) float64 tensor")
        if field.dtype != torch.complex64 or tuple(field.shape) != self.shape or not field.is_contiguous():
            raise ValueError("field must be the contiguous complex64 (n0, n1) output of forward")
        if target.dtype != torch.float32 or tuple(target.shape) != self.shape or not target.is_contiguous():
            raise ValueError("target must be a contiguous float32 (n0, n1) tensor")
        self._on_device(emitters=emit, amplitudes=amp, field=field, target=target)
        loss = torch.empty((), dtype=torch.float64, device=emit.device)
        gpos = torch.empty((n, 3), dtype=torch.float32, device=emit.device)
        gph = torch.empty((n,), dtype=torch.float32, device=emit.device)
        gamp = torch.empty((n,), dtype=torch.float32, device=emit.device) if want_amp else None
        check(self._lib.idisim_adjoint_l2(self._h, _ptr(emit), _ptr(amp), int(n), float(extent), _ptr(field),
                                          _ptr(target), _ptr(loss), _ptr(gpos), _ptr(gph), _ptr(gamp), _stream()))
        return loss, gpos, gph, gamp

    def loss_grad_l2(self, emit, target, amp=None, extent=0.0, field=None, want_amp=True):
        """forward(emit, amp, field, OUT_FIELD) and adjoint_l2 on that field in one call, the same
        results bit for bit (idisim_loss_grad_l2): returns (L as a 0-d float64 CUDA tensor, pos (N, 3),
        phase (N,), amp (N,) or None); `field` (complex64 (n0, n1)), when given, receives the field."""
        import torch

        n = emit.shape[0]
        if n < 1:
            raise ValueError("the loss needs at least one emitter")
        if emit.dtype != torch.float64 or emit.dim() != 2 or emit.shape[1] != 4 or not emit.is_contiguous():
            raise ValueError("emitters must be a contiguous (N, 4) float64 tensor")
        if amp is not None and (amp.dtype != torch.float64 or amp.shape != (n,) or not amp.is_contiguous()):
            raise ValueError("amplitudes must be a contiguous (N,) float64 tensor")
        if target.dtype != torch.float32 or tuple(target.shape) != self.shape or not target.is_contiguous():
            raise ValueError("target must be a contiguous float32 (n0, n1) tensor")
        if field is not None and (field.dtype != torch.complex64 or tuple(field.shape) != self.shape or
                                  not field.is_contiguous()):
            raise ValueError("field must be a contiguous complex64 (n0, n1) tensor")
        self._on_device(emitters=emit, amplitudes=amp, target=target, field=field)
        loss = torch.empty((), dtype=torch.float64, device=emit.device)
        gpos = torch.empty((n, 3), dtype=torch.float32, device=emit.device)
        gph = torch.empty((n,), dtype=torch.float32, device=emit.device)
        gamp = torch.empty((n,), dtype=torch.float32, device=emit.device) if want_amp else None
        check(self._lib.idisim_loss_grad_l2(self._h, _ptr(emit), _ptr(amp), int(n), float(extent), _ptr(target),
                                            _ptr(field), _ptr(loss), _ptr(gpos), _ptr(gph), _ptr(gamp), _stream()))
        return loss, gpos, gph, gamp

    def time_forward(self, emit, times, c, tau, out=None):
        """Time-integrated intensity (n0, n1) float64 of decaying emitters excited at `times`."""
        import torch

        n = emit.shape[0]
        if emit.dtype != torch.float64 or emit.dim() != 2 or emit.shape[1] != 4 or not emit.is_contiguous():
            raise ValueError("emit must be a contiguous (N, 4) float64 CUDA tensor")
        if times.dtype != torch.float64 or tuple(times.shape) != (n,) or not times.is_contiguous():
            raise ValueError("times must be a contiguous (N,) float64 CUDA tensor")
        self._on_device(emit=emit, times=times, out=out)
        if out is None:
            out = torch.empty(self.shape, dtype=torch.float64, device=emit.device)
        elif out.dtype != torch.float64 or tuple(out.shape) != self.shape or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 tensor of the detector shape")
        check(self._lib.idisim_time_forward(self._h, _ptr(emit), _ptr(times), int(n), float(c), float(tau), _ptr(out),
                                            _stream()))
        return out

    def _check_out(self, out, kind, nshots):
        import torch

        want = self.field_dtype if kind == OUT_FIELD else self.real_dtype
        shape = self.shape if nshots == 1 and out.dim() == 2 else (nshots,) + self.shape
        if out.dtype != want or tuple(out.shape) != shape or not out.is_contiguous() or not out.is_cuda:
            raise ValueError(f"output must be a contiguous CUDA {want} tensor of shape {shape}")
        self._on_device(out=out)
