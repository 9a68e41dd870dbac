"""HIP backend of the direct-sum wavefield: drop-in for idi/simulation/cuda.py (and cpu.py).

    simulate(Nimg, simobject, Ndet, pixelsize, detz, k=None, settings="double", verbose=True,
             detangles=(0, 0), detoffset=(0, 0), returngpu=False, *args, **kwargs)
    simulate_gen(simobject, Ndet, pixelsize, detz, k=None, settings="double", maximg=inf,
                 detangles=(0, 0), detoffset=(0, 0), returngpu=False, *args, **kwargs)

have the signatures and return types of cuda.py:141-202 / cuda.py:16-138:
``simulate`` returns complex128 (Nimg, N0, N1) on the host; ``simulate_gen``
yields one (N0, N1) field per image, complex64 for "single" and complex128
otherwise, or a device tensor with ``returngpu=True``.

``settings`` is the reference's substring string (cuda.py:38-74); unknown
tokens are ignored:
    single       FP32-split kernels: the large phase is formed in FP64 and
                 range-reduced, sin/cos and sums run in FP32 (complex64 out)
    double       (default) literal FP64 kernels (complex128 out)
    nf           near field: exact distance |r_p - r_j|
    scale        1/r amplitude
    secondorder  far-field Fresnel term k/(2|r_p|)(x^2 + y^2)
    nofast       accepted, no effect (the kernels never use fast-math)

Extensions (keyword arguments; the reference signatures are unchanged):
    amplitudes   (N,) per-emitter amplitudes, or a callable returning them per
                 image; default: ``simobject.get_amplitudes()`` if present
    intensity    True: return |E|^2 (float) instead of the field
    device_rng   True: draw sphere emitters / random phases on the GPU
    simulate_batched(...)  many shots with uneven emitter counts in one
                 persistent, graph-replayed launch.

There is no CPU path: without a ROCm GPU and the built library every call
raises RuntimeError.
"""
from __future__ import annotations

import numpy as _np

from .. import _lib
from ..util import rotation as _rotation

__all__ = ["simulate", "simulate_gen", "simulate_batched", "get_handle", "parse_settings"]

_HANDLES = {}
_MAX_HANDLES = 8


def parse_settings(settings):
    """settings substrings -> dict of flags (cuda.py:53-74)."""
    s = settings or ""
    return {
        "single": "single" in s,
        "nf": "nf" in s,
        "scale": "scale" in s,
        "secondorder": "secondorder" in s,
    }


def _ndet(Ndet):
    if _np.size(Ndet) == 1:
        n = int(_np.asarray(Ndet).ravel()[0])
        return n, n
    return int(Ndet[0]), int(Ndet[1])


def get_handle(Ndet, pixelsize, detz, k, settings="double", detangles=(0, 0), detoffset=(0, 0), device=None):
    """Cached library handle for one geometry / precision (created on first use)."""
    import torch

    n0, n1 = _ndet(Ndet)
    f = parse_settings(settings)
    rot = _rotation(detangles[0], detangles[1], 0) if _np.any(detangles) else None
    dev = torch.cuda.current_device() if device is None else int(device)
    key = (n0, n1, float(pixelsize), float(detz), float(k), tuple(float(v) for v in detoffset),
           None if rot is None else tuple(rot.ravel()), f["nf"], f["scale"], f["secondorder"], f["single"], dev)
    h = _HANDLES.get(key)
    if h is None:
        if len(_HANDLES) >= _MAX_HANDLES:
            # dropped from the cache only: a generator or an autograd graph may still hold it,
            # and Handle.__del__ frees the library handle once the last reference goes
            _HANDLES.pop(next(iter(_HANDLES)))
        h = _lib.Handle(n0, n1, pixelsize, detz, k, detoffset=detoffset, rot=rot, nf=f["nf"], scale=f["scale"],
                        secondorder=f["secondorder"],
                        precision=_lib.PREC_FP32SPLIT if f["single"] else _lib.PREC_FP64, device=dev)
        _HANDLES[key] = h
    return h


def _amplitudes(simobject, amplitudes):
    if amplitudes is not None:
        a = amplitudes() if callable(amplitudes) else amplitudes
    elif hasattr(simobject, "get_amplitudes"):
        a = simobject.get_amplitudes()
    else:
        a = None
    return None if a is None else _np.ascontiguousarray(a, dtype=_np.float64).reshape(-1)


def _extent(atoms):
    return float(_np.sqrt(_np.max(_np.einsum("ij,ij->i", atoms[:, :3], atoms[:, :3])))) if len(atoms) else 0.0


def _host_extent(simobject, pos, cache):
    """Upper bound of |r_j| for the planner without a pass over the emitters where the object
    fixes one: a crystal's sites all lie within its largest site norm (rotations keep norms),
    a sphere's within its radius.  Otherwise the emitters are measured."""
    from . import simobj as _so

    if isinstance(simobject, _so.crystal):
        allpos = simobject._allpos
        key = (id(allpos), allpos.shape)
        if cache.get("sites") != key:
            cache["sites"] = key
            a64 = _np.asarray(allpos, dtype=_np.float64)  # the sites may be float32; measure in FP64
            cache["site_extent"] = float(_np.sqrt(_np.max(_np.einsum("ij,ij->i", a64, a64)))) if len(a64) else 0.0
        return cache["site_extent"] * (1.0 + 1e-12)
    if type(simobject) is _so.sphere:
        return float(simobject.r) * (1.0 + 1e-12)
    return _extent(pos)


class _Staging:
    """Pinned host buffer + device buffer for one image's emitters."""

    def __init__(self, n, with_amp, device):
        import torch

        self.h = torch.empty((n, 4), dtype=torch.float64, pin_memory=True)
        self.d = torch.empty((n, 4), dtype=torch.float64, device=device)
        self.ha = torch.empty((n,), dtype=torch.float64, pin_memory=True) if with_amp else None
        self.da = torch.empty((n,), dtype=torch.float64, device=device) if with_amp else None
        self.ready = None  # event after the H2D copy; the host buffer may be rewritten once it has passed

    def upload(self, atoms, amp, stream, after=None):
        """Copy one image's emitters (an (N, 4) array, a (positions, phases) pair as get2()
        returns, or None: they are drawn on the device) and amplitudes (None: none) through
        the pinned buffers, asynchronously on `stream`, once the event `after` (the last
        kernel that read the device buffers) has passed.  Returns the event that marks the
        copies done."""
        import torch

        if self.ready is not None:
            self.ready.synchronize()  # the previous copy out of the pinned buffers has finished
        with torch.cuda.stream(stream):
            if after is not None:
                stream.wait_event(after)
            if atoms is not None:
                hv = self.h.numpy()
                if isinstance(atoms, tuple):  # get2(): written straight into the pinned buffer
                    hv[:, :3], hv[:, 3:] = atoms
                else:
                    hv[...] = atoms
                self.d.copy_(self.h, non_blocking=True)
            if amp is not None:
                self.ha.numpy()[...] = amp
                self.da.copy_(self.ha, non_blocking=True)
            self.ready = torch.cuda.Event()
            self.ready.record(stream)
        return self.ready


def _device_source(simobject):
    """How the device can draw this object's emitters: 'ball' (simobj.sphere), 'balls'
    (simobj.multisphere: host centres and counts), 'gauss' (simobj.gauss / gnorm with
    rho = 2), each with fresh positions and phases; 'phases' (fixed positions, random
    phases); or None (host draws)."""
    from . import simobj as _so

    rnd_pos = getattr(simobject, "rndPos", True)
    rnd_phase = getattr(simobject, "rndPhase", False)
    if not rnd_pos and rnd_phase:
        return "phases"
    if not (rnd_pos and rnd_phase):
        return None
    if type(simobject) is _so.sphere:
        return "ball"
    if type(simobject) is _so.multisphere:
        return "balls"
    if isinstance(simobject, _so.crystal):
        return "choice"
    if isinstance(simobject, _so.gnorm) and _np.all(simobject.rho == 2):
        return "gauss"
    return None


def _device_draw(source, simobject, out, seed, cache=None):
    """Draw the emitters of one image into `out` on the device; returns the extent for the planner
    (0.0: let the library measure it).  `cache` keeps device copies (crystal sites) between images."""
    if source == "choice":
        cache = {} if cache is None else cache
        sites = _np.asarray(simobject._allpos, dtype=_np.float64)
        if "sampler" not in cache:
            cache["sampler"] = _lib.ChoiceSampler(sites, device=out.device.index)
            cache["extent"] = float(_np.linalg.norm(sites, axis=1).max())
            cache["weights_id"] = None
        smp = cache["sampler"]
        rot = simobject.orientation()
        if simobject.fwhm is not None:
            w = simobject.site_weights(sites if rot is None else sites @ rot.T)
            if rot is not None or cache["weights_id"] != id(w):
                smp.set_weights(w)
                cache["weights_id"] = id(w)
        smp.draw(out, seed, rot=rot)
        return cache["extent"]
    if source == "ball":
        _lib.generate_ball(out, seed, 0, simobject.r)
        return float(simobject.r)
    if source == "balls":
        centres, counts = simobject.centres_counts()
        _lib.generate_balls(out, seed, centres, counts, radius=simobject.rsphere)
        used = _np.asarray(centres)[_np.asarray(counts) > 0]
        return float(_np.linalg.norm(used, axis=1).max() + simobject.rsphere) if len(used) else 0.0
    if source == "gauss":
        sigma = _np.broadcast_to(simobject.fwhm / (2 * _np.sqrt(2 * _np.log(2))), (3,))
        rot = simobject._rotmatrix if simobject._rotmatrix is not False else None
        _lib.generate_gauss(out, seed, sigma, rot)
        return 0.0
    raise ValueError(source)


def simulate_gen(simobject, Ndet, pixelsize, detz, k=None, settings="double", maximg=_np.inf, detangles=(0, 0),
                 detoffset=(0, 0), returngpu=False, *args, amplitudes=None, intensity=False, device_rng=False,
                 **kwargs):
    """Generator of simulated wavefields (cuda.py:16-138).

    Two-deep pipeline: kernels run on the current stream, copies on a second stream
    ordered with events.  While the GPU computes image i, the host draws the
    emitters of image i+1 and they are copied in through pinned memory; image i+1
    is launched behind image i, image i is copied out into one of two pinned
    buffers (or cloned on the device) and yielded as soon as that copy has landed,
    so the caller works on it while image i+1 computes.

    device_rng=True (extension): ``simobj.sphere``, ``simobj.multisphere`` (ball
    centres and counts still drawn on the host), ``simobj.gauss`` / ``gnorm``
    with rho = 2 and the crystals (site choice on the device, orientation and
    weights on the host) are drawn on the GPU (idisim_generate_ball / _balls /
    _gauss / _choice,
    Philox4x32-10, seed of image i drawn from ``simobject.rng``), and an object
    with fixed positions (rndPos False) and random phases gets its phases drawn
    there; no emitters are copied per image.
    """
    import torch

    _lib._require_cuda()
    if k is None:
        k = simobject.k
    n0, n1 = _ndet(Ndet)
    h = get_handle((n0, n1), pixelsize, detz, k, settings, detangles, detoffset)
    dev = torch.device("cuda", h.device)
    kind = _lib.OUT_INTENSITY if intensity else _lib.OUT_FIELD
    dtype = h.real_dtype if intensity else h.field_dtype

    def _gen():
        count = 0
        if maximg <= 0:
            return
        with torch.cuda.device(dev):
            cs = torch.cuda.current_stream()
            xs = torch.cuda.Stream()  # host <-> device copies (default priority)
            a0 = _amplitudes(simobject, amplitudes)
            stages = [_Staging(simobject.N, a0 is not None, dev) for _ in range(2)]
            outs = [torch.empty((n0, n1), dtype=dtype, device=dev) for _ in range(2)]
            hosts = [] if returngpu else [torch.empty((n0, n1), dtype=dtype, pin_memory=True) for _ in range(2)]
            for t in [st.d for st in stages] + [st.da for st in stages if st.da is not None] + outs:
                t.record_stream(xs)  # device buffers the copy stream uses
            fwd_done = [None, None]  # kernel of the image in slot k finished reading stages[k], writing outs[k]
            copied = [None, None]  # outs[k] copied out to hosts[k]

            source = _device_source(simobject) if device_rng else None
            draw_cache = {}
            if source == "phases":
                fixed = simobject.get()
                for st in stages:
                    cs.wait_event(st.upload(fixed, a0, xs))
                fixed_extent = _extent(fixed)

            def launch(slot, amp):
                st = stages[slot]
                if source is None:
                    pos, phase = simobject.get2()
                    ext = _host_extent(simobject, pos, draw_cache)
                    cs.wait_event(st.upload((pos, phase), amp, xs, fwd_done[slot]))
                else:
                    seed = int(simobject.rng.integers(2**63))
                    if source == "phases":
                        _lib.generate_ball(st.d, seed, 0, positions=False)
                        ext = fixed_extent
                    else:
                        ext = _device_draw(source, simobject, st.d, seed, draw_cache)
                    if amp is not None:
                        cs.wait_event(st.upload(None, amp, xs, fwd_done[slot]))
                if copied[slot] is not None:
                    cs.wait_event(copied[slot])  # outs[slot] still held the image before
                h.forward(st.d, st.da if amp is not None else None, outs[slot], kind, ext)
                fwd_done[slot] = torch.cuda.Event()
                fwd_done[slot].record(cs)

            try:
                launch(0, a0)
                slot = 0
                while count < maximg:
                    count += 1
                    if returngpu:
                        ret = outs[slot].clone()  # on the compute stream, ahead of the next kernels
                    if count < maximg:
                        # the upload of image count+1 is queued on the copy stream ahead of the
                        # copy-out of image count, so it overlaps the kernel of image count
                        launch(1 - slot, _amplitudes(simobject, amplitudes))
                    if not returngpu:
                        with torch.cuda.stream(xs):
                            xs.wait_event(fwd_done[slot])
                            hosts[slot].copy_(outs[slot], non_blocking=True)
                            copied[slot] = torch.cuda.Event()
                            copied[slot].record(xs)
                        copied[slot].synchronize()
                        ret = hosts[slot].numpy().copy()  # pageable, as cupy's get() returns
                    yield ret
                    slot = 1 - slot
            finally:
                xs.synchronize()

    return _gen()


def simulate(Nimg, simobject, Ndet, pixelsize, detz, k=None, settings="double", verbose=True, detangles=(0, 0),
             detoffset=(0, 0), returngpu=False, *args, amplitudes=None, intensity=False, **kwargs):
    """Array of Nimg simulated wavefields (cuda.py:141-202): complex128 (Nimg, N0, N1) on the host,
    or a device tensor with ``returngpu=True``.  Ctrl-C returns the images finished so far."""
    import torch

    n0, n1 = _ndet(Ndet)
    gen = simulate_gen(simobject, (n0, n1), pixelsize, detz, k, settings=settings, maximg=Nimg, detangles=detangles,
                       detoffset=detoffset, returngpu=returngpu, amplitudes=amplitudes, intensity=intensity,
                       device_rng=kwargs.get("device_rng", False))
    if returngpu:
        h = get_handle((n0, n1), pixelsize, detz, simobject.k if k is None else k, settings, detangles, detoffset)
        dt = h.real_dtype if intensity else h.field_dtype
        result = torch.empty((Nimg, n0, n1), dtype=dt, device=torch.device("cuda", h.device))
    else:
        result = _np.empty((Nimg, n0, n1), _np.float64 if intensity else _np.complex128)
    i = -1
    try:
        for i, img in enumerate(gen):
            if verbose:
                print(i, end=". ", flush=True)
            result[i, ...] = img
        return result
    except KeyboardInterrupt:
        if verbose:
            print("Interrupted")
        return result[: max(i, 0)]


def simulate_batched(emitters, Ndet, pixelsize, detz, k, settings="single", detangles=(0, 0), detoffset=(0, 0),
                     amplitudes=None, intensity=False, returngpu=False, extent=None):
    """Many shots with uneven emitter counts in one persistent launch (no reference counterpart).

    emitters:   list of (N_s, 4) float64 arrays (x, y, z, phase), or one
                concatenated array/tensor together with ``offsets`` given as
                ``(array, offsets)``
    amplitudes: None, a list of (N_s,) arrays, or one concatenated array
    returns     (nshots, N0, N1) field (complex64 for "single") or intensity,
                on the host (numpy) or as a device tensor with ``returngpu``.
    Repeated calls with the same shot sizes replay a captured hipGraph.
    """
    import torch

    _lib._require_cuda()
    n0, n1 = _ndet(Ndet)
    h = get_handle((n0, n1), pixelsize, detz, k, settings, detangles, detoffset)
    dev = torch.device("cuda", h.device)
    if isinstance(emitters, tuple):
        cat, offsets = emitters
        offsets = _np.asarray(offsets, dtype=_np.int64)
    else:
        sizes = [len(e) for e in emitters]
        offsets = _np.concatenate([[0], _np.cumsum(sizes)]).astype(_np.int64)
        cat = _np.concatenate([_np.asarray(e, dtype=_np.float64).reshape(-1, 4) for e in emitters])
    if isinstance(cat, torch.Tensor):
        d = cat.to(device=dev, dtype=torch.float64).contiguous()
        ext = float(torch.linalg.vector_norm(d[:, :3], dim=1).max()) if extent is None and len(d) else extent
    else:
        cat = _np.ascontiguousarray(cat, dtype=_np.float64)
        ext = _extent(cat) if extent is None else extent
        d = torch.from_numpy(cat).to(dev)
    da = None
    if amplitudes is not None:
        if isinstance(amplitudes, (list, tuple)):
            amplitudes = _np.concatenate([_np.asarray(a, dtype=_np.float64).ravel() for a in amplitudes])
        da = torch.as_tensor(amplitudes, dtype=torch.float64).to(dev).contiguous()
    kind = _lib.OUT_INTENSITY if intensity else _lib.OUT_FIELD
    out = h.forward_batched(d, offsets, da, None, kind, ext or 0.0)
    return out if returngpu else out.cpu().numpy()
