"""Helpers shared by the GPU tests: workloads, handles and comparisons."""
import numpy as np

from conftest import K_UM


def ball(rng, n, r):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v * r * rng.uniform(0, 1, (n, 1)) ** (1 / 3)


def atoms_ball(rng, n, r):
    return np.concatenate([ball(rng, n, r), rng.uniform(0, 2 * np.pi, (n, 1))], 1)


def sc_ball(rng, n, a, r):
    """n of the simple-cubic sites (spacing a) inside |x| <= r, without replacement, random phases."""
    m = int(np.ceil(r / a))
    g = np.arange(-m, m + 1) * a
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[np.linalg.norm(pts, axis=1) <= r * (1 + 1e-12)]
    idx = rng.choice(len(pts), n, replace=False)
    return np.concatenate([pts[idx], rng.uniform(0, 2 * np.pi, (n, 1))], 1), len(pts)


def make_handle(lib, geom, precision):
    """idi_amd Handle from an oracle.Geometry."""
    return lib.Handle(geom.n0, geom.n1, geom.pixelsize, geom.detz, geom.k, detoffset=geom.detoffset,
                      rot=geom.M if geom.rot else None, nf=geom.nf, scale=geom.scale, secondorder=geom.secondorder,
                      precision=precision)


def run(lib, geom, atoms, amp=None, precision=None, kind=None, handle=None, **opts):
    import torch

    precision = lib.PREC_FP32SPLIT if precision is None else precision
    kind = lib.OUT_FIELD if kind is None else kind
    h = make_handle(lib, geom, precision) if handle is None else handle
    for k, v in opts.items():
        h.set_option(getattr(lib, "OPT_" + k.upper()), v)
    d = torch.from_numpy(np.ascontiguousarray(atoms, dtype=np.float64)).cuda()
    da = None if amp is None else torch.from_numpy(np.ascontiguousarray(amp, dtype=np.float64)).cuda()
    out = h.forward(d, da, None, kind)
    torch.cuda.synchronize()
    return out.cpu().numpy(), h


def fp64_tol(geom):
    """relL2 bound between two FP64 evaluations of the literal near-field formula: a few ulps
    of k |r_p| per term (D ~ |r_p| is rounded before the subtraction of z_p)."""
    if not geom.nf:
        return 1e-9
    half = np.hypot(geom.n0 * geom.pixelsize, geom.n1 * geom.pixelsize) / 2
    R = np.hypot(geom.detz, half + np.hypot(*geom.detoffset))
    return 4 * 2.0**-52 * R * geom.k


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def intensity(f):
    return np.abs(f) ** 2


K = K_UM
