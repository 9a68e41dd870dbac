"""Direct-sum wavefield simulation on MI355X: drop-in for ``idi.simulation``.

    from idi_amd import simulation as sim
    obj = sim.simobj.sphere(E=8000, N=1000, r=0.05)
    fields = sim.auto.simulate(4, obj, 128, 75, 1e5, settings="single nf")

``hip`` is the backend (simulate, simulate_gen, simulate_batched) and ``auto``
refers to it, as ``idi.simulation.auto`` refers to the CUDA backend when one is
available (idi/simulation/__init__.py:9-31).  ``cuda`` is the same module under the
name the reference's scripts call (scripts/idi_sim.py:83), and ``cpu`` keeps the
reference CPU backend's interface and defaults (far field, FP64, one reused
output buffer; cpu.py:112-212) on the FP64 GPU kernels.  Unlike the reference
there is no CPU fallback: on a machine without a ROCm GPU the calls raise
RuntimeError.
``autograd`` holds the torch.autograd.Function with the hand-written adjoint.
``cutime`` / ``time`` are the time-dependent simulation (decaying emitters; ``autotime``
refers to ``cutime``), and
``common`` holds photon sampling and detector charge sharing.
"""
from . import common, cpu, cutime, hip, simobj, time  # noqa: F401
from .common import chargesharing, randomphotons  # noqa: F401

auto = hip
cuda = hip  # scripts/idi_sim.py:83 calls sim.cuda.simulate
autotime = cutime  # idi/simulation/__init__.py:14,30: the time-dependent backend

__all__ = ["hip", "cpu", "cuda", "simobj", "auto", "autotime", "autograd", "common", "cutime", "time", "randomphotons",
           "chargesharing"]


def __getattr__(name):
    if name == "autograd":  # imported on first use: it imports torch
        import importlib

        return importlib.import_module(".autograd", __name__)
    raise AttributeError(name)
