"""idi_amd -- MI355X-native direct-sum wavefield simulation with the API of idi.simulation.

Layout
  idi_amd/csrc/         HIP kernels for gfx950 and the C ABI (include/idisim.h)
  idi_amd/_lib.py       ctypes binding of libidisim.so
  idi_amd/simulation/   drop-in for idi.simulation: hip (= auto), simobj, autograd
  idi_amd/util/         rotation and random samplers used by the generators
  idi_amd/build.py      builds libidisim.so in-tree with hipcc
"""
__version__ = "0.1.0"

from . import simulation, util  # noqa: F401,E402

__all__ = ["simulation", "util"]
