"""Correlators of photon images on the GPU (drop-in for idi.reconstruction's cusimple and cucor).

simple: FFT autocorrelation ``corr`` and the streaming ``correlator2d`` (idi/reconstruction/cusimple.py)
qcor:   q-corrected direct autocorrelation ``corrfunction`` (idi/reconstruction/cucor.py)
"""
from . import qcor, simple  # noqa: F401

cusimple = simple
cucor = qcor
