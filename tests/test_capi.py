"""The C-ABI library without a GPU: it loads, exports every declared symbol, and rejects bad
arguments with error codes and messages (no compute calls)."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    from idi_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        from idi_amd.build import build

        build()
    return _lib.load()


def test_exports_every_declared_symbol(lib):
    from idi_amd import _lib

    names = _lib.declared_functions()
    assert len(names) == 22
    for n in names:
        assert hasattr(lib, n), n
    assert lib.idisim_version() == _lib.ABI_VERSION == 2


def test_exports_only_the_c_abi(lib):
    """Internal C++ symbols are hidden: the dynamic symbol table holds exactly the header's functions."""
    import shutil
    import subprocess

    from idi_amd import _lib

    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    text_syms = sorted({ln.split()[-1] for ln in out.stdout.splitlines() if ln.split()[1:2] == ["T"]})
    assert text_syms == _lib.declared_functions()


def test_argument_errors_without_device(lib):
    from idi_amd import _lib

    h = ctypes.c_void_p()
    assert lib.idisim_create(ctypes.byref(h), None, 1, 0) == -1
    assert b"NULL" in lib.idisim_last_error()
    g = _lib.GeometryT()
    g.n0, g.n1, g.pixelsize, g.detz, g.k = 0, 4, 1.0, 1.0, 1.0
    assert lib.idisim_create(ctypes.byref(h), ctypes.byref(g), 1, 0) == -1
    g.n0 = 4
    assert lib.idisim_create(ctypes.byref(h), ctypes.byref(g), 7, 0) == -1
    g.k = -1.0
    assert lib.idisim_create(ctypes.byref(h), ctypes.byref(g), 1, 0) == -1
    assert lib.idisim_forward(None, None, None, 0, 0.0, None, 0, None) == -1
    assert lib.idisim_forward_noalloc(None, None, None, 0, 0.0, None, 0, None) == -1
    assert lib.idisim_forward_batched(None, None, None, None, 1, 0.0, None, 0, None) == -1
    assert lib.idisim_adjoint(None, None, None, 0, 0.0, None, None, None, None, None, None) == -1
    assert lib.idisim_adjoint_l2(None, None, None, 0, 0.0, None, None, None, None, None, None, None) == -1
    assert lib.idisim_loss_grad_l2(None, None, None, 0, 0.0, None, None, None, None, None, None, None) == -1
    assert lib.idisim_set_option(None, 1, 1) == -1
    assert lib.idisim_get_plan(None, None) == -1
    assert lib.idisim_extent(None, 5, None, None) == -1
    assert lib.idisim_issue_rates(0, None, None) == -1
    assert lib.idisim_destroy(None) == 0


def test_no_cpu_fallback():
    import torch

    if torch.cuda.is_available():
        pytest.skip("this machine has a GPU")
    from idi_amd.simulation import hip, simobj

    with pytest.raises(RuntimeError):
        hip.simulate(1, simobj.sphere(100, 10, 1), 4, 10, 1000, verbose=False)


def test_header_declares_reference_counterparts():
    """Each entry point's comment names what it replaces in the reference (file:line) or says it has none."""
    text = open(os.path.join(ROOT, "include", "idisim.h")).read()
    for token in ("sim.cu:141-149", "cpu.py:68-107", "cuda.py:53-86", "no reference counterpart", "No reference"):
        assert token in text, token


def test_plan_struct_matches_header():
    """The ctypes mirror of idisim_plan has the header's size (int32 x 8, double x 2, int32 x 2)."""
    from idi_amd import _lib

    assert ctypes.sizeof(_lib.PlanT) == 8 * 4 + 2 * 8 + 2 * 4
    assert [f for f, _ in _lib.PlanT._fields_][-2:] == ["extent_exceeded", "reserved"]


def test_inputs_on_another_device_are_rejected_before_any_launch():
    """Every tensor a Handle method passes to the library must live on the handle's GPU: a host
    tensor is refused with ValueError before the ctypes call (checked on the CPU with a handle
    object that holds no library handle, so any call into the library would fail differently)."""
    import torch

    from idi_amd import _lib

    h = object.__new__(_lib.Handle)
    h.device, h.shape, h.precision = 0, (4, 4), _lib.PREC_FP32SPLIT
    h._lib = None  # a library call would raise AttributeError, not ValueError
    h._h = ctypes.c_void_p()
    emit = torch.zeros((5, 4), dtype=torch.float64)
    amp = torch.ones(5, dtype=torch.float64)
    with pytest.raises(ValueError, match="cuda:0"):
        h.forward(emit)
    with pytest.raises(ValueError, match="cuda:0"):
        h.forward_batched(emit, [0, 2, 5])
    with pytest.raises(ValueError, match="cuda:0"):
        h.adjoint(emit, torch.zeros((4, 4), dtype=torch.complex64), torch.zeros((4, 4), dtype=torch.float32), amp)
    with pytest.raises(ValueError, match="cuda:0"):
        h.time_forward(emit, torch.zeros(5, dtype=torch.float64), 1.0, 1.0)
    with pytest.raises(ValueError, match="(?i)cuda"):
        h._check_out(torch.zeros((4, 4), dtype=torch.complex64), _lib.OUT_FIELD, 1)


def test_simulation_namespace():
    """The names idi.simulation offers (simulation/__init__.py:4-31) and the ones the reference
    scripts call: auto, cuda and cpu, and the lazily imported autograd module."""
    from idi_amd import simulation as sim
    from idi_amd.simulation import autograd

    assert sim.auto is sim.hip and sim.cuda is sim.hip
    assert sim.cpu.simulate.__module__ == "idi_amd.simulation.cpu"
    assert sim.autograd is autograd and callable(autograd.intensity)
