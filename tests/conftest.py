import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

LAMBDA_UM = 1.54e-4               # 0.154 nm
K_UM = 2 * np.pi / LAMBDA_UM      # 4.0799e4 1/um


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libidisim.so")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def golden():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "wavefield_golden.json")) as f:
        return json.load(f)["cases"]


def golden_geometry(case):
    import oracle

    return oracle.Geometry(case["n0"], case["n1"], case["pixelsize"], case["detz"], case["k"],
                           tuple(case["detoffset"]), tuple(case["detangles"]), case["nf"], case["scale"],
                           case["secondorder"], M=np.array(case["M"]).reshape(3, 3))


def golden_field(case):
    f = np.array(case["field"])
    return f[..., 0] + 1j * f[..., 1]


def golden_term_scale(case):
    """Sum over emitters of |a_j m_pj| (an upper bound of |E_p|), max over pixels."""
    amp = np.ones(len(case["atoms"])) if case["amp"] is None else np.array(case["amp"])
    s = amp.sum()
    if case["scale"]:
        s /= np.sqrt(case["detz"] ** 2)  # |r_p| >= detz for these detectors (|1/D| ~ 1/detz)
        s *= 1.01
    return s


def fp64_phase_tolerance(case):
    """Per-term phase error bound of the literal FP64 near-field formula: a few ulps of k |r_p|."""
    R = np.hypot(case["detz"], max(case["n0"], case["n1"]) * case["pixelsize"]) + np.hypot(*case["detoffset"])
    return 4 * 2.0**-52 * R * case["k"] + 1e-13
