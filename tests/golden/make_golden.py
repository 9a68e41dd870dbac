"""Write tests/golden/wavefield_golden.json: 50-digit mpmath evaluations of the direct sum.

Each case is a tiny detector (<= 7 x 5 pixels) at a realistic geometry, seven
emitters, and the exact field
    E_p = sum_j a_j m_pj exp(i theta_pj)
with the phase models of idi/simulation/cpu.py:42-66 (near field "nf" and far
field), the 1/r "scale" option (cpu.py:52-54, 64-65), the far-field
"secondorder" term (sim.cu:77-88) and detector rotation/offset (cpu.py:76-80).
Inputs are float64 values; mpmath converts them exactly, so the fixture pins
the mathematics of the formula, not any particular rounding.

Run in a container with mpmath:  python tests/golden/make_golden.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))
LAMBDA_UM = 1.54e-4  # 0.154 nm
K_CFG = 2 * np.pi / LAMBDA_UM


def rotation(alpha, beta, gamma):
    ca, cb, cg = np.cos([alpha, beta, gamma])
    sa, sb, sg = np.sin([alpha, beta, gamma])
    return np.array(
        [
            [cb * cg, sa * sb * cg - ca * sg, ca * sb * cg + sa * sg],
            [cb * sg, sa * sb * sg + ca * cg, ca * sb * sg - sa * cg],
            [-sb, sa * cb, ca * cb],
        ]
    )


def ball(rng, n, r):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v * (r * rng.uniform(0, 1, (n, 1)) ** (1 / 3))


def exact_field(case):
    n0, n1 = case["n0"], case["n1"]
    p, z, k = (mp.mpf(case[key]) for key in ("pixelsize", "detz", "k"))
    off = [mp.mpf(v) for v in case["detoffset"]]
    M = [mp.mpf(v) for v in case["M"]]
    atoms = [[mp.mpf(v) for v in row] for row in case["atoms"]]
    amp = [mp.mpf(v) for v in case["amp"]] if case["amp"] is not None else [mp.mpf(1)] * len(atoms)
    out = []
    for x in range(n0):
        row = []
        for y in range(n1):
            dx = (x - mp.mpf(n0) / 2) * p + off[0]
            dy = (y - mp.mpf(n1) / 2) * p + off[1]
            dz = z
            if case["rot"]:
                dx, dy, dz = (M[3 * i] * dx + M[3 * i + 1] * dy + M[3 * i + 2] * dz for i in range(3))
            R = mp.sqrt(dx * dx + dy * dy + dz * dz)
            acc = mp.mpc(0)
            for (ax, ay, az, aphi), a in zip(atoms, amp):
                if case["nf"]:
                    D = mp.sqrt((dx - ax) ** 2 + (dy - ay) ** 2 + (dz - az) ** 2)
                    theta = (D - dz + az) * k + aphi
                    mag = a / D if case["scale"] else a
                else:
                    qx, qy, qz = k * dx / R, k * dy / R, k * (dz / R - 1)
                    theta = -(ax * qx + ay * qy + az * qz) + aphi
                    if case["secondorder"]:
                        theta += k / (2 * R) * (ax * ax + ay * ay)
                    mag = a / R if case["scale"] else a
                acc += mag * mp.expj(theta)
            row.append([float(acc.real), float(acc.imag)])
        out.append(row)
    return out


def make_cases():
    rng = np.random.default_rng(20240601)
    cases = []

    def add(name, n0, n1, pixelsize, detz, k, radius, nf, scale=False, secondorder=False,
            detangles=(0.0, 0.0), detoffset=(0.0, 0.0), amp=False, n=7):
        pos = ball(rng, n, radius)
        phase = rng.uniform(0, 2 * np.pi, (n, 1))
        atoms = np.concatenate([pos, phase], 1)
        rot = bool(np.any(detangles))
        M = rotation(detangles[0], detangles[1], 0.0) if rot else np.eye(3)
        cases.append(
            dict(
                name=name, n0=n0, n1=n1, pixelsize=float(pixelsize), detz=float(detz), k=float(k),
                detoffset=[float(v) for v in detoffset], detangles=[float(v) for v in detangles],
                rot=rot, M=[float(v) for v in M.ravel()], nf=nf, scale=scale, secondorder=secondorder,
                atoms=atoms.tolist(), amp=rng.uniform(0.5, 1.0, n).tolist() if amp else None,
            )
        )

    # real pixel pitch, detector centre (ball radius, distance and pitch of each workload)
    add("nf_ball50nm_75um_z0.1m_centre", 6, 5, 75.0, 1e5, K_CFG, 0.05, True)
    # same angular range as the full detectors: pitch x (N/6)
    add("nf_ball50nm_128px_z0.1m_wide", 6, 5, 75.0 * 128 / 6, 1e5, K_CFG, 0.05, True)
    add("nf_ball10nm_512px_z0.2m_wide", 6, 5, 75.0 * 512 / 6, 2e5, K_CFG, 0.01, True)
    add("nf_ball30nm_256px_z0.1m_wide", 6, 5, 100.0 * 256 / 6, 1e5, K_CFG, 0.03, True)
    add("nf_ball20nm_256px_z0.1m_wide", 6, 5, 75.0 * 256 / 6, 1e5, K_CFG, 0.02, True)
    add("nf_ball75nm_1024px_z0.5m_wide_amp", 6, 5, 50.0 * 1024 / 6, 5e5, K_CFG, 0.075, True, amp=True)
    add("ff_ball50nm_128px_z0.1m_wide", 6, 5, 75.0 * 128 / 6, 1e5, K_CFG, 0.05, False)
    add("ff_scale_secondorder", 6, 5, 75.0 * 128 / 6, 1e5, K_CFG, 0.05, False, scale=True, secondorder=True)
    add("ff_secondorder_amp", 7, 5, 75.0 * 128 / 7, 1e5, K_CFG, 0.05, False, secondorder=True, amp=True)
    add("nf_scale", 6, 5, 75.0 * 128 / 6, 1e5, K_CFG, 0.05, True, scale=True)
    add("nf_rotated_offset", 6, 5, 2000.0, 1e5, K_CFG, 0.05, True, detangles=(0.3, -0.2),
        detoffset=(500.0, -300.0))
    add("ff_rotated_offset", 6, 5, 2000.0, 1e5, K_CFG, 0.05, False, detangles=(-0.25, 0.4),
        detoffset=(-800.0, 250.0))
    add("nf_odd_7x5", 7, 5, 75.0 * 128 / 7, 1e5, K_CFG, 0.05, True)
    # geometry of the reference's own shape test (tests/test_sim.py:103-108):
    # sphere E=100 eV, r=10 um, 4x4 pixels of 10 um at 1000 um
    add("nf_reftest_4x4_sphere10um", 4, 4, 10.0, 1000.0, 2 * np.pi / (1.24 / 100), 10.0, True)
    add("ff_reftest_4x4_sphere10um", 4, 4, 10.0, 1000.0, 2 * np.pi / (1.24 / 100), 10.0, False)
    return cases


def main():
    cases = make_cases()
    for c in cases:
        c["field"] = exact_field(c)
    path = os.path.join(HERE, "wavefield_golden.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "mp_dps": mp.mp.dps, "cases": cases}, f)
    print("wrote", path, len(cases), "cases")


if __name__ == "__main__":
    main()
