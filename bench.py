"""Benchmark: emitter-pixel pair evaluations per second on one MI355X, FP32-split vs FP64.

Default workload (one step = one near-field image through the C ABI, emitters
already resident in HBM):
  100,000 of the 155,331 simple-cubic sites (a = 0.3 nm) inside a 10 nm ball,
  random phases, unit amplitudes; 512 x 512 pixels of 75 um at z = 0.2 m;
  lambda = 0.154 nm.  2.62144e10 pairs per step.
The FP32-split kernel is timed for --steps steps after --warmup steps, with HIP
events on the stream the library launches on (torch's current stream).  The
literal FP64 kernel is timed the same way as the comparator.  The FP32-split
image is checked against the FP64 GPU image (full) and against the FP64 oracle
on the rows the CPU baseline computes.

Other workloads (--workload): c1 (1000 emitters, 128^2 px, the plain FP64 kernel and
the CPU restatement on the full image), batched (256 shots, N_s ~ U{1000..50000} in a
30 nm ball, 256^2 px of 100 um at 0.1 m, one persistent graph-replayed launch),
large (1e6 emitters in three balls, amplitudes U[0.5, 1], 1024^2 px of 50 um at
0.5 m, 16 shots), adjoint (4096 emitters, 256^2 px of 75 um at 0.1 m), timedep (time-dependent
simulation: 10000 emitters, 128^2 px, near field, float32 arrival keys), e2e (simulate() on the
default workload's object per image: host or device draw, copies and kernel), photons (randomphotons +
chargesharing of a 512^2 image), correlate (streaming FFT correlator and q-corrected correlator), generate (emitters drawn on
the device).

Prints ONE JSON line (rank 0).  --gpus must be 1: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LAMBDA_UM = 1.54e-4
K = 2 * np.pi / LAMBDA_UM
METRIC = "emitter-pixel pair evals/s, 1 GPU: FP32-split vs FP64 at <=1e-4 rel err"
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (published, packed)
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector (published)
OPS_PER_PAIR_NF = 23       # near field, unit amplitude: 3 sub, 3 mul + 2 add, sqrt, 4 (phase), cos, sin, 8 (Kahan)
OPS_PER_PAIR_NF_AMP = 25
OPS_PER_PAIR_ADJ = 33
# The published FP32 peak counts an FMA as two flops: 157.3 TFLOP/s is 7.865e13 FMA issue slots per
# second over the chip's lanes.  v_sin_f32 / v_cos_f32 issue at a quarter of that rate.
PEAK_SLOTS_PER_S = FP32_PEAK_TFLOPS * 1e12 / 2
TRANS_SLOTS = 4
MIX_FILE = os.path.join(ROOT, "profiles", "valu_mix.json")


def load_mix(family, tile, amp=False):
    """VALU and transcendental instructions per pair of an FP32-split kernel, from the rocprofv3
    counter summaries (profiles/valu_mix.json); None when no summary covers this kernel."""
    try:
        with open(MIX_FILE) as f:
            entries = json.load(f)["entries"]
    except (OSError, KeyError, ValueError):
        return None
    found = [e for e in entries if e["family"] == family and e["tile"] == tile]
    found.sort(key=lambda e: e.get("amp", False) != bool(amp))  # the entry with the same amplitude mode first
    if not found:
        return None
    e = found[0]
    return {"valu": e["valu_per_pair"], "trans": e["trans_per_pair"], "kernel": e["kernel"], "source": e["source"],
            "amp": e.get("amp", False)}


def issue_fields(lib, mix, pairs_per_s):
    """Issue ceiling of an instruction mix (VALU and transcendental per pair) from the v_fma_f32 and
    v_sin_f32 / v_cos_f32 rates measured on this device (the two do not overlap in issue)."""
    fma_rate, trans_rate = lib.issue_rates()
    issue = {"fma_per_s": fma_rate, "trans_per_s": trans_rate}
    if mix is not None:
        ceiling = 1.0 / ((mix["valu"] - mix["trans"]) / fma_rate + mix["trans"] / trans_rate)
        issue.update({"valu_per_pair": mix["valu"], "trans_per_pair": mix["trans"],
                      "ceiling_pairs_per_s": ceiling, "frac": pairs_per_s / ceiling})
    return issue


def roofline_valu(lib, mix, pairs_per_s, ops_per_pair, traffic=None, kernel_ms=None):
    """Roofline fields of a VALU-bound kernel.

    achieved / frac: the FMA issue slots the kernel's instruction stream occupies per second, from its
    measured mix (each VALU instruction one slot per lane, a transcendental TRANS_SLOTS), at two flops per
    slot as the published peak counts them; it cannot exceed 1.  algorithmic: the reference formula's
    ops per pair at the achieved pair rate (the mirror-pair tiles share work across pixels, so this one
    can exceed the peak).  issue: the same mix against the v_fma_f32 / v_sin_f32 rates measured on this
    device in the same run."""
    r = {"bound": "valu", "achieved": None, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": None,
         "traffic": traffic, "kernel_ms": kernel_ms}
    if mix is not None:
        slots = mix["valu"] - mix["trans"] + TRANS_SLOTS * mix["trans"]
        achieved = 2.0 * slots * pairs_per_s / 1e12
        r.update({"achieved": achieved, "frac": achieved / FP32_PEAK_TFLOPS, "slots_per_pair": slots,
                  "valu_per_pair": mix["valu"], "trans_per_pair": mix["trans"], "mix_kernel": mix["kernel"],
                  "mix_source": mix["source"]})
    alg = pairs_per_s * ops_per_pair / 1e12
    r["algorithmic"] = {"ops_per_pair": ops_per_pair, "tflops": alg, "frac": alg / FP32_PEAK_TFLOPS}
    r["issue"] = issue_fields(lib, mix, pairs_per_s)
    return r


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sc_ball(rng, n, a, r):
    m = int(np.ceil(r / a))
    g = np.arange(-m, m + 1) * a
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[np.linalg.norm(pts, axis=1) <= r * (1 + 1e-12)]
    idx = rng.choice(len(pts), n, replace=False)
    return np.concatenate([pts[idx], rng.uniform(0, 2 * np.pi, (n, 1))], 1), len(pts)


def ball(rng, n, r):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v * r * rng.uniform(0, 1, (n, 1)) ** (1 / 3)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def host_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class StepTimer:
    """Wall time of `steps` calls between two synchronisations and, with per_step, HIP events around
    each call on torch's current stream (the library's launch stream).  A pair of event records
    between two steps leaves the GPU idle for about 10 us before the next step's first kernel
    (measured in a kernel trace of the 0.15 ms adjoint step), so the short steps take their wall
    time from a run without them."""

    def __init__(self):
        import torch

        self.torch = torch

    def run(self, fn, steps, per_step=True):
        torch = self.torch
        pairs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                 for _ in range(steps if per_step else 0)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if per_step:
            for a, b in pairs:
                a.record()
                fn()
                b.record()
        else:
            for _ in range(steps):
                fn()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, [a.elapsed_time(b) for a, b in pairs]


def cpu_baseline_rows(oracle, geom, atoms, amp, budget_s, nthreads):
    """Time the FP64 oracle on evenly spaced rows until ~budget_s of CPU work; returns rows, field, rate."""
    n0 = geom.n0
    probe = np.unique(np.linspace(0, n0 - 1, 4).round().astype(np.int64))
    oracle.field_c(geom, atoms, amp, rows=probe[:1], nthreads=nthreads)  # thread start-up, page-in
    t0 = time.perf_counter()
    f0 = oracle.field_c(geom, atoms, amp, rows=probe, nthreads=nthreads)
    dt = time.perf_counter() - t0
    per_row = max(dt / len(probe), 1e-3)
    nrows = int(min(n0, max(4, budget_s / per_row)))
    rows = np.unique(np.linspace(0, n0 - 1, nrows).round().astype(np.int64))
    t0 = time.perf_counter()
    f = oracle.field_c(geom, atoms, amp, rows=rows, nthreads=nthreads)
    dt = time.perf_counter() - t0
    del f0
    return rows, f, len(rows) * geom.n1 * len(atoms) / dt, dt


def bench_single(args):
    import torch

    import oracle
    from idi_amd import _lib

    rng = np.random.default_rng(2000)
    atoms, nsites = sc_ball(rng, 100_000, 3e-4, 0.01)
    n0 = n1 = 512
    geom = oracle.Geometry(n0, n1, 75.0, 2e5, K, nf=True)
    pairs = len(atoms) * n0 * n1
    d = torch.from_numpy(atoms).cuda()
    ext = float(np.linalg.norm(atoms[:, :3], axis=1).max())

    hs = _lib.Handle(n0, n1, 75.0, 2e5, K, nf=True, precision=_lib.PREC_FP32SPLIT)
    if args.tile is not None:  # pin a tile variant (profiling one kernel); the planner's split
        hs.set_option(_lib.OPT_TILE, args.tile)
    if args.split is not None:  # pin the emitter split count
        hs.set_option(_lib.OPT_SPLIT, args.split)
    out = hs.forward(d, None, None, _lib.OUT_FIELD, ext)  # autotune + first run
    torch.cuda.synchronize()
    plan = hs.plan()
    log("fp32-split plan", plan)
    timer = StepTimer()
    timer.run(lambda: hs.forward(d, None, out, _lib.OUT_FIELD, ext), args.warmup)
    wall, _ = timer.run(lambda: hs.forward(d, None, out, _lib.OUT_FIELD, ext), args.steps, per_step=False)
    _, ms = timer.run(lambda: hs.forward(d, None, out, _lib.OUT_FIELD, ext), args.steps)
    split_ms = float(np.mean(ms))
    value = pairs * args.steps / wall
    field_split = out.cpu().numpy()
    log(f"fp32-split: {split_ms:.3f} ms/step (events), wall {wall * 1e3 / args.steps:.3f} ms/step")

    h64 = _lib.Handle(n0, n1, 75.0, 2e5, K, nf=True, precision=_lib.PREC_FP64)
    out64 = h64.forward(d, None, None, _lib.OUT_FIELD, ext)
    torch.cuda.synchronize()
    log("fp64 plan", h64.plan())
    s64 = max(1, min(args.steps, 5))
    timer.run(lambda: h64.forward(d, None, out64, _lib.OUT_FIELD, ext), 1)
    wall64, ms64 = timer.run(lambda: h64.forward(d, None, out64, _lib.OUT_FIELD, ext), s64)
    value64 = pairs * s64 / wall64
    field64 = out64.cpu().numpy()
    log(f"fp64: {float(np.mean(ms64)):.3f} ms/step")

    I_s, I_64 = np.abs(field_split) ** 2, np.abs(field64) ** 2
    rel_vs_fp64 = rel_l2(I_s, I_64)

    cpu = None
    rel_vs_oracle = None
    rel64_vs_oracle = None
    if not args.no_cpu:
        nthreads = host_threads()
        rows, fref, rate, dt = cpu_baseline_rows(oracle, geom, atoms, None, args.cpu_budget, nthreads)
        Iref = np.abs(fref) ** 2
        rel_vs_oracle = rel_l2(I_s[rows], Iref)
        rel64_vs_oracle = rel_l2(I_64[rows], Iref)
        cpu = {"value": rate, "unit": "pair-evals/s", "cores": nthreads, "kind": "port",
               "sample": f"{len(rows)} evenly spaced rows x {n1} px x {len(atoms)} emitters "
                         f"({len(rows) * n1 * len(atoms):.3e} pairs, {dt:.1f} s) of the same image; "
                         f"FP64 C/OpenMP restatement of cpu.py:30-107, {cpu_model()}"}
    rate = pairs / (split_ms * 1e-3)
    roof = roofline_valu(_lib, load_mix("single", plan["tile"]), rate, OPS_PER_PAIR_NF, kernel_ms=split_ms)
    log("roofline", roof)
    traffic = args.traffic
    if traffic is None:
        # HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE for this plan
        for name in ("fwd_lin_sc100k_512px_traffic.json", "fwd_ref_sc100k_512px_traffic.json",
                     "fwd_pair_sc100k_512px_traffic.json", "fwd_split_sc100k_512px_traffic.json"):
            try:
                with open(os.path.join(ROOT, "profiles", name)) as f:
                    rec = json.load(f)
                if rec["plan"]["tile"] == plan["tile"] and rec["plan"]["split"] == plan["split"]:
                    traffic = rec["hbm_bytes_per_launch"]
                    break
            except (OSError, KeyError, ValueError):
                continue
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "near-field image: 100000 of 155331 simple-cubic sites (a=0.3 nm) in a 10 nm ball, "
                               "random phases, 512x512 px of 75 um at z=0.2 m, lambda=0.154 nm",
                   "emitters": len(atoms), "pixels": n0 * n1, "pairs_per_step": pairs,
                   "precision": "FP32-split (FP64 range-reduced phase, FP32 sincos + chunked Kahan sums)",
                   "plan": plan},
        "fp64": {"value": value64, "ms_per_step": wall64 * 1e3 / s64, "kernel_ms": float(np.mean(ms64)),
                 "plan": h64.plan()},
        "speedup_vs_fp64": value / value64,
        "rel_l2_intensity": {"fp32split_vs_fp64_gpu_full": rel_vs_fp64,
                             "fp32split_vs_oracle_rows": rel_vs_oracle,
                             "fp64_gpu_vs_oracle_rows": rel64_vs_oracle},
        "roofline": dict(roof, traffic=traffic,
                         fp64_algorithmic_frac=pairs * OPS_PER_PAIR_NF / (float(np.mean(ms64)) * 1e-3) / 1e12 /
                         FP64_PEAK_TFLOPS),
        "cpu_baseline": cpu,
    }
    return result


def bench_batched(args):
    import torch

    from idi_amd import _lib

    rng = np.random.default_rng(3000)
    nshots = 256
    sizes = rng.integers(1000, 50001, nshots)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n0 = n1 = 256
    pairs = int(offsets[-1]) * n0 * n1
    ext = 0.03
    d = torch.empty((int(offsets[-1]), 4), dtype=torch.float64, device="cuda")
    h = _lib.Handle(n0, n1, 100.0, 1e5, K, nf=True, precision=_lib.PREC_FP32SPLIT)
    if args.tile is not None:  # pin a tile variant (profiling one kernel)
        h.set_option(_lib.OPT_TILE, args.tile)
    out = torch.empty((nshots, n0, n1), dtype=torch.float32, device="cuda")
    seeds = iter(range(1, 1 << 30))

    def step():
        # fresh positions (uniform in the 30 nm ball) and phases for every shot, drawn
        # on the device, then one persistent graph-replayed launch over all shots
        _lib.generate_ball(d, next(seeds), 0, ext)
        h.forward_batched(d, offsets, None, out, _lib.OUT_INTENSITY, ext)

    step()
    torch.cuda.synchronize()
    timer = StepTimer()
    timer.run(step, args.warmup)
    wall, ms = timer.run(step, args.steps)
    # the plain FP64 kernel on the same batch (the emitters of the last step): throughput and the
    # relative L2 error of the FP32-split intensities over all 256 shots
    hf = _lib.Handle(n0, n1, 100.0, 1e5, K, nf=True, precision=_lib.PREC_FP64)
    outf = torch.empty((nshots, n0, n1), dtype=torch.float64, device="cuda")

    def step64():
        hf.forward_batched(d, offsets, None, outf, _lib.OUT_INTENSITY, ext)

    step64()
    torch.cuda.synchronize()
    f64_steps = max(1, args.steps // 3)
    wall64, _ = StepTimer().run(step64, f64_steps)
    rel = float((out.double() - outf).norm() / outf.norm())
    f64 = {"value": pairs * f64_steps / wall64, "ms_per_step": wall64 * 1e3 / f64_steps, "plan": hf.plan()}
    print(f"fp64 batched: {f64['ms_per_step']:.1f} ms/step; relL2(I) fp32-split vs fp64 over 256 shots {rel:.3e}",
          file=sys.stderr)
    # the FP64 oracle on rows of the smallest, a middle and the largest shot of the last step
    rel_oracle = None
    if not args.no_cpu:
        import oracle

        geom = oracle.Geometry(n0, n1, 100.0, 1e5, K, nf=True)
        order = np.argsort(sizes)
        shots = [int(order[0]), int(order[nshots // 2]), int(order[-1])]
        rows = [0, 97, 128, 255]
        host = d.cpu().numpy()
        got, ref = [], []
        for s_ in shots:
            at = host[offsets[s_]:offsets[s_ + 1]]
            ref.append(np.abs(oracle.field_c(geom, at, rows=rows, nthreads=host_threads())) ** 2)
            got.append(out[s_].cpu().numpy()[rows])
        rel_oracle = {"shots": shots, "emitters": [int(sizes[s_]) for s_ in shots], "rows": rows,
                      "rel_l2": rel_l2(np.stack(got), np.stack(ref))}
        log("batched vs oracle rows", rel_oracle)
    return {"metric": METRIC + " [batched 256 shots]", "value": pairs * args.steps / wall, "unit": "pair-evals/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "256 shots, N_s~U{1000..50000} uniform in a 30 nm ball, fresh positions and "
                                   "phases per step drawn on the GPU (Philox), 256x256 px of 100 um at z=0.1 m, "
                                   "near field, intensity out", "pairs_per_step": pairs,
                       "plan": h.plan()},
            "fp64": f64, "speedup_vs_fp64": (pairs * args.steps / wall) / f64["value"],
            "rel_l2_intensity": {"fp32split_vs_fp64_gpu_all_shots": rel, "fp32split_vs_oracle_rows": rel_oracle},
            "roofline": roofline_valu(_lib, load_mix("batched", h.plan()["tile"]), pairs / (np.mean(ms) * 1e-3),
                                      OPS_PER_PAIR_NF, kernel_ms=float(np.mean(ms)))}


def three_balls(rng, n_total=1_000_000):
    """Emitters uniform in three non-overlapping balls of radius 15/20/25 nm (5 nm gaps),
    counts proportional to volume: 125000 / 296296 / 578704."""
    radii = np.array([0.015, 0.020, 0.025])
    centres = np.array([[-0.040, 0, 0], [0.0, 0, 0], [0.050, 0, 0]])
    counts = np.rint(n_total * radii**3 / np.sum(radii**3)).astype(int)
    counts[-1] = n_total - counts[:-1].sum()
    return np.concatenate([c + ball(rng, k, r) for c, k, r in zip(centres, counts, radii)]), counts


def bench_large(args):
    import torch

    from idi_amd import _lib

    rng = np.random.default_rng(5000)
    nshots = 16
    pos, counts = three_balls(rng)
    n = len(pos)
    n0 = n1 = 1024
    pairs = nshots * n * n0 * n1
    offsets = (np.arange(nshots + 1) * n).astype(np.int64)
    d = torch.empty((nshots * n, 4), dtype=torch.float64, device="cuda")
    dpos = torch.from_numpy(pos).cuda()
    for s in range(nshots):
        d[s * n:(s + 1) * n, :3] = dpos
    amp = torch.from_numpy(rng.uniform(0.5, 1.0, nshots * n)).cuda()
    ext = float(np.linalg.norm(pos, axis=1).max())
    h = _lib.Handle(n0, n1, 50.0, 5e5, K, nf=True, precision=_lib.PREC_FP32SPLIT)
    if args.tile is not None:
        h.set_option(_lib.OPT_TILE, args.tile)
    out = torch.empty((nshots, n0, n1), dtype=torch.float32, device="cuda")
    seeds = iter(range(1, 1 << 30))

    def step():
        # fresh phases for every shot (positions fixed), drawn on the device
        _lib.generate_ball(d, next(seeds), 0, positions=False)
        h.forward_batched(d, offsets, amp, out, _lib.OUT_INTENSITY, ext)

    step()
    torch.cuda.synchronize()
    timer = StepTimer()
    timer.run(step, args.warmup)
    wall, ms = timer.run(step, args.steps)
    # the FP64 oracle on two rows of the first and the last shot of the last step, at the full emitter count
    rel_oracle = None
    if not args.no_cpu:
        import oracle

        geom = oracle.Geometry(n0, n1, 50.0, 5e5, K, nf=True)
        rows = [300, 512]
        got, ref = [], []
        ahost = amp.cpu().numpy()
        for s_ in (0, nshots - 1):
            at = d[s_ * n:(s_ + 1) * n].cpu().numpy()
            ref.append(np.abs(oracle.field_c(geom, at, ahost[s_ * n:(s_ + 1) * n], rows=rows,
                                             nthreads=host_threads())) ** 2)
            got.append(out[s_].cpu().numpy()[rows])
        rel_oracle = {"shots": [0, nshots - 1], "rows": rows, "rel_l2": rel_l2(np.stack(got), np.stack(ref))}
        log("large vs oracle rows", rel_oracle)
    return {"metric": METRIC + " [1e6 emitters, 16 shots, amplitudes]", "value": pairs * args.steps / wall,
            "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "1e6 emitters in three balls (r=15/20/25 nm, counts %s), amplitudes U[0.5,1], "
                                   "fresh phases, 16 shots, 1024x1024 px of 50 um at z=0.5 m, near field"
                                   % counts.tolist(), "pairs_per_step": pairs, "plan": h.plan()},
            "rel_l2_intensity": {"fp32split_vs_oracle_rows": rel_oracle},
            "roofline": roofline_valu(_lib, load_mix("batched", h.plan()["tile"], amp=True),
                                      pairs / (np.mean(ms) * 1e-3), OPS_PER_PAIR_NF_AMP, kernel_ms=float(np.mean(ms)))}


def bench_adjoint(args):
    import torch

    from idi_amd import _lib

    rng = np.random.default_rng(4000)
    n = 4096
    n0 = n1 = 256
    pos0 = ball(rng, n, 0.02)
    phase = rng.uniform(0, 2 * np.pi, n)
    at0 = np.concatenate([pos0, phase[:, None]], 1)
    at = at0.copy()
    at[:, :3] += rng.normal(0, 5e-5, (n, 3))
    h = _lib.Handle(n0, n1, 75.0, 1e5, K, nf=True, precision=_lib.PREC_FP32SPLIT)
    if args.tile is not None:  # pin the forward's tile variant
        h.set_option(_lib.OPT_TILE, args.tile)
    if args.split is not None:
        h.set_option(_lib.OPT_SPLIT, args.split)
    target = h.forward(torch.from_numpy(at0).cuda(), None, None, _lib.OUT_INTENSITY).double()
    d = torch.from_numpy(at).cuda()
    ext = float(np.linalg.norm(at[:, :3], axis=1).max())
    field = torch.empty((n0, n1), dtype=torch.complex64, device="cuda")
    pairs = 2 * n * n0 * n1

    target32 = target.float()

    def step():
        # one call: the forward kernel, then the adjoint pass that sums the forward's split partials,
        # forms dL/dI = 2 (I - T) of L = sum (I - T)^2 itself and returns L with the gradients
        return h.loss_grad_l2(d, target32, None, ext, field=field, want_amp=False)

    step()
    torch.cuda.synchronize()
    timer = StepTimer()
    timer.run(step, args.warmup, per_step=False)
    wall, _ = timer.run(step, args.steps, per_step=False)
    h.forward(d, None, field, _lib.OUT_FIELD, ext)
    fwd_plan = h.plan()
    # the adjoint call alone (pixel table, k_adjoint, finalize) on the last field, for the roofline of
    # its dominant kernel
    _, ms_adj = StepTimer().run(lambda: h.adjoint_l2(d, field, target32, None, ext, want_amp=False), args.steps)
    adj_ms = float(np.mean(ms_adj))
    roof = roofline_valu(_lib, load_mix("adjoint", 0), n * n0 * n1 / (adj_ms * 1e-3), OPS_PER_PAIR_ADJ,
                         kernel_ms=adj_ms)
    roof["kernel_ms_note"] = "the adjoint call: pixel table, k_adjoint and its finalize"
    log("roofline", roof)
    # the gradients of one step against long-double central differences of the loss (h = 1e-4 nm for
    # positions, 1e-4 rad for phases) on 24 position and 8 phase coordinates, and against the FP64
    # analytic gradient of the oracle over all 4N coordinates
    check = cpu = None
    if not args.no_cpu:
        import oracle

        geom = oracle.Geometry(n0, n1, 75.0, 1e5, K, nf=True)
        loss, gpos, gph, _ = step()
        loss = float(loss)
        gpos, gph = gpos.cpu().numpy().astype(np.float64), gph.cpu().numpy().astype(np.float64)
        T = target32.double().cpu().numpy()
        crng = np.random.default_rng(4001)
        sel = [(int(j), int(c)) for j, c in zip(crng.choice(n, 32, replace=False), crng.integers(0, 3, 32))]
        sel = sel[:24] + [(j, 3) for j, _ in sel[24:]]
        nt = host_threads()
        t0 = time.perf_counter()
        fd = oracle.loss_central_diff(geom, at, T, sel, [1e-7 if c < 3 else 1e-4 for _, c in sel], nthreads=nt)
        fd_s = time.perf_counter() - t0
        gpu = np.array([gpos[j, c] if c < 3 else gph[j] for j, c in sel])
        rms = np.array([np.sqrt(np.mean(gpos**2)) if c < 3 else np.sqrt(np.mean(gph**2)) for _, c in sel])
        t0 = time.perf_counter()
        fref = oracle.field_c(geom, at, nthreads=nt)
        rpos, rph, _ = oracle.adjoint_c(geom, at, 2 * (np.abs(fref) ** 2 - T), field=fref, nthreads=nt)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": pairs / cpu_s, "unit": "pair-evals/s", "cores": nt, "kind": "port",
               "sample": f"the whole step: FP64 forward and analytic FP64 adjoint of the oracle (C/OpenMP "
                         f"restatement of cpu.py:30-107 and its gradient), {pairs:.3e} pairs, "
                         f"{cpu_s:.2f} s, {cpu_model()}"}
        check = {"fd_coordinates": len(sel), "fd_seconds": fd_s,
                 "loss_rel_err_vs_fp64": abs(loss - float(np.sum((np.abs(fref) ** 2 - T) ** 2))) /
                 float(np.sum((np.abs(fref) ** 2 - T) ** 2)),
                 "max_rel_err_vs_fd": float(np.max(np.abs(gpu - fd) / np.abs(fd))),
                 "max_err_vs_fd_over_abs_plus_rms": float(np.max(np.abs(gpu - fd) / (np.abs(fd) + rms))),
                 "rel_l2_vs_fp64_analytic": {"pos": rel_l2(gpos, rpos), "phase": rel_l2(gph, rph)}}
        log("adjoint check", check)
    return {"metric": METRIC + " [forward + adjoint]", "value": pairs * args.steps / wall, "unit": "pair-evals/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "4096 emitters in a 20 nm ball perturbed by N(0, 0.05 nm), loss "
                                   "sum (I - I_target)^2, 256x256 px of 75 um at z=0.1 m, near field; "
                                   "step = forward + adjoint in one idisim_loss_grad_l2 call (pairs counted "
                                   "twice)", "pairs_per_step": pairs,
                       "forward_plan": fwd_plan, "plan": h.plan()},
            "gradient_check": check, "roofline": roof, "cpu_baseline": cpu}


def bench_c1(args):
    """The reference's CPU-runnable case: 1000 emitters in a 50 nm ball, 128^2 px of 75 um at 0.1 m,
    near field.  Plain FP64 kernel and FP32-split on the GPU, the FP64 CPU restatement on the full image."""
    import torch

    import oracle
    from idi_amd import _lib

    rng = np.random.default_rng(1000)
    n, n0 = 1000, 128
    atoms = np.concatenate([ball(rng, n, 0.05), rng.uniform(0, 2 * np.pi, (n, 1))], 1)
    geom = oracle.Geometry(n0, n0, 75.0, 1e5, K, nf=True)
    pairs = n * n0 * n0
    d = torch.from_numpy(atoms).cuda()
    ext = float(np.linalg.norm(atoms[:, :3], axis=1).max())
    res = {}
    for name, prec in (("fp64", _lib.PREC_FP64), ("fp32split", _lib.PREC_FP32SPLIT)):
        h = _lib.Handle(n0, n0, 75.0, 1e5, K, nf=True, precision=prec)
        out = h.forward(d, None, None, _lib.OUT_FIELD, ext)
        torch.cuda.synchronize()
        timer = StepTimer()
        timer.run(lambda: h.forward(d, None, out, _lib.OUT_FIELD, ext), args.warmup)
        wall, _ = timer.run(lambda: h.forward(d, None, out, _lib.OUT_FIELD, ext), args.steps)
        res[name] = {"value": pairs * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps, "plan": h.plan(),
                     "field": out.cpu().numpy()}
    nthreads = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    ref = oracle.field_c(geom, atoms, nthreads=nthreads)
    cpu_s = time.perf_counter() - t0
    I_ref = np.abs(ref) ** 2
    rel = {k: rel_l2(np.abs(v.pop("field")) ** 2, I_ref) for k, v in res.items()}
    return {"metric": METRIC + " [C1: plain FP64, 1000 emitters, 128^2]", "value": res["fp64"]["value"],
            "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": res["fp64"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
            "config": {"workload": "1000 emitters uniform in a 50 nm ball, random phases, 128x128 px of 75 um at "
                                   "z=0.1 m, near field, plain FP64 kernel", "pairs_per_step": pairs,
                       "plan": res["fp64"]["plan"]},
            "fp32split": res["fp32split"], "rel_l2_intensity_vs_oracle": rel,
            "cpu_baseline": {"value": pairs / cpu_s, "unit": "pair-evals/s", "cores": nthreads, "kind": "port",
                             "sample": "the full 128x128 image; FP64 C/OpenMP restatement of cpu.py:30-107"}}


def bench_timedep(args):
    """Time-dependent simulation: 10000 emitters in a 20 nm ball, 128^2 px of 300 um at z = 0.1 m,
    near field, tau = 1 fs-scale decay comparable to the arrival spread, float32 arrival keys."""
    import torch

    from idi_amd import _lib

    rng = np.random.default_rng(5000)
    n, n0 = 10000, 128
    at = np.concatenate([ball(rng, n, 0.02), rng.uniform(0, 2 * np.pi, (n, 1))], 1)
    times = rng.normal(0, 5e-4, n)  # fs
    h = _lib.Handle(n0, n0, 300.0, 1e5, K, nf=True, precision=_lib.PREC_FP32SPLIT)
    d = torch.from_numpy(at).cuda()
    t = torch.from_numpy(times).cuda()
    out = torch.empty((n0, n0), dtype=torch.float64, device="cuda")
    pairs = n * n0 * n0

    def step():
        h.time_forward(d, t, 0.299792458, 1e-3, out)

    step()
    torch.cuda.synchronize()
    timer = StepTimer()
    timer.run(step, args.warmup)
    wall, ms = timer.run(step, args.steps)
    return {"metric": METRIC + " [time-dependent, decaying emitters]", "value": pairs * args.steps / wall,
            "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp64+f32 keys", "data": "synthetic",
            "config": {"workload": "10000 emitters in a 20 nm ball, excitation times N(0, 0.5 as), tau 1 as, "
                                   "128x128 px of 300 um at z=0.1 m, near field; per pixel: FP64 arrival and "
                                   "amplitude, segmented sort by arrival, decay scan + integral",
                       "pairs_per_step": pairs}}


def bench_e2e(args):
    """End-to-end simulate() on the default workload's object: each image draws 100000 of the
    155331 simple-cubic sites and fresh phases, runs the FP32-split kernel and returns the
    complex64 field to the host.  Host draw (numpy, as the reference's simobj does) and device
    draw (device_rng=True: Philox choice on the GPU) are both timed; the rate includes the draw,
    the PCIe copies and the launches."""
    import torch

    from idi_amd.simulation import hip, simobj

    m = int(np.ceil(0.01 / 3e-4))
    g = np.arange(-m, m + 1) * 3e-4
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[np.linalg.norm(pts, axis=1) <= 0.01 * (1 + 1e-12)]
    obj = simobj.sc(1.24 / LAMBDA_UM, 100_000, 3e-4, repeats=[2 * m + 1] * 3)
    obj._allpos = pts
    obj.rng = np.random.default_rng(2001)
    n0 = 512
    pairs = obj.N * n0 * n0
    kw = dict(settings="single nf", verbose=False)
    rates = {}
    for name, dev in (("host_draw", False), ("device_draw", True)):
        hip.simulate(max(1, args.warmup), obj, n0, 75.0, 2e5, device_rng=dev, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = hip.simulate(args.steps, obj, n0, 75.0, 2e5, device_rng=dev, **kw)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        assert res.shape == (args.steps, n0, n0)
        rates[name] = {"value": pairs * args.steps / wall, "ms_per_image": wall * 1e3 / args.steps}
        log(name, rates[name])
    h = hip.get_handle(n0, 75.0, 2e5, obj.k, "single nf")
    return {"metric": METRIC + " [end-to-end simulate(): draw, copies, kernel]",
            "value": rates["host_draw"]["value"], "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": rates["host_draw"]["ms_per_image"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "simulate(Nimg, sc crystal clipped to a 10 nm ball: 100000 of 155331 sites "
                                   "per image, random phases), 512x512 px of 75 um at z=0.2 m, settings "
                                   "'single nf', complex64 field returned to the host per image",
                       "pairs_per_step": pairs, "sites": int(len(pts)), "plan": h.plan()},
            "device_draw": rates["device_draw"]}


def _sized_cpu_sample(run, sizes, budget_s):
    """Run `run(size)` on growing sizes until one takes about budget_s; returns (size, seconds, result)."""
    res = None
    for size in sizes:
        t0 = time.perf_counter()
        res = run(size)
        dt = time.perf_counter() - t0
        if dt > budget_s / 4:
            return size, dt, res
    return size, dt, res


def bench_photons(args):
    """Photon sampling and detector charge sharing (drop-in for idi/simulation/common.py): a 512x512
    speckle intensity image, randomphotons with 1e6 photons (Poisson on the device), chargesharing
    with sigma 0.75 px (5x5 deposits per hit) and gain noise 0.1.  The unit is charge deposits
    (hit x output pixel) per second."""
    import torch

    import oracle.photons as oph
    from idi_amd.simulation import common

    rng = np.random.default_rng(6000)
    n0, nph, s, ns = 512, 1e6, 0.75, 0.1
    w = int(np.ceil(2 * s))
    inten = torch.from_numpy(rng.exponential(1.0, (n0, n0))).cuda()
    seeds = iter(range(1, 1 << 30))

    def step():
        counts = common.randomphotons(inten, nph, rng=next(seeds))
        return common.chargesharing(counts, s, ns, rng=next(seeds))

    out = step()
    torch.cuda.synchronize()
    timer = StepTimer()
    timer.run(step, args.warmup)
    wall, ms = timer.run(step, args.steps)
    deposits = nph * (2 * w + 1) ** 2
    cpu = None
    if not args.no_cpu:
        counts = common.randomphotons(inten, nph, rng=12345).cpu().numpy()

        def run(m):
            sub = counts[:m, :m]
            oph.chargesharing(sub, s, ns, 7)
            return float(sub.sum())

        m, dt, hits = _sized_cpu_sample(run, (32, 64, 96, 128, 192, 256), args.cpu_budget)
        cpu = {"value": hits * (2 * w + 1) ** 2 / dt, "unit": "charge deposits/s", "cores": 1, "kind": "port",
               "sample": f"{m}x{m} corner of one photon image ({hits:.0f} hits, {dt:.1f} s); the per-hit restatement "
                         f"of common.py:26-67 in Python with math.erf (oracle/photons.py)"}
    return {"metric": "charge deposits/s [photon sampling + charge sharing]", "value": deposits * args.steps / wall,
            "unit": "charge deposits/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
            "config": {"workload": "512x512 exponential speckle intensity, randomphotons 1e6 photons (torch.poisson on "
                                   "the device), chargesharing sigma 0.75 px, gain noise 0.1 (idisim_chargesharing: "
                                   "one thread per output pixel gathers the hits of its 5x5 neighbourhood)",
                       "photons": nph, "deposits_per_step": deposits, "event_ms": float(np.mean(ms))},
            "roofline": None, "cpu_baseline": cpu}


def bench_correlate(args):
    """Correlators of photon images (drop-in for idi/reconstruction/cusimple.py and cucor.py): the
    streaming FFT correlator (correlator2d.add: normalisation, rocFFT autocorrelation of the image,
    masked divide and Welford update) on 512x512 float32 images, and the q-corrected pair correlator
    (corrfunction) on 256x256 images at z = 2000 px with maxq 64."""
    import torch

    import oracle.correlate as ocr
    from idi_amd.reconstruction import qcor, simple

    rng = np.random.default_rng(7000)
    n = 512
    imgs = [torch.from_numpy(rng.poisson(2.0, (n, n)).astype(np.float32)).cuda() for _ in range(4)]
    c2d = simple.correlator2d()
    it = iter(range(1 << 30))

    def step():
        c2d.add(imgs[next(it) % len(imgs)])

    step()
    torch.cuda.synchronize()
    timer = StepTimer()
    timer.run(step, args.warmup)
    wall, ms = timer.run(step, args.steps)

    nq, zq, maxq = 256, 2000.0, 64
    qimgs = [torch.from_numpy(rng.poisson(2.0, (nq, nq)).astype(np.float32)).cuda() for _ in range(2)]
    with qcor.corrfunction((nq, nq), zq, maxq) as qc:
        qc(qimgs[0])
        zd, xc, yc, qmax, maxdqz = qcor.geometry((nq, nq), zq, maxq)
        dzd = torch.from_numpy(zd).cuda()
        qout = torch.empty((2 * maxdqz, 2 * qmax, 2 * qmax), dtype=torch.float64, device="cuda")
        from idi_amd import _lib

        qtimer = StepTimer()
        qtimer.run(lambda: _lib.qcorr(dzd, qimgs[0], xc, yc, zq, qmax, maxdqz, qout), 2)
        qwall, qms = qtimer.run(lambda: _lib.qcorr(dzd, qimgs[0], xc, yc, zq, qmax, maxdqz, qout), 10)
    qpairs = sum(min(nq - r, qmax) for r in range(nq)) * sum(nq - abs(dc) for dc in range(-(qmax - 1), qmax))
    cpu = qcpu = None
    if not args.no_cpu:
        import scipy.signal

        host = imgs[0].cpu().numpy()
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < args.cpu_budget / 2:
            scipy.signal.correlate(host, host, method="fft")
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": reps / dt, "unit": "images/s", "cores": 1, "kind": "port",
               "sample": f"{reps} autocorrelations of one {n}x{n} image with scipy.signal.correlate(method='fft') "
                         f"({dt:.1f} s), the comparator of the reference's tests/test_recon.py:106-123"}
        qhost = qimgs[0].cpu().numpy()

        def run(rows):
            ocr.qcorr(zd[:rows], qhost[:rows], xc, yc, zq, qmax, maxdqz)
            return rows

        rows, qdt, _ = _sized_cpu_sample(run, (4, 8, 16, 32, 64, 128, 256), args.cpu_budget / 2)
        sub_pairs = sum(min(rows - r, qmax) for r in range(rows)) * sum(nq - abs(dc) for dc in range(-(qmax - 1), qmax))
        qcpu = {"value": sub_pairs / qdt, "unit": "pixel pairs/s", "cores": 1, "kind": "port",
                "sample": f"first {rows} of {nq} rows ({sub_pairs:.3e} pairs, {qdt:.1f} s); NumPy restatement of "
                          f"cucor.py:54-96 (oracle/correlate.py)"}
    return {"metric": "images/s [streaming FFT correlator] + pixel pairs/s [q-corrected correlator]",
            "value": args.steps / wall, "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": f"correlator2d.add on {n}x{n} Poisson(2) float32 images (FFT length "
                                   f"{simple._fastlen(2 * n)}); corrfunction on {nq}x{nq} images, z {zq:.0f} px, maxq {maxq}",
                       "event_ms": float(np.mean(ms))},
            "qcorr": {"value": qpairs * 10 / qwall, "unit": "pixel pairs/s", "ms_per_image": qwall * 1e3 / 10,
                      "pairs_per_image": qpairs, "cpu_baseline": qcpu},
            "roofline": None, "cpu_baseline": cpu}


def bench_generate(args):
    """Emitters drawn on the device (drop-in for the draws of idi/simulation/simobj.py): 6.5e6 emitters
    uniform in a 30 nm ball with phases (the emitter total of the batched 256-shot workload), 1e6 in a
    Gaussian cloud, and 100000 of the 155331 simple-cubic sites of a 10 nm ball without replacement.
    The unit is emitters (x, y, z, phase) written per second; the CPU baseline is the numpy draw of the
    same simobj objects on the host."""
    import torch

    from idi_amd import _lib
    from idi_amd.simulation import simobj

    nb, ng = 6_500_000, 1_000_000
    dball = torch.empty((nb, 4), dtype=torch.float64, device="cuda")
    dgauss = torch.empty((ng, 4), dtype=torch.float64, device="cuda")
    m = int(np.ceil(0.01 / 3e-4))
    g = np.arange(-m, m + 1) * 3e-4
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts[np.linalg.norm(pts, axis=1) <= 0.01 * (1 + 1e-12)]
    sampler = _lib.ChoiceSampler(pts)
    dchoice = torch.empty((100_000, 4), dtype=torch.float64, device="cuda")
    seeds = iter(range(1, 1 << 30))
    sigma = np.full(3, 0.02 / (2 * np.sqrt(2 * np.log(2))))
    rates = {}
    for name, n, fn in (("ball", nb, lambda: _lib.generate_ball(dball, next(seeds), 0, 0.03)),
                        ("gauss", ng, lambda: _lib.generate_gauss(dgauss, next(seeds), sigma)),
                        ("choice", 100_000, lambda: sampler.draw(dchoice, next(seeds)))):
        fn()
        torch.cuda.synchronize()
        timer = StepTimer()
        timer.run(fn, args.warmup)
        wall, ms = timer.run(fn, args.steps)
        rates[name] = {"value": n * args.steps / wall, "ms_per_draw": wall * 1e3 / args.steps}
    cpu = None
    if not args.no_cpu:
        ball = simobj.sphere(8048.0, nb, 0.03)
        ball.rng = np.random.default_rng(1)
        t0 = time.perf_counter()
        ball.get()
        dt = time.perf_counter() - t0
        cpu = {"value": nb / dt, "unit": "emitters/s", "cores": 1, "kind": "port",
               "sample": f"one simobj.sphere(N=6.5e6, r=30 nm).get() with numpy on the host ({dt:.1f} s)"}
        crys = simobj.sc(8048.0, 100_000, 3e-4, repeats=[2 * m + 1] * 3)
        crys._allpos = pts
        crys.rng = np.random.default_rng(2)
        t0 = time.perf_counter()
        for _ in range(5):
            crys.get()
        rates["choice"]["cpu"] = 5 * 100_000 / (time.perf_counter() - t0)
    return {"metric": "emitters/s [on-device emitter generation]", "value": rates["ball"]["value"],
            "unit": "emitters/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": rates["ball"]["ms_per_draw"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
            "config": {"workload": "6.5e6 emitters uniform in a 30 nm ball with phases (Philox4x32-10, "
                                   "idisim_generate_ball); also a 1e6 Gaussian cloud and the 100000-of-155331 "
                                   "site choice", "emitters": nb},
            "gauss": rates["gauss"], "choice": rates["choice"], "roofline": None, "cpu_baseline": cpu}


def bench_issue(args):
    """The issue-rate probes alone (idisim_issue_rates: eight independent v_fma_f32 or v_sin_f32 /
    v_cos_f32 chains per lane, 16 workgroups per CU), --steps times: the rates the FP32-split kernels'
    issue ceilings are computed from, in a run short enough to profile with counters."""
    from idi_amd import _lib

    rates = [_lib.issue_rates() for _ in range(max(1, args.steps))]
    fma = max(r[0] for r in rates)
    trans = max(r[1] for r in rates)
    return {"metric": "VALU lane operations/s [issue-rate probes]", "value": fma, "unit": "v_fma_f32 lane ops/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "issue-rate probes: 16 workgroups of 256 lanes per CU, eight independent chains "
                                   "per lane"},
            "trans_per_s": trans, "fma_per_s_runs": [r[0] for r in rates], "trans_per_s_runs": [r[1] for r in rates],
            "published_fma_slots_per_s": PEAK_SLOTS_PER_S, "fma_frac_of_published": fma / PEAK_SLOTS_PER_S,
            "roofline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="single", choices=["single", "c1", "batched", "large", "adjoint", "timedep", "e2e", "photons",
                                                         "correlate", "generate", "issue"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tile", type=int, default=None,
                    help="single, batched, large and adjoint (its forward) workloads: run this FP32-split tile "
                         "variant instead of the planner's choice")
    ap.add_argument("--split", type=int, default=None,
                    help="single and adjoint workloads: run this many emitter splits instead of the planner's choice")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--traffic", type=float, default=None, help="HBM bytes per launch from rocprofv3 --pmc")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the bench, run one elementwise copy of MIB MiB (reads and writes MIB MiB) "
                         "to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is design only)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    res = {"single": bench_single, "batched": bench_batched, "large": bench_large,
           "adjoint": bench_adjoint, "timedep": bench_timedep, "c1": bench_c1, "e2e": bench_e2e,
           "photons": bench_photons, "correlate": bench_correlate, "generate": bench_generate,
           "issue": bench_issue}[args.workload](args)
    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.rand(n, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)  # one elementwise kernel: n*4 bytes read, n*4 bytes written
        torch.cuda.synchronize()
        res["copy_probe_bytes"] = n * 4
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
