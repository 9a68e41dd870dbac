"""Build libidisim.so (HIP, gfx950) in-tree with hipcc.

    python -m idi_amd.build          # or: from idi_amd.build import build; build()

The library is a plain C-ABI shared object (include/idisim.h) loaded with
ctypes; it links only the HIP runtime.  Flags:
  --offload-arch=gfx950                  MI355X only
  -O3 -std=c++17 -fPIC -shared
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                         no packed FP32 (v_pk_*_f32) code
  no -ffast-math: the Kahan compensation must survive.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libidisim.so")
SOURCES = ["kernels.hip", "adjoint.hip", "generate.hip", "photons.hip", "correlate.hip", "timesim.hip", "choice.hip", "issue.hip", "idisim.hip"]
ARCH = "gfx950"
FILE_FLAGS = {}


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def compile_flags():
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-fvisibility=hidden",
        "-Wall",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
    ]


def build(verbose=False, jobs=3):
    """Compile every HIP source to an object (in parallel), then link libidisim.so."""
    hipcc = hipcc_path()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        extra = FILE_FLAGS.get(src, [])
        cmd = [hipcc, *compile_flags(), *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait() if procs[0][1].poll() is None else None
            for _, p in procs:
                if p.poll() is None:
                    p.wait()
                    break
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32-ops target feature
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append(src)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
