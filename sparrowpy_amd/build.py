"""Compile the HIP kernels into ``sparrowpy_amd/lib/libsparrow_hip.so``.

Cross-compiles for gfx950 (MI355X) with hipcc; no GPU is needed to build.
Every product and sum in the geometry kernels must be its own rounding
(``-ffp-contract=off``) for bit-exact visibility and delays; packed FP32
instructions are disabled (``-packed-fp32-ops``).
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libsparrow_hip.so")
SOURCES = ["util.hip", "bake.hip", "exchange.hip", "tiles.hip",
           "collect.hip", "adjoint.hip"]
ARCH = os.environ.get("SPARROW_HIP_ARCH", "gfx950")

HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core accumulators in VGPRs: with AGPRs the compiler copied the
    # per-pair kernel's 64 accumulator registers out and back every step
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]


def _hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; install ROCm to build sparrow_hip")
    return path


def _compile(src, obj):
    cmd = [_hipcc(), *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l)
    return warn


def build(verbose=False):
    """Compile all sources and link the shared library; returns its path."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, s.replace(".hip", ".o")) for s in SOURCES]
    with concurrent.futures.ThreadPoolExecutor(len(SOURCES)) as ex:
        futs = [ex.submit(_compile, s, o) for s, o in zip(SOURCES, objs)]
        for f in futs:
            w = f.result()
            if verbose and w.strip():
                print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
           "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
