"""Builds libbdhip.so (the HIP step library) in-tree with hipcc for gfx950.

    python -m brownian_dynamics_cuda_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbdhip.so")
SOURCES = ["bd_forces.hip", "bd_integrate.hip", "bd_chol.hip", "bd_engine.hip", "bd_legacy.hip",
           "bd_rng.hip", "bd_potrf.hip", "bd_adjoint.hip"]

# -packed-fp32-ops: no v_pk_*_f32 (packed FP32) instructions in any kernel.
HIPFLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPFLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libbdhip.so")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=False):
    """The CPU oracle (and the reference library where the reference is
    present) -- test infrastructure, built beside the product."""
    odir = os.path.join(ROOT, "oracle")
    res = subprocess.run(["make", "-C", odir, "-s"], capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
    if res.returncode != 0:
        raise RuntimeError("building oracle/ failed")


if __name__ == "__main__":
    print(build_library(verbose=True))
    build_oracle(verbose=True)
