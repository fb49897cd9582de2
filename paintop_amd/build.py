"""Build libpaintop_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: ``python -m paintop_amd.build`` (or ``paintop_amd.build.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpaintop_hip.so")
SOURCES = [
    "sor_kernels.hip", "fp64_kernels.hip", "gen_kernels.hip", "solver.hip",
    "unet_kernels.hip", "unet.hip", "warm_kernels.hip", "assert_kernels.hip",
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    # no implicit a*b+c contraction: every fused multiply-add is written as fma()
    "-ffp-contract=off",
    # keep packed FP32 (v_pk_*_f32) out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the register-ring sweep kernels must unroll completely
    "-mllvm", "-pragma-unroll-threshold=1000000",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = [os.path.join(CSRC, src.replace(".hip", ".o")) for src in SOURCES]
    # the translation units are independent: compile them side by side
    jobs = max(1, min(len(SOURCES), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, so[0]),
                                     "-o", so[1]]), zip(SOURCES, objs)))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
