"""Build recipe for the HIP library (gfx950).

    python famrdp-jax_amd/build.py

Compiles csrc/famrdp_capi.hip (which includes the kernels) into
famrdp_jax/_lib/libfamrdp_hip.so with hipcc.  Packed FP32 instructions are
disabled (the kernels are FP64; this keeps the compiler from forming them in
any FP32 helper code).  The machine scheduler favours instruction-level
parallelism (max-ilp): the stage kernels are issue-bound FP64 code at two
workgroups per CU, and this schedule measured 2 % faster on them.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "famrdp_jax", "_lib", "libfamrdp_hip.so")
SRC = os.path.join(HERE, "csrc", "famrdp_capi.hip")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-sched-strategy=max-ilp",
         "-I" + os.path.join(HERE, "csrc")]


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc, *FLAGS, "-o", tmp, SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed for " + SRC)
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
