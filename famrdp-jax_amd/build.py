"""Build recipe for the HIP library (gfx950).

    python famrdp-jax_amd/build.py

Compiles two translation units with hipcc and links them into
famrdp_jax/_lib/libfamrdp_hip.so:

* csrc/famrdp_capi.hip (the C-ABI, which includes the forward, BC, halo and
  adjoint-support kernels) with the machine scheduler that favours
  instruction-level parallelism (max-ilp): the stage kernels are issue-bound
  FP64 code at two workgroups per CU, and this schedule measured 2 % faster on
  the fused kernel and 2.4 % faster on the marching kernel than the default;
* csrc/famrdp_reverse.hip (the two reverse stage kernels) with the default
  scheduler, which measured 2.5 % (faces) and 6.5 % (gather) faster on them.

Packed FP32 instructions are disabled in both (the kernels are FP64; this keeps
the compiler from forming them in any FP32 helper code).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "famrdp_jax", "_lib", "libfamrdp_hip.so")
CSRC = os.path.join(HERE, "csrc")
COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + CSRC]
UNITS = [("famrdp_capi.hip", ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]),
         ("famrdp_reverse.hip", [])]
SRC = os.path.join(CSRC, UNITS[0][0])


def _run(cmd, what: str, verbose: bool) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed for " + what)
    if verbose:
        sys.stderr.write(res.stderr)


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(os.path.dirname(OUT), os.path.splitext(src)[0] + ".o")
        _run([hipcc, *COMMON, *extra, "-c", "-o", obj, os.path.join(CSRC, src)], src, verbose)
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], "the link", verbose)
    for obj in objs:
        os.remove(obj)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
