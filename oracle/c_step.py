"""Oracle: the TVD-RK3 step of one block whose faces are all CUT1TO1
connections to itself (periodic), in C + OpenMP (``oracle/c/step_omp.c``).

TEST INFRASTRUCTURE ONLY.  Used by ``bench.py``'s CPU-baseline leg (the
reference's CPU step timed on all host cores) and by
``tests/test_oracle_c.py``, which pins it against the numpy oracle
(``oracle.blocks.step``).  The product path never imports this module.

Same arithmetic, order and rounding as the numpy oracle (no FMA contraction);
every output value is computed by one thread, so results do not depend on
the thread count.  Citations: famrdp_jax/solver/step.py:18-56,
mesh/halo.py:56-92 (through the C source).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC_DIR = os.path.join(HERE, "c")
LIB = os.path.join(SRC_DIR, "libstep_omp.so")
CUT1TO1 = -1                      # famrdp_jax/core/constants.py:55-61
GAS_KEYS = ("gamma", "R", "mu0", "T0", "Ts", "Pr_lam")

_lib = None


def build() -> str:
    """Compile oracle/c/step_omp.c with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", SRC_DIR, "-B", "libstep_omp.so"], check=True)
    return LIB


def available() -> bool:
    return os.path.exists(LIB)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"C oracle not built: {LIB} is missing (run oracle.c_step.build())")
        lib = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.oracle_step_self_block.argtypes = [dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip,
                                               ctypes.c_int, dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        lib.oracle_step_self_block.restype = ctypes.c_int
        lib.oracle_max_threads.argtypes = []
        lib.oracle_max_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def max_threads() -> int:
    return int(load().oracle_max_threads())


def self_cuts(topo):
    """(target face, source face) of every CUT1TO1 face in the topology's dict
    order (halo.py:76-90); every face must be one, connected to the block
    itself with orient 0."""
    cuts = []
    for face, info in topo.neighbors.items():
        bct = topo.bc_type.get(face)
        if info is None or bct is None or int(bct) != CUT1TO1:
            raise NotImplementedError("the C oracle steps self-connected (periodic) blocks only")
        nb, src_face, orient = info
        if nb != topo.block_id or orient != 0:
            raise NotImplementedError("the C oracle steps self-connected (periodic) blocks only")
        cuts.append((int(face), int(src_face)))
    return cuts


def step(q, kxyz, vol, topo, gas, dt, *, limiter="minmod", threads=0):
    """One step of oracle.blocks.step for a single periodic block; returns the
    new q (the input is not modified)."""
    lib = load()
    out = np.array(q, dtype=np.float64, order="C", copy=True)
    k = np.ascontiguousarray(kxyz, dtype=np.float64)
    v = np.ascontiguousarray(vol, dtype=np.float64)
    _, ni, nj, nk = out.shape
    if k.shape != (3, 3, ni, nj, nk) or v.shape != (ni, nj, nk):
        raise ValueError("kxyz / vol shapes do not match q")
    cuts = self_cuts(topo)
    c = (ctypes.c_int * max(1, 2 * len(cuts)))(*[x for pair in cuts for x in pair])
    g6 = (ctypes.c_double * 6)(*[float(gas[key]) for key in GAS_KEYS])
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.oracle_step_self_block(out.ctypes.data_as(dp), k.ctypes.data_as(dp), v.ctypes.data_as(dp), ni, nj, nk,
                                    c, len(cuts), g6, float(dt), 1 if limiter == "vanleer" else 0, int(threads))
    if rc != 0:
        raise ValueError("oracle_step_self_block refused its arguments")
    return out
