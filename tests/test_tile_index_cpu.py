"""CPU check of the fused stage kernels' cell indexing (famrdp_tiles.h): the
per-thread tables, evaluated at compile time from the LDS slot maps, and the
three wrap modes (modulo, one conditional wrap, none for tiles inside the
block).  For every tile of a block and every thread, the staged cells, the
viscous-tensor cells and the thread's own cell must be the periodic (roll)
image of the cell the tile needs (64-bit modular arithmetic as the
reference's jnp.roll, rhs/viscous.py:12-16) and lie in the block; the slot
tables must match the slot maps, which must invert the cell maps.

The header is compiled for the host into a test-only library
(tests/hostcheck/tile_index_hostcheck.hip); the product never loads it.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "famrdp-jax_amd", "csrc")


@pytest.fixture(scope="module")
def hc(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("tic") / "libtic.so")
    subprocess.run([hipcc, "--cuda-host-only", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + CSRC,
                    os.path.join(HERE, "hostcheck", "tile_index_hostcheck.hip"), "-o", out],
                   check=True, capture_output=True)
    return ctypes.CDLL(out)


# block sizes: single-wrap blocks whose last tiles end exactly 1, 2 or 3 cells
# before the block end (the no-wrap test's edge), blocks below the single-wrap
# bound (modulo), tiny blocks, the C2-like 100^3
@pytest.mark.parametrize("dims", [(20, 20, 20), (23, 37, 41), (25, 25, 25), (21, 29, 37), (40, 18, 70),
                                  (13, 30, 25), (5, 5, 5), (1, 2, 3), (36, 36, 36), (100, 100, 100)])
def test_fused_tile_indices(hc, dims):
    out = np.zeros(4, dtype=np.int64)
    hc.hc_tile_index_check(*dims, out.ctypes.data_as(ctypes.c_void_p))
    bad, checked, tiles, inside = (int(v) for v in out)
    assert checked > 0 and tiles > 0
    assert bad == 0, f"{bad} of {checked} indices differ from the periodic image or leave the block"
    if min(dims) >= 36:
        assert inside > 0          # the no-wrap path is exercised
