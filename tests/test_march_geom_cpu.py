"""CPU check of the i-marching stage kernel's geometry (famrdp_march_geom.h).

For every workgroup of a block and every thread, each LDS slot the kernel
writes (primitive plane, viscous-tensor plane, face buffer) lies in its
buffer and is written once per step; each slot it reads holds what the read
needs: the periodic (roll) image of the neighbour for the viscous terms
(rhs/viscous.py:12-16), the unwrapped cell inside the block for the four
cells of a face (rhs/invscd.py:53-97), the face between the right cells for
the inviscid difference; and the updates cover every cell of the block
exactly once over all segments.  The workgroups are the ones the kernel
launches: TJ x TK tiles and, where the k remainder is narrow, transposed
TK x TJ strip tiles.  Block sizes: partial tiles at both tile widths, narrow
k remainders (strip tiles), blocks smaller than a tile, tiny blocks, segment
lengths that do and do not divide the block.

The header is compiled for the host into a test-only library
(tests/hostcheck/march_geom_hostcheck.hip); the product never loads it.
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
    out = str(tmp_path_factory.mktemp("mgc") / "libmgc.so")
    subprocess.run([hipcc, "--cuda-host-only", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + CSRC,
                    os.path.join(HERE, "hostcheck", "march_geom_hostcheck.hip"), "-o", out],
                   check=True, capture_output=True)
    return ctypes.CDLL(out)


# (dims, whether the 8 x 32 launches use strip tiles)
CASES = [((20, 20, 20), False), ((36, 36, 36), True), ((23, 37, 41), True), ((70, 18, 45), True),
         ((5, 5, 5), False), ((6, 7, 9), False), ((17, 17, 17), False), ((65, 33, 31), False),
         ((100, 52, 84), True), ((260, 20, 24), False), ((20, 70, 260), True), ((12, 41, 70), True),
         ((100, 100, 100), True)]


@pytest.mark.parametrize("dims,strips", CASES)
def test_march_geometry(hc, dims, strips):
    out = np.zeros(4, dtype=np.int64)
    hc.hc_march_geom_check(*dims, out.ctypes.data_as(ctypes.c_void_p))
    bad, checked, cover, nstrip = (int(v) for v in out)
    assert checked > 0
    assert bad == 0, f"{bad} of {checked} slot checks failed"
    assert cover == 0, f"{cover} cells not updated exactly once"
    assert (nstrip > 0) == strips, nstrip
