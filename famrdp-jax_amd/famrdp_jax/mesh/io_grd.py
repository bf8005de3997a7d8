"""PLOT3D multi-block grid reader (famrdp_jax/mesh/io_grd.py:5-29).

File layout (little-endian, no Fortran record markers): int32 nblocks, then
int32 (ni, nj, nk) per block, then per block the x, y and z arrays of
ni*nj*nk float64 values each in Fortran (i fastest) order.
"""
from __future__ import annotations

import numpy as np


def read_grd(path) -> list:
    """Returns one (3, ni, nj, nk) float64 C-ordered array per block
    (axis 0: x, y, z)."""
    raw = np.fromfile(path, dtype=np.uint8)
    nb = int(raw[:4].view("<i4")[0])
    dims = raw[4:4 + 12 * nb].view("<i4").reshape(nb, 3).astype(np.int64)
    pos = 4 + 12 * nb
    out = []
    for ni, nj, nk in dims:
        n = int(ni * nj * nk)
        nbytes = 3 * n * 8
        if pos + nbytes > raw.size:
            raise ValueError(f"{path}: truncated grid file (block of {ni}x{nj}x{nk} needs {nbytes} bytes)")
        vals = raw[pos:pos + nbytes].view("<f8").reshape(3, int(nk), int(nj), int(ni))
        # Fortran order: the last numpy axis of this view is i; move it first
        out.append(np.ascontiguousarray(vals.transpose(0, 3, 2, 1), dtype=np.float64))
        pos += nbytes
    return out
