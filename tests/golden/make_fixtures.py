"""Write the small golden fixtures under tests/golden/ from the reference's data.

Reads (as data, never executes) the Fortran probe files that the reference
ships for its validation tests (validation/references/test3/{xyz,metric}/,
used by tests/validation/test_m10_grid_match.py and test_m11_metric_match.py
there), and stores a sub-box of each block:

* ``test3_metric_subbox.npz``: xyz[:, 0:24, 0:24, :] and jac[0:24, 0:24, :]
  for blocks 1 and 2 (51x51x11 each in the reference).  A 4th-order metric
  evaluated on the sub-box agrees with the full-block Jacobian wherever the
  stencil stays inside the sub-box in i and j (cells 2..21); k is full extent.

Run once in a container where /root/reference exists:
    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

REF = Path("/root/reference/validation/references/test3")
OUT = Path(__file__).resolve().parent


def _read_probe(path, ndim):
    with open(path, "rb") as f:
        shape = tuple(int(x) for x in np.fromfile(f, dtype="<i4", count=ndim))
        return np.fromfile(f, dtype="<f8", count=int(np.prod(shape))).reshape(shape, order="F")


def _read_xyz(path):
    with open(path, "rb") as f:
        ni, nj, nk = (int(v) for v in np.fromfile(f, dtype="<i4", count=3))
        return np.stack([np.fromfile(f, dtype="<f8", count=ni * nj * nk).reshape((ni, nj, nk), order="F")
                         for _ in range(3)])


def main():
    data = {}
    for b in (1, 2):
        xyz = _read_xyz(REF / "xyz" / f"block_{b:04d}.bin")
        jac = _read_probe(REF / "metric" / f"block_{b:04d}_jac.bin", 3)
        data[f"xyz{b}"] = np.ascontiguousarray(xyz[:, 0:24, 0:24, :])
        data[f"jac{b}"] = np.ascontiguousarray(jac[0:24, 0:24, :])
        data[f"full_shape{b}"] = np.asarray(jac.shape, dtype=np.int64)
    np.savez_compressed(OUT / "test3_metric_subbox.npz", **data)
    print("wrote", OUT / "test3_metric_subbox.npz")


if __name__ == "__main__":
    main()
