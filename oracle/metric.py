"""Oracle: grid metrics, probe readers, CFL time step and objectives.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def central4(f, axis):
    """famrdp_jax/mesh/metric.py:18-28 (4th-order central, roll wrap)."""
    fp2 = np.roll(f, -2, axis=axis)
    fp1 = np.roll(f, -1, axis=axis)
    fm1 = np.roll(f, 1, axis=axis)
    fm2 = np.roll(f, 2, axis=axis)
    return (-fp2 + 8.0 * fp1 - 8.0 * fm1 + fm2) / 12.0


def compute_metrics(xyz):
    """famrdp_jax/mesh/metric.py:31-70.

    Returns (jac, kxyz, vol) with J[m, l] = dx_m/dxi_l, jac = det J,
    kxyz = inv(J) laid out [l, m], vol = |jac|.  LU-based det/inv as in
    jnp.linalg (numpy.linalg is LAPACK getrf as well).
    """
    J = np.stack([np.stack([central4(xyz[m], l) for l in range(3)], axis=0) for m in range(3)], axis=0)
    Jc = np.transpose(J, (2, 3, 4, 0, 1))
    jac = np.linalg.det(Jc)
    kinv = np.linalg.inv(Jc)
    kxyz = np.ascontiguousarray(np.transpose(kinv, (3, 4, 0, 1, 2)))
    return jac, kxyz, np.abs(jac)


def read_probe(path, expect_ndim):
    """validation/fortran_ref.py:6-15: int32[ndim] shape + f64 Fortran order."""
    with open(path, "rb") as f:
        shape = tuple(int(x) for x in np.fromfile(f, dtype="<i4", count=expect_ndim))
        arr = np.fromfile(f, dtype="<f8", count=int(np.prod(shape))).reshape(shape, order="F")
    return arr


def read_xyz_probe(probe_dir, block_id):
    """validation/fortran_ref.py:18-31: int32[3] + x, y, z in Fortran order."""
    path = Path(probe_dir) / f"block_{block_id:04d}.bin"
    with open(path, "rb") as f:
        ni, nj, nk = (int(v) for v in np.fromfile(f, dtype="<i4", count=3))
        comps = [np.fromfile(f, dtype="<f8", count=ni * nj * nk).reshape((ni, nj, nk), order="F")
                 for _ in range(3)]
    return np.stack(comps, axis=0)


def compare(a, ref, *, atol=0.0, rtol=0.0):
    """validation/diff.py:5-19 (max_abs / max_rel summary)."""
    diff = np.abs(a - ref)
    rel = diff / np.maximum(np.abs(ref), 1e-300)
    return {"ok": bool(diff.max() <= atol and (rtol == 0.0 or rel.max() <= rtol)),
            "max_abs": float(diff.max()), "max_rel": float(rel.max())}


def cfl_dt(q, kxyz, vol, gas, cfl, ghost=2):
    """Build-defined global time step (the reference has no CFL->dt map).

    dt = cfl / max over interior cells of
         sum_d (|u . k_d| + c |k_d|) / vol  +  max(4/3, gamma/Pr) * mu/rho * sum_d |k_d|^2
    with k_d = kxyz[d, :], c = sqrt(gamma p / rho), mu from Sutherland.
    """
    g = gas["gamma"]
    sl = (slice(ghost, -ghost),) * 3
    qi = q[(slice(None),) + sl]
    k = kxyz[(slice(None), slice(None)) + sl]
    v = vol[sl]
    rho = qi[0]
    u = qi[1] / rho; vv = qi[2] / rho; w = qi[3] / rho
    p = (g - 1.0) * (qi[4] - rho * (0.5 * (u * u + vv * vv + w * w)))
    c = np.sqrt(g * p / rho)
    T = p / (rho * gas["R"])
    mu = gas["mu0"] * (T / gas["T0"]) ** 1.5 * (gas["T0"] + gas["Ts"]) / (T + gas["Ts"])
    inv = np.zeros_like(rho)
    ksq = np.zeros_like(rho)
    for d in range(3):
        kn2 = k[d, 0] * k[d, 0] + k[d, 1] * k[d, 1] + k[d, 2] * k[d, 2]
        inv = inv + (np.abs(u * k[d, 0] + vv * k[d, 1] + w * k[d, 2]) + c * np.sqrt(kn2))
        ksq = ksq + kn2
    lam = inv / v + max(4.0 / 3.0, g / gas["Pr_lam"]) * mu / rho * ksq
    return cfl / float(lam.max())
