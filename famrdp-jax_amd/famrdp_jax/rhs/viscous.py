"""Viscous RHS on the GPU (famrdp_jax/rhs/viscous.py:19-119)."""
from __future__ import annotations

import ctypes

from .. import _lib as L


def compute_viscous_rhs(q, metrics, *, gas, ghost: int):
    """Chain-rule central-difference viscous terms with the reference's
    periodic roll at the padded-array edges; mass row is zero."""
    torch = L._torch()
    qd = L.device_f64(q)
    kx = L.device_f64(metrics.kxyz)
    _, ni, nj, nk = qd.shape
    rhs = torch.empty_like(qd)
    work = torch.empty((9, ni, nj, nk), dtype=torch.float64, device=qd.device)
    g = L.gas_struct(gas)
    L.check(L.lib().famrdp_viscous_rhs(L.ptr(qd), L.ptr(kx), L.ptr(rhs), ni, nj, nk, ctypes.byref(g), L.ptr(work),
                                       L.stream()))
    return L.like_input(rhs, q)
