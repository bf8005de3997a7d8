"""Inviscid RHS on the GPU (famrdp_jax/rhs/invscd.py:18-133)."""
from __future__ import annotations

from .. import _lib as L


def compute_invscd_rhs(q, metrics, *, ghost: int, gamma: float, reconstruct: str, flux: str, limiter: str):
    """Same signature and result as the reference: MUSCL-2 + Roe per axis,
    cells 2..n-3 along each axis receive -(dF)/(vol + 1e-300).  As in the
    reference, ``ghost``, ``reconstruct`` and ``flux`` do not change the
    result."""
    torch = L._torch()
    code = L.limiter_code(limiter)
    qd = L.device_f64(q)
    kx = L.device_f64(metrics.kxyz)
    vo = L.device_f64(metrics.vol)
    _, ni, nj, nk = qd.shape
    rhs = torch.empty_like(qd)
    L.check(L.lib().famrdp_invscd_rhs(L.ptr(qd), L.ptr(kx), L.ptr(vo), L.ptr(rhs), ni, nj, nk, float(gamma), code,
                                      L.stream()))
    return L.like_input(rhs, q)
