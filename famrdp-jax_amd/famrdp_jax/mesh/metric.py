"""Grid metrics on the GPU (famrdp_jax/mesh/metric.py:31-70)."""
from __future__ import annotations

from .. import _lib as L
from ..core.types import Metrics


def compute_metrics(xyz, *, ghost: int) -> Metrics:
    """4th-order central differences (roll wrap) of xyz, J[m, l] = dx_m/dxi_l,
    jac = det J, kxyz = inv(J) laid out [l, m], vol = |jac|."""
    torch = L._torch()
    x = L.device_f64(xyz)
    _, ni, nj, nk = x.shape
    jac = torch.empty((ni, nj, nk), dtype=torch.float64, device=x.device)
    vol = torch.empty_like(jac)
    kxyz = torch.empty((3, 3, ni, nj, nk), dtype=torch.float64, device=x.device)
    L.check(L.lib().famrdp_compute_metrics(L.ptr(x), L.ptr(jac), L.ptr(kxyz), L.ptr(vol), ni, nj, nk, L.stream()))
    return Metrics(jac=L.like_input(jac, xyz), kxyz=L.like_input(kxyz, xyz), vol=L.like_input(vol, xyz))
