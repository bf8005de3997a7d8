"""Roe flux with Harten entropy fix on the GPU (famrdp_jax/flux/roe.py:21-85)."""
from __future__ import annotations

from .. import _lib as L


def roe_flux(UL, UR, nxyz, *, gamma, eps_entropy=0.1):
    """F(UL, UR; n) for unit normals n.  UL/UR: (5, ...); nxyz: (3, ...),
    broadcast against the states' trailing shape like the reference's jnp
    expressions."""
    torch = L._torch()
    a, b, n = L.device_f64(UL), L.device_f64(UR), L.device_f64(nxyz)
    if a.shape[0] != 5 or b.shape[0] != 5 or n.shape[0] != 3:
        raise ValueError("roe_flux: expected UL/UR with 5 and nxyz with 3 leading components")
    tail = torch.broadcast_shapes(tuple(a.shape[1:]), tuple(b.shape[1:]), tuple(n.shape[1:]))
    a = a.expand((5,) + tuple(tail)).reshape(5, -1).contiguous()
    b = b.expand((5,) + tuple(tail)).reshape(5, -1).contiguous()
    n = n.expand((3,) + tuple(tail)).reshape(3, -1).contiguous()
    F = torch.empty_like(a)
    L.check(L.lib().famrdp_roe_flux(L.ptr(a), L.ptr(b), L.ptr(n), L.ptr(F), a.shape[1], float(gamma),
                                    float(eps_entropy), L.stream()))
    return L.like_input(F.reshape((5,) + tuple(tail)), UL if isinstance(UL, torch.Tensor) else UR)
