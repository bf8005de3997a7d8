"""MUSCL-2 reconstruction on the GPU (famrdp_jax/physics/reconstruct/muscl2.py)."""
from __future__ import annotations

from ... import _lib as L

_LIMITERS = ("minmod", "vanleer")


def _as4d(pv, axis):
    """View ``pv`` as (nvar, n0, n1, n2) with the reconstruction axis kept."""
    shape = tuple(pv.shape)
    if len(shape) < 2 or not (1 <= axis < len(shape)):
        raise ValueError(f"axis {axis} invalid for shape {shape}")
    nvar = shape[0]
    before = 1
    for s in shape[1:axis]:
        before *= s
    after = 1
    for s in shape[axis + 1:]:
        after *= s
    return nvar, before, shape[axis], after


def muscl2_reconstruct(pv, *, axis: int, limiter: str = "minmod"):
    """muscl2.py:18-37: returns (pv_L, pv_R) at faces j = 0..n-4 (between
    cells j+1 and j+2 along ``axis``)."""
    if limiter not in _LIMITERS:
        raise KeyError(limiter)
    torch = L._torch()
    d = L.device_f64(pv)
    nvar, n0, n1, n2 = _as4d(d, axis)
    out_shape = list(d.shape)
    out_shape[axis] -= 3
    pL = torch.empty(out_shape, dtype=torch.float64, device=d.device)
    pR = torch.empty_like(pL)
    L.check(L.lib().famrdp_muscl2(L.ptr(d), L.ptr(pL), L.ptr(pR), nvar, n0, n1, n2, 2,
                                  L.limiter_code(limiter), L.stream()))
    return L.like_input(pL, pv), L.like_input(pR, pv)


def _limiter(a, b, limiter):
    torch = L._torch()
    ta, tb = L.device_f64(a), L.device_f64(b)
    ta, tb = torch.broadcast_tensors(ta, tb)
    ta, tb = ta.contiguous(), tb.contiguous()
    out = torch.empty_like(ta)
    L.check(L.lib().famrdp_limiter(L.ptr(ta), L.ptr(tb), L.ptr(out), ta.numel(), L.limiter_code(limiter),
                                   L.stream()))
    return L.like_input(out, a if isinstance(a, torch.Tensor) else b)


def minmod(a, b):
    """muscl2.py:5-7 (evaluated on the GPU)."""
    return _limiter(a, b, "minmod")


def vanleer(a, b):
    """muscl2.py:10-12 (evaluated on the GPU)."""
    return _limiter(a, b, "vanleer")
