"""Equation of state on the GPU (famrdp_jax/physics/eos.py:5-25).

Arrays may be torch tensors (results stay on the device) or numpy arrays
(copied in and back).  The leading axis of ``prim``/``q`` holds the 5
variables; the remaining axes are flattened for the kernel.
"""
from __future__ import annotations

from .. import _lib as L


def _run5(fn, x, gamma):
    torch = L._torch()
    d = L.device_f64(x)
    if d.shape[0] != 5:
        raise ValueError(f"expected 5 variables on axis 0, got shape {tuple(d.shape)}")
    out = torch.empty_like(d)
    n = d[0].numel()
    L.check(fn(L.ptr(d), L.ptr(out), n, float(gamma), L.stream()))
    return L.like_input(out, x)


def prim_to_cons(prim, gamma):
    """eos.py:5-9."""
    return _run5(L.lib().famrdp_prim_to_cons, prim, gamma)


def cons_to_prim(q, gamma):
    """eos.py:12-17."""
    return _run5(L.lib().famrdp_cons_to_prim, q, gamma)


def speed_of_sound(rho, p, gamma):
    """eos.py:20-21: sqrt(gamma p / rho)."""
    torch = L._torch()
    r = L.device_f64(rho)
    pp = L.device_f64(p)
    r, pp = torch.broadcast_tensors(r, pp)
    r, pp = r.contiguous(), pp.contiguous()
    out = torch.empty_like(r)
    L.check(L.lib().famrdp_speed_of_sound(L.ptr(r), L.ptr(pp), L.ptr(out), r.numel(), float(gamma), L.stream()))
    res = L.like_input(out, rho if isinstance(rho, torch.Tensor) else p)
    return res.item() if hasattr(res, "ndim") and res.ndim == 0 and not isinstance(res, torch.Tensor) else res


def sutherland_viscosity(T, mu0, T0, Ts):
    """eos.py:24-25."""
    torch = L._torch()
    t = L.device_f64(T)
    out = torch.empty_like(t)
    L.check(L.lib().famrdp_sutherland(L.ptr(t), L.ptr(out), t.numel(), float(mu0), float(T0), float(Ts), L.stream()))
    return L.like_input(out, T)
