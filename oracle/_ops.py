"""Array-backend adapter for the oracle (test infrastructure only).

The oracle's formulas are written once and evaluated either with numpy
(the forward oracle) or with torch float64 on the CPU (the autograd oracle
that stands in for ``jax.grad``).  Only the handful of primitives the
reference uses are wrapped here, with JAX's derivative conventions where
torch's differ:

* ``abs``: JAX differentiates ``|x|`` as ``select(x >= 0, g, -g)``, so the
  derivative at 0 is +1; torch's own ``abs`` gives 0 there.
* ``maximum``/``minimum``: both JAX (``_balanced_eq``) and torch split the
  cotangent 1/2-1/2 on exact ties, so the library functions are used.
* ``sign``: derivative 0 in both.
* ``where``: cotangent flows only to the selected branch in both.
"""
from __future__ import annotations

import numpy as np

try:  # torch is only needed by the autograd oracle
    import torch
except Exception:  # pragma: no cover - torch is installed in this image
    torch = None


def is_torch(x) -> bool:
    return torch is not None and isinstance(x, torch.Tensor)


if torch is not None:
    class _JaxAbs(torch.autograd.Function):
        """|x| with derivative select(x >= 0, 1, -1) (JAX's abs JVP rule)."""

        @staticmethod
        def forward(ctx, x):
            ctx.save_for_backward(x)
            return x.abs()

        @staticmethod
        def backward(ctx, g):
            (x,) = ctx.saved_tensors
            return torch.where(x >= 0, g, -g)


def stack(xs, axis=0):
    if is_torch(xs[0]):
        return torch.stack(list(xs), dim=axis)
    return np.stack(xs, axis=axis)


def sqrt(x):
    return torch.sqrt(x) if is_torch(x) else np.sqrt(x)


def abs_(x):
    return _JaxAbs.apply(x) if is_torch(x) else np.abs(x)


def sign(x):
    return torch.sign(x) if is_torch(x) else np.sign(x)


def _as_like(v, ref):
    if is_torch(ref) and not is_torch(v):
        return torch.full_like(ref, float(v))
    return v


def maximum(a, b):
    if is_torch(a) or is_torch(b):
        ref = a if is_torch(a) else b
        return torch.maximum(_as_like(a, ref), _as_like(b, ref))
    return np.maximum(a, b)


def minimum(a, b):
    if is_torch(a) or is_torch(b):
        ref = a if is_torch(a) else b
        return torch.minimum(_as_like(a, ref), _as_like(b, ref))
    return np.minimum(a, b)


def where(c, a, b):
    if is_torch(c) or is_torch(a) or is_torch(b):
        ref = next(t for t in (a, b, c) if is_torch(t))
        if not is_torch(ref) or ref.dtype == torch.bool:
            ref = a if is_torch(a) else b
        return torch.where(c, _as_like(a, ref), _as_like(b, ref))
    return np.where(c, a, b)


def roll(x, shift, axis):
    if is_torch(x):
        return torch.roll(x, shifts=shift, dims=axis)
    return np.roll(x, shift, axis=axis)


def zeros_like(x):
    return torch.zeros_like(x) if is_torch(x) else np.zeros_like(x)


def ones_like(x):
    return torch.ones_like(x) if is_torch(x) else np.ones_like(x)


def copy(x):
    return x.clone() if is_torch(x) else np.array(x, copy=True)


def asarray_like(vals, ref):
    """A float64 array of ``vals`` on the same backend as ``ref``."""
    if is_torch(ref):
        return torch.as_tensor(np.asarray(vals, dtype=np.float64), dtype=torch.float64)
    return np.asarray(vals, dtype=np.float64)
