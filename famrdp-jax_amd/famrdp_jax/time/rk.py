"""TVD-RK3 (famrdp_jax/time/rk.py:5-20) with the stage combinations on the GPU.

``q`` is an array or a pytree of arrays (nested tuples, lists and dicts),
as in the reference; ``L`` maps a state of that structure to one of the same
structure.  Every leaf is combined on the GPU with the reference's
expressions: q1 = q + dt L(q); q2 = 0.75 q + 0.25 (q1 + dt L(q1));
q_new = 1/3 q + 2/3 (q2 + dt L(q2)).
"""
from __future__ import annotations

from .. import _lib as _L


def _comb(out, a, x, b, y, c, z, mode):
    _L.check(_L.lib().famrdp_axpby(_L.ptr(out), float(a), _L.ptr(x), float(b), _L.ptr(y), float(c), _L.ptr(z),
                                   out.numel(), mode, _L.stream()))


def _flatten(tree):
    """Leaves in order and a function that rebuilds the structure from new leaves."""
    if isinstance(tree, (tuple, list)):
        parts = [_flatten(t) for t in tree]
        sizes = [len(p[0]) for p in parts]
        leaves = [x for p in parts for x in p[0]]

        def rebuild(new):
            out, i = [], 0
            for (_, rb), n in zip(parts, sizes):
                out.append(rb(new[i:i + n]))
                i += n
            if hasattr(tree, "_fields"):          # namedtuple
                return type(tree)(*out)
            return type(tree)(out)
        return leaves, rebuild
    if isinstance(tree, dict):
        keys = list(tree.keys())
        parts = [_flatten(tree[k]) for k in keys]
        sizes = [len(p[0]) for p in parts]
        leaves = [x for p in parts for x in p[0]]

        def rebuild(new):
            out, i = {}, 0
            for k, (_, rb), n in zip(keys, parts, sizes):
                out[k] = rb(new[i:i + n])
                i += n
            return type(tree)(out)
        return leaves, rebuild
    return [tree], lambda new: new[0]


def rk3_step(q, L, dt):
    """One Shu-Osher TVD-RK3 step of dq/dt = L(q) (time/rk.py:5-20)."""
    torch = _L._torch()
    leaves, rebuild = _flatten(q)
    q0 = [_L.device_f64(x) for x in leaves]

    def rhs(state_leaves):
        out, _ = _flatten(L(rebuild(state_leaves)))
        if len(out) != len(state_leaves):
            raise ValueError("rk3_step: L(q) must have the structure of q")
        r = [_L.device_f64(x).contiguous() for x in out]
        for a, b in zip(r, state_leaves):
            if tuple(a.shape) != tuple(b.shape):
                raise ValueError(f"rk3_step: L(q) leaf of shape {tuple(a.shape)} for a state leaf of shape "
                                 f"{tuple(b.shape)}")
        return r

    r = rhs(q0)
    q1 = [torch.empty_like(x) for x in q0]
    for o, x, z in zip(q1, q0, r):
        _comb(o, 0.0, None, 0.0, x, dt, z, 1)
    r = rhs(q1)
    q2 = [torch.empty_like(x) for x in q0]
    for o, x, y, z in zip(q2, q0, q1, r):
        _comb(o, 0.75, x, 0.25, y, dt, z, 2)
    r = rhs(q2)
    qn = [torch.empty_like(x) for x in q0]
    for o, x, y, z in zip(qn, q0, q2, r):
        _comb(o, 1.0 / 3.0, x, 2.0 / 3.0, y, dt, z, 2)
    return rebuild([_L.like_input(o, x) for o, x in zip(qn, leaves)])
