"""One TVD-RK3 step on the GPU (famrdp_jax/solver/step.py:18-56)."""
from __future__ import annotations

from .. import _lib as L
from .._plan import plan_for
from ..core.types import Block, Config, State


def step(state: State, cfg: Config, dt: float) -> State:
    """Halo exchange once, BCs once, then three RHS evaluations combined by
    Shu-Osher TVD-RK3 over every padded cell of every block.  Returns a new
    State (t + dt, step + 1); the input arrays are not modified.  With
    ``cfg.nan_check`` (the reference's default) the new state is monitored on
    the GPU and a non-finite or non-positive (rho, p) interior cell raises
    FloatingPointError."""
    qs = [L.device_f64(b.q).clone() for b in state.blocks]
    kx = [L.device_f64(m.kxyz) for m in state.metrics]
    vo = [L.device_f64(m.vol) for m in state.metrics]
    dims = [tuple(q.shape[1:]) for q in qs]
    plan = plan_for(dims, cfg.topology, cfg.bc_params, cfg.gas, cfg.scheme.limiter, cfg.ghost)
    plan.step(qs, kx, vo, dt)
    if cfg.nan_check:
        plan.raise_if_bad(qs)
    blocks = tuple(Block(q=L.like_input(q, b.q), xyz=b.xyz) for q, b in zip(qs, state.blocks))
    return State(blocks=blocks, metrics=state.metrics, t=state.t + dt, step=state.step + 1)
