"""BC dispatch on the GPU (famrdp_jax/physics/bc/dispatch.py:11-32)."""
from __future__ import annotations

from ... import _lib as L
from ..._plan import Plan, bcs_from_config
from ...core.types import Block


def apply_bc(block: Block, topo, cfg) -> Block:
    """Faces in Face order, sequential and in place on a copy: WALL
    (adiabatic / isothermal), SYMMETRY, FARFIELD; CUT1TO1 and the other kinds
    are no-ops (OUTFLOW with bc_params {"extrapolate": True} is the
    extrapolation extension)."""
    qd = L.device_f64(block.q).clone()
    bcs = bcs_from_config((topo,), cfg.bc_params)
    if bcs:
        plan = Plan([tuple(qd.shape[1:])], [], bcs, cfg.gas, "minmod", cfg.ghost, workspace=False)
        plan.apply_bc([qd])
    return Block(q=L.like_input(qd, block.q), xyz=block.xyz)
