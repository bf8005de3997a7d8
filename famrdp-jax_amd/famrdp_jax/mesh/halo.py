"""1-to-1 halo exchange on the GPU (famrdp_jax/mesh/halo.py:56-92)."""
from __future__ import annotations

from .. import _lib as L
from .._plan import Plan, cuts_from_topology
from ..core.types import Block, GasModel


def halo_exchange(blocks, topos, ghost: int):
    """Whole-face ghost slabs copied from neighbours' interior-adjacent slabs,
    all reading the pre-exchange arrays, later faces (dict order) winning at
    edges and corners; orient != 0 raises NotImplementedError.  Returns new
    Blocks; the inputs are not modified."""
    qs = [L.device_f64(b.q).clone() for b in blocks]
    dims = [tuple(q.shape[1:]) for q in qs]
    plan = Plan(dims, cuts_from_topology(topos), [], GasModel(), "minmod", ghost, workspace=False)
    plan.halo_exchange(qs)
    return tuple(Block(q=L.like_input(q, b.q), xyz=b.xyz) for q, b in zip(qs, blocks))
